"""A(r,z) of the HIP engine against the oracle, through the C ABI.

- Identical photons, the benchmark's own grids: the GPU's A(r,z) (plain and wave-aggregated
  tally paths) against the float32 oracle, within the reference's cross-engine limits. A(r,z) is
  compared flattened depth-major (total and cumulative shape through every cell) and as its
  radial marginal, beside A_z, R_r and T_r (tissuemc_amd.cross_engine).
- Every populated bin within 3 combined standard errors of the float64 oracle on the same
  photons: ten replicas (seeds) give each bin a mean and a standard error on both sides.
- Independent photons: the bins' z-scores against the float64 oracle follow the t distribution
  of ten-replica estimates (about 0.8% beyond 3; mean square about 1.1).

The oracle's own A(r,z) binning is pinned by tests/test_oracle_arz.py.
"""

import math

import numpy as np
import pytest

import oracle
from conftest import load_spec
from tissuemc_amd import cross_engine, stats

pytestmark = pytest.mark.gpu

CASES = [("three_layer_c2", 1_000_000, 10), ("v3_three_layer", 800_000, 20),
         ("slab_c1", 400_000, 10)]


@pytest.mark.parametrize("aggregate", [False, True], ids=["plain", "aggregated"])
@pytest.mark.parametrize("case,photons,batches", CASES, ids=[c[0] for c in CASES])
def test_arz_identical_photons_within_reference_limits(gpu, case, photons, batches, aggregate):
    spec = load_spec(case)
    g = gpu.run(spec, photons=photons, batches=batches, seed=11, arz=True, aggregate=aggregate)
    o = oracle.run(spec, photons=photons, batches=batches, seed=11, precision="float32",
                   arz=True, threads=16)
    gaps = cross_engine.compare_runs(o, g)
    names = {x.name.split(":")[0] for x in gaps}
    assert {"a_z", "rd_r", "t_r", "a_rz", "a_rz_r"} <= names
    failed = [x.line() for x in gaps if not x.passed]
    assert not failed, "\n".join(failed)
    assert g["anomalies"] == 0


K, N = 10, 100_000  # replicas (seeds) and photons per replica


def _replicas(run, seeds):
    return np.stack([run(s) for s in seeds]) / N  # per launched photon


def _estimate(reps):
    return reps.mean(axis=0), reps.std(axis=0, ddof=1) / math.sqrt(len(reps))


@pytest.fixture(scope="module")
def c2_replicas(gpu):
    spec = load_spec("three_layer_c2")
    same = range(100, 100 + K)

    def g(seed):
        return gpu.run(spec, photons=N, batches=16, seed=seed, arz=True)["a_rz"]

    def o(seed):
        return oracle.run(spec, photons=N, batches=16, seed=seed, precision="float64",
                          arz=True, threads=16)["a_rz"]

    return (_replicas(g, same), _replicas(o, same), _replicas(g, range(300, 300 + K)))


def test_arz_bins_within_3_sigma_of_float64_oracle_same_photons(c2_replicas):
    gpu_same, ora, _ = c2_replicas
    mg, sg = _estimate(gpu_same)
    mo, so = _estimate(ora)
    keep = (sg > 0) & (so > 0)
    assert keep.sum() > 20_000  # most of the 200 x 200 grid inside the stack is populated
    z = stats.profile_z(mg, sg, mo, so)[keep]
    assert np.all(np.abs(z) <= 3.0), float(np.max(np.abs(z)))
    # radial marginal
    mg, sg = _estimate(gpu_same.sum(axis=1))
    mo, so = _estimate(ora.sum(axis=1))
    keep = (sg > 0) & (so > 0)
    assert np.all(np.abs(stats.profile_z(mg, sg, mo, so)[keep]) <= 3.0)


def test_arz_bins_agree_statistically_with_independent_photons(c2_replicas):
    _, ora, gpu_other = c2_replicas
    mg, sg = _estimate(gpu_other)
    mo, so = _estimate(ora)
    keep = (sg > 0) & (so > 0)
    z = stats.profile_z(mg, sg, mo, so)[keep]
    # ten-replica standard errors: z follows a t distribution with ~18 degrees of freedom
    # (P(|t| > 3) = 0.0077, E[t^2] = 18/16)
    assert np.mean(np.abs(z) > 3) <= 0.02, float(np.mean(np.abs(z) > 3))
    assert np.mean(z ** 2) <= 1.25, float(np.mean(z ** 2))
    mg, sg = _estimate(gpu_other.sum(axis=1))
    mo, so = _estimate(ora.sum(axis=1))
    keep = (sg > 0) & (so > 0)
    zr = stats.profile_z(mg, sg, mo, so)[keep]
    assert np.mean(np.abs(zr) > 3) <= 0.05
    assert np.mean(zr ** 2) <= 1.6, float(np.mean(zr ** 2))
