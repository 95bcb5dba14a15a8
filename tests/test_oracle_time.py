"""The oracle's time-of-flight profiles (no reference counterpart; the reference lists time-resolved
tallies as a non-goal). A photon's time is its optical path / c, the path summing n s over its
hops; escaping weight is binned in nt bins of dt ns (the last bin takes later photons).

Pinned by:
- exact identities: per batch, the time bins of R_d and T_t sum to R_d and T_t; switching the
  profiles on changes no other tally;
- a non-scattering slab (mua -> 0, mus = 0): every escape happens after exactly m traversals, at
  optical path m n d, transmitted for odd m and reflected for even m, with probability
  (1 - R) R^(m - 1) (R the normal-incidence Fresnel reflectance), so the bins of the other parity
  are exactly empty and the filled ones are binomial;
- a non-scattering two-layer stack with different indices: the first transmitted arrivals sit at
  n1 d1 + n2 d2 (each hop weighted by the index of its own layer), with probability
  (1 - R12)(1 - Rb), and nothing is transmitted earlier or before two more passes.
"""

import math

import numpy as np
import pytest

import oracle
from conftest import load_spec

C_CM_PER_NS = 29.9792458


def slab(n: float, d: float, layers=None) -> dict:
    return {"photons": 1, "dz": 0.01, "dr": 0.01, "nz": 10, "nr": 10, "na": 1,
            "n_above": 1.0, "n_below": 1.0,
            "layers": layers or [{"n": n, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d}]}


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_time_bins_sum_to_the_totals(precision):
    spec = load_spec("v3_three_layer")
    r = oracle.run(spec, photons=20_000, batches=4, seed=2, precision=precision, time=(40, 0.005))
    for b, batch in enumerate(r["batches"]):
        assert r["rt_b"][b].sum() == pytest.approx(batch["rd"], rel=1e-12)
        assert r["tt_b"][b].sum() == pytest.approx(batch["tt"], rel=1e-12)
    plain = oracle.run(spec, photons=20_000, batches=4, seed=2, precision=precision)
    assert plain["batches"] == r["batches"] and plain["a_z"] == r["a_z"]
    assert plain["rd_r"] == r["rd_r"] and plain["events"] == r["events"]


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_non_scattering_slab_arrives_at_whole_traversals(precision):
    n, d, N = 1.5, 0.1, 200_000
    R = ((n - 1.0) / (n + 1.0)) ** 2
    w0 = 1.0 - R  # launch weight: the specular reflection is taken off at entry
    # bins of 0.95 traversal: traversal m (< 19) lands in bin m, never on a bin edge
    dt = 0.95 * n * d / C_CM_PER_NS
    r = oracle.run(slab(n, d), photons=N, batches=4, seed=5, precision=precision, time=(16, dt))
    rt, tt = r["rd_t"], r["t_t"]
    assert rt[0] == 0 and tt[0] == 0
    for m in range(1, 15):
        if m % 2:  # odd: transmitted only
            assert rt[m] == 0, m
        else:
            assert tt[m] == 0, m
    for m in (1, 2, 3):
        p = (1 - R) * R ** (m - 1)
        got = (tt if m % 2 else rt)[m] / w0
        assert abs(got - N * p) <= 4 * math.sqrt(N * p * (1 - p)) + 1e-6, (m, got, N * p)


def test_two_layers_use_the_index_of_the_layer_crossed():
    n1, d1, n2, d2, N = 1.33, 0.1, 1.6, 0.05, 100_000
    layers = [{"n": n1, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d1},
              {"n": n2, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d2}]
    w0 = 1.0 - ((1.0 - n1) / (1.0 + n1)) ** 2
    r12 = ((n1 - n2) / (n1 + n2)) ** 2
    rb = ((n2 - 1.0) / (n2 + 1.0)) ** 2
    width = 0.017  # cm of optical path per bin
    # first transmitted arrivals: n1 d1 + n2 d2 = 0.213 (bin 12); with n1 for both layers they
    # would be at 0.1995 (bin 11), with n2 for both at 0.24 (bin 14); any later transmitted path
    # adds at least two passes, 2 n2 d2 = 0.16 (bin >= 21)
    first = n1 * d1 + n2 * d2
    r = oracle.run(slab(0, 0, layers), photons=N, batches=4, seed=9,
                   time=(40, width / C_CM_PER_NS))
    tt = r["t_t"] / w0
    k = int(first / width)
    assert k == 12 and int(n1 * (d1 + d2) / width) == 11 and int(n2 * (d1 + d2) / width) == 14
    assert np.all(tt[:k] == 0) and np.all(tt[k + 1:21] == 0)
    p = (1 - r12) * (1 - rb)
    assert abs(tt[k] - N * p) <= 4 * math.sqrt(N * p * (1 - p))


def test_float32_and_float64_time_profiles_agree():
    spec = load_spec("tissue_g09")
    kw = dict(photons=4000, batches=4, seed=3, time=(30, 0.02))
    a = oracle.run(spec, precision="float64", **kw)
    b = oracle.run(spec, precision="float32", **kw)
    for key in ("rd_t", "t_t"):
        ta, tb = a[key], b[key]
        assert abs(ta.sum() - tb.sum()) <= 5e-3 * max(ta.sum(), 1e-12)
        assert np.abs(ta - tb).max() <= 2e-2 * max(ta.sum(), 1e-12)


def test_time_options_are_checked():
    spec = load_spec("v1_slab")
    with pytest.raises(ValueError, match="nt >= 1"):
        oracle.run(spec, photons=100, batches=1, time=(0, 0.01))
    with pytest.raises(ValueError, match="dt > 0"):
        oracle.run(spec, photons=100, batches=1, time=(10, 0.0))
