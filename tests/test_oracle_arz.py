"""The oracle's A(r,z) and T_r binning, pinned by an independent photon walker.

A(r,z) and T_r have no reference counterpart, so no golden fixture holds them. The pure-Python
walker (tests/photon_walk.py) is first pinned photon for photon against the reference's own
float64 engine (the one-photon-per-batch fixture v3_three_layer_n40_b40_s5), then its recorded
absorption and exit positions, binned here by their own code, must reproduce the oracle's A(r,z),
T_r and R_r. A slip in the oracle's radial binning (bin width, the position used, the bin clamp)
fails these; so does the same slip planted in the walker's binning (last test).
"""

import json
import math

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_spec
from photon_walk import arz_of, bin_index, walk


@pytest.fixture(scope="module")
def v3_walk():
    fx = json.loads((GOLDEN / "ref" / "v3_three_layer_n40_b40_s5_float64.json").read_text())
    spec = load_spec("v3_three_layer")
    return fx, spec, [walk(spec, fx["seed"], i) for i in range(fx["photons"])]


def test_walker_reproduces_reference_engine_photon_for_photon(v3_walk):
    fx, spec, recs = v3_walk
    nl = len(spec["layers"])
    for b, rec in zip(fx["batches"], recs):
        assert b["photons"] == 1
        assert rec.rd == b["rd"] and rec.tt == b["tt"]
        absorbed = [math.fsum(a[4] for a in rec.absorptions if a[3] == k) for k in range(nl)]
        assert absorbed == pytest.approx(b["absorbed"], rel=1e-15, abs=0)
        assert len(rec.absorptions) == b["events"]
    # the reference's own depth and radius profiles, from the walker's positions
    az = np.zeros(spec["nz"])
    rr = np.zeros(spec["nr"])
    for rec in recs:
        for (x, y, z, _l, dwa) in rec.absorptions:
            az[bin_index(z, spec["dz"], spec["nz"])] += dwa
        if rec.rd:
            x, y = rec.exit_xy
            rr[bin_index(math.sqrt(x * x + y * y), spec["dr"], spec["nr"])] += rec.rd
    assert np.allclose(az, fx["a_z"], rtol=1e-13, atol=1e-17)
    assert np.allclose(rr, fx["rd_r"], rtol=1e-13, atol=1e-17)
    assert np.count_nonzero(rr) >= 10  # many radial bins are exercised


def _oracle_vs_walker(spec, seed, n, recs):
    o = oracle.run(spec, photons=n, batches=n, seed=seed, precision="float64", arz=True)
    nr, dr = spec["nr"], spec["dr"]
    for b, rec in enumerate(recs):
        ob = o["batches"][b]
        assert ob["rd"] == rec.rd and ob["tt"] == rec.tt and ob["events"] == len(rec.absorptions)
        want = arz_of([rec], spec)
        assert np.allclose(o["a_rz_b"][b], want, rtol=1e-13, atol=1e-18), b
        tr = np.zeros(nr)
        if rec.tt:
            x, y = rec.exit_xy
            tr[bin_index(math.sqrt(x * x + y * y), dr, nr)] += rec.tt
        assert np.allclose(o["tr_b"][b], tr, rtol=1e-13, atol=1e-18), b
    return o


def test_oracle_arz_and_tr_match_walker_on_reference_photons(v3_walk):
    fx, spec, recs = v3_walk
    o = _oracle_vs_walker(spec, fx["seed"], fx["photons"], recs)
    assert np.count_nonzero(o["a_rz"]) > 200  # many (z, r) cells are exercised


@pytest.mark.parametrize("case,n,seed", [("three_layer_c2", 60, 3), ("slab_c1", 300, 2)])
def test_oracle_arz_and_tr_match_walker_on_benchmark_grids(case, n, seed):
    """The benchmark grids: 200 x 200 bins of 0.005 cm (three layers) and 100 x 100 bins of
    0.001 cm, whose last radial bin collects photons beyond 0.1 cm (the clamp)."""
    spec = load_spec(case)
    recs = [walk(spec, seed, i) for i in range(n)]
    o = _oracle_vs_walker(spec, seed, n, recs)
    cells = np.count_nonzero(o["a_rz"])
    assert cells > 300, cells
    if case == "slab_c1":
        assert o["a_rz"][:, -1].sum() > 0  # the clamped last radial bin is exercised


def test_planted_binning_slips_are_caught(v3_walk):
    """The comparison above is sensitive: the walker's binning with a slip planted in it (the
    position before the hop, a 1% wider radial bin, r taken as |x| + |y|) disagrees with the
    oracle."""
    fx, spec, recs = v3_walk
    o = oracle.run(spec, photons=fx["photons"], batches=1, seed=fx["seed"], arz=True)
    good = arz_of(recs, spec)
    assert np.allclose(o["a_rz"], good, rtol=1e-13, atol=1e-18)
    assert not np.allclose(o["a_rz"], arz_of(recs, spec, positions="before"), rtol=1e-6)
    assert not np.allclose(o["a_rz"], arz_of(recs, dict(spec, dr=spec["dr"] * 1.01)), rtol=1e-6)
    manhattan = np.zeros_like(good)
    for rec in recs:
        for (x, y, z, _l, dwa) in rec.absorptions:
            manhattan[bin_index(z, spec["dz"], spec["nz"]),
                      bin_index(abs(x) + abs(y), spec["dr"], spec["nr"])] += dwa
    assert not np.allclose(o["a_rz"], manhattan, rtol=1e-6)
