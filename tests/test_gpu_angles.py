"""Exit-angle profiles on the GPU (MCML's Rd_a, Tt_a, Rd_ra, Tt_ra), through the C ABI.

- Identical photons against the float32 oracle: profile totals and cumulative shapes within the
  reference's cross-engine profile limits (the same random words, so the same exits).
- Exact integer identities with the pinned tallies: the angle bins sum to R_d and T_t, and per
  radius to R_r and T_r, in units of 2^-32.
- Chandrasekhar's closed form for a matched isotropic half-space, per angle bin, within 3.5
  standard errors estimated from ten independent runs.
"""

import math

import numpy as np
import pytest

import oracle
import tissuemc_amd.engine as engine
from conftest import load_spec
from tissuemc_amd import cross_engine, hfunction

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,na,photons", [("v3_three_layer", 12, 300_000),
                                             ("v2_semi_infinite", 30, 200_000),
                                             ("v1_slab", 20, 300_000)])
def test_identical_photons_with_oracle(gpu, name, na, photons):
    spec = dict(load_spec(name), na=na)
    g = engine.run(spec, photons=photons, batches=10, seed=7, angles=True)
    o = oracle.run(spec, photons=photons, batches=10, seed=7, precision="float32", angles=True)
    for key in ("rd_a", "t_a"):
        for gap in cross_engine.profile_gaps(o, g, key):
            assert gap.passed, (key, gap.line())
    for key in ("rd_ra", "t_ra"):
        a, b = np.asarray(o[key]).ravel(), np.asarray(g[key]).ravel()
        if a.sum() > 0:
            assert abs(a.sum() - b.sum()) / a.sum() < cross_engine.MAX_PROFILE_TOTAL_REL
            shape = np.abs(np.cumsum(a) - np.cumsum(b)).max() / a.sum()
            assert shape < cross_engine.MAX_PROFILE_SHAPE, (key, shape)


def test_exact_identities(gpu):
    spec = dict(load_spec("v3_three_layer"), na=9)
    g = engine.run(spec, photons=200_000, batches=8, seed=3, angles=True, arz=True)
    f = g["fixed"]
    assert int(f["ra"].sum()) == int(f["rd"].sum())
    assert int(f["ta"].sum()) == int(f["tt"].sum())
    assert np.array_equal(f["rra"].sum(axis=1), f["rr"].sum(axis=0))
    assert np.array_equal(f["tra"].sum(axis=1), f["tr"].sum(axis=0))
    assert np.array_equal(f["rra"].sum(axis=0), f["ra"])
    # the angle profiles change nothing else
    h = engine.run(spec, photons=200_000, batches=8, seed=3, arz=True)
    for key in ("rd", "tt", "absorbed", "az", "rr", "tr", "arz"):
        assert np.array_equal(f[key], h["fixed"][key]), key


def test_matched_half_space_follows_the_h_function(gpu):
    na, runs, photons = 18, 10, 400_000
    spec = dict(load_spec("o1_semi_infinite_matched"), na=na)
    per = np.array([engine.run(spec, photons=photons, batches=2, seed=100 + k, angles=True)["rd_a"]
                    for k in range(runs)]) / photons
    want = hfunction.escape_angle_fractions(10.0, 90.0, na)
    mean, se = per.mean(axis=0), per.std(axis=0, ddof=1) / math.sqrt(runs)
    z = (mean - want) / se
    assert np.all(np.abs(z) < 3.5), z


MCI = """1.0
1
mco_case.mco A
200000
0.01 0.02
40 25 12
3
1.0
1.37 1 100 0.90 0.1
1.37 1 10 0 0.1
1.37 2 10 0.70 0.2
1.0
"""


def test_cli_writes_mcml_output(gpu, tmp_path):
    import json
    import subprocess

    from conftest import ROOT
    from tissuemc_amd import mco

    (tmp_path / "case.mci").write_text(MCI)
    cli = ROOT / "tissuemc_amd" / "bin" / "tissuemc"
    p = subprocess.run([str(cli), "run", "case.mci", "--batches", "4", "--seed", "2", "--mco",
                        "--output", "case.json"], cwd=tmp_path, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    assert "wrote mco_case.mco" in p.stdout
    m = mco.read_mco((tmp_path / "mco_case.mco").read_text())
    j = json.loads((tmp_path / "case.json").read_text())["runs"][0]
    e = m["echo"]
    assert e["photons"] == 200000 and (e["nz"], e["nr"], e["na"]) == (40, 25, 12)
    assert (e["dz"], e["dr"]) == (0.01, 0.02)
    assert e["layers"][2] == [1.37, 2.0, 10.0, 0.7, 0.2]
    N = 200000
    rd = sum(b["rd"] for b in j["batches"]) / N
    tt = sum(b["tt"] for b in j["batches"]) / N
    ab = sum(sum(b["absorbed"]) for b in j["batches"]) / N
    assert m["rsp"] == pytest.approx(j["specular"], rel=1e-5)
    assert m["rd"] == pytest.approx(rd, rel=1e-5) and m["tt"] == pytest.approx(tt, rel=1e-5)
    assert m["absorbed"] == pytest.approx(ab, rel=1e-5)
    # MCML's normalizations, undone: every profile integrates back to its total (the values carry
    # five significant digits)
    dz, dr, nz, nr, na = 0.01, 0.02, 40, 25, 12
    da = math.pi / (2 * na)
    area = 2 * math.pi * (np.arange(nr) + 0.5) * dr * dr
    omega = 4 * math.pi * np.sin((np.arange(na) + 0.5) * da) * math.sin(da / 2)
    tol = dict(rel=2e-4)
    assert m["A_l"].sum() == pytest.approx(ab, **tol)
    assert (m["A_z"] * dz).sum() == pytest.approx(ab, **tol)
    assert (m["Rd_r"] * area).sum() == pytest.approx(rd, **tol)
    assert (m["Tt_r"] * area).sum() == pytest.approx(tt, **tol)
    assert (m["Rd_a"] * omega).sum() == pytest.approx(rd, **tol)
    assert (m["Tt_a"] * omega).sum() == pytest.approx(tt, **tol)
    assert np.allclose((m["A_rz"] * area[:, None] * dz).sum(axis=0), m["A_z"] * dz, rtol=2e-4,
                       atol=1e-9)
    assert np.allclose((m["Rd_ra"] * area[:, None] * omega[None, :]).sum(axis=1),
                       m["Rd_r"] * area, rtol=2e-4, atol=1e-9)
    assert np.allclose(m["Rd_a"] * omega * N, j["rd_a"], rtol=1e-4, atol=1e-3)
    assert np.allclose(m["Tt_ra"] * area[:, None] * omega[None, :] * N,
                       np.asarray(j["t_ra"]).reshape(nr, na), rtol=1e-4, atol=1e-3)
