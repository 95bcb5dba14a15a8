"""Time-of-flight profiles (TMC_FLAG_TIME) on the GPU, through the C ABI.

- Exact integer identities: the time bins of R_d and T_t sum to the batch sums of R_d and T_t;
  switching the profiles on changes no other tally.
- The non-scattering slab and two-layer stack of tests/test_oracle_time.py, whose arrival times
  and weights are fixed by the Fresnel series.
- Identical photons against the float32 oracle (same streams): the profiles' shapes agree within
  the cross-engine limits, the totals exactly as the other tallies do.
"""

import math

import numpy as np
import pytest

import oracle
import tissuemc_amd.engine as engine
from conftest import load_spec

pytestmark = pytest.mark.gpu

C_CM_PER_NS = 29.9792458


def slab(layers) -> dict:
    return {"photons": 1, "dz": 0.01, "dr": 0.01, "nz": 10, "nr": 10, "na": 1,
            "n_above": 1.0, "n_below": 1.0, "layers": layers}


@pytest.mark.parametrize("name,arz", [("v3_three_layer", True), ("tissue_g09", False)])
def test_identities_and_no_side_effects(gpu, name, arz):
    spec = load_spec(name)
    kw = dict(photons=300_000, batches=10, seed=4, arz=arz)
    t = engine.run(spec, time=(64, 0.004), **kw)
    plain = engine.run(spec, **kw)
    fx = t["fixed"]
    assert int(fx["rt"].sum()) == int(fx["rd"].sum())
    assert int(fx["tt_t"].sum()) == int(fx["tt"].sum())
    for key in ("rd", "tt", "absorbed", "az", "rr", "tr") + (("arz",) if arz else ()):
        assert np.array_equal(fx[key], plain["fixed"][key]), key
    assert t["events"] == plain["events"]


def test_many_bins_share_lds_with_the_wave_slots(gpu):
    """The angle and time bins live in each workgroup's LDS ahead of the wave slots: thousands of
    them take LDS from the A(r,z) tile and, past that, waves from the workgroup; the identities
    hold all the same. Bins that cannot fit beside even one wave are rejected."""
    spec = dict(load_spec("v3_three_layer"), na=500)
    kw = dict(photons=200_000, batches=4, seed=12, arz=True, angles=True)
    r = engine.run(spec, time=(7000, 0.0002), **kw)  # 2 (500 + 7000) counters = 120 KB
    fx = r["fixed"]
    assert int(fx["rt"].sum()) == int(fx["rd"].sum()) == int(fx["ra"].sum())
    assert int(fx["tt_t"].sum()) == int(fx["tt"].sum()) == int(fx["ta"].sum())
    plain = engine.run(spec, **kw)
    for key in ("rd", "tt", "absorbed", "az", "rr", "tr", "arz", "ra", "ta", "rra", "tra"):
        assert np.array_equal(fx[key], plain["fixed"][key]), key
    with pytest.raises(engine.ConfigError, match="do not fit in LDS"):
        engine.run(spec, photons=1000, batches=2, time=(20000, 0.01))


def test_non_scattering_slab(gpu):
    n, d, N = 1.5, 0.1, 400_000
    R = ((n - 1.0) / (n + 1.0)) ** 2
    w0 = 1.0 - R
    dt = 0.95 * n * d / C_CM_PER_NS
    spec = slab([{"n": n, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d}])
    r = engine.run(spec, photons=N, batches=4, seed=5, time=(16, dt))
    rt, tt = r["rd_t"], r["t_t"]
    assert rt[0] == 0 and tt[0] == 0
    for m in range(1, 15):
        assert (rt if m % 2 else tt)[m] == 0, m
    for m in (1, 2, 3):
        p = (1 - R) * R ** (m - 1)
        got = (tt if m % 2 else rt)[m] / w0
        assert abs(got - N * p) <= 4 * math.sqrt(N * p * (1 - p)) + 1e-3, (m, got, N * p)


def test_two_layers(gpu):
    n1, d1, n2, d2, N = 1.33, 0.1, 1.6, 0.05, 200_000
    spec = slab([{"n": n1, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d1},
                 {"n": n2, "mua": 1e-10, "mus": 0.0, "g": 0.0, "d": d2}])
    width = 0.017
    r = engine.run(spec, photons=N, batches=4, seed=9, time=(40, width / C_CM_PER_NS))
    w0 = 1.0 - ((1.0 - n1) / (1.0 + n1)) ** 2
    tt = r["t_t"] / w0
    assert np.all(tt[:12] == 0) and np.all(tt[13:21] == 0)
    p = (1 - ((n1 - n2) / (n1 + n2)) ** 2) * (1 - ((n2 - 1.0) / (n2 + 1.0)) ** 2)
    assert abs(tt[12] - N * p) <= 4 * math.sqrt(N * p * (1 - p)) + 1e-3


@pytest.mark.parametrize("name,grid", [("v3_three_layer", (60, 0.004)), ("tissue_g09", (50, 0.02))])
def test_identical_photons_with_oracle(gpu, name, grid):
    spec = load_spec(name)
    kw = dict(photons=200_000, batches=10, seed=7, time=grid)
    g = engine.run(spec, **kw)
    o = oracle.run(spec, precision="float32", **kw)
    for key, total in (("rd_t", "rd"), ("t_t", "tt")):
        gt, ot = np.asarray(g[key]), np.asarray(o[key])
        tot = sum(b[total] for b in o["batches"])
        if tot == 0:
            assert gt.sum() == 0
            continue
        assert abs(gt.sum() - ot.sum()) <= 1e-3 * tot
        # shapes: per-bin difference as a fraction of the total (cross-engine profile limit)
        assert np.abs(gt - ot).max() <= 2e-3 * tot, (key, np.abs(gt - ot).max() / tot)


def test_time_options_are_checked(gpu):
    spec = load_spec("v1_slab")
    with pytest.raises(engine.ConfigError, match="nt >= 1"):
        engine.run(spec, photons=1000, batches=2, time=(0, 0.01))
    with pytest.raises(engine.ConfigError, match="dt"):
        engine.run(spec, photons=1000, batches=2, time=(10, -1.0))


def test_cli_time(gpu, tmp_path):
    import json
    import subprocess

    from conftest import CONFIGS, ROOT

    cli = ROOT / "tissuemc_amd" / "bin" / "tissuemc"
    p = subprocess.run([str(cli), "run", str(CONFIGS / "v3_three_layer.mci"), "--photons", "100000",
                        "--batches", "4", "--time", "30", "0.005", "--output",
                        str(tmp_path / "t.json")], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    run = json.loads((tmp_path / "t.json").read_text())["runs"][0]
    assert len(run["rd_t"]) == 30 and run["time_dt_ns"] == 0.005
    assert sum(run["rd_t"]) == pytest.approx(sum(b["rd"] for b in run["batches"]), rel=1e-12)
    bad = subprocess.run([str(cli), "run", str(CONFIGS / "v3_three_layer.mci"), "--time", "30", "x"],
                         capture_output=True, text=True)
    assert bad.returncode == 2
