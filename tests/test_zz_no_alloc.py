"""The no-allocation entry point (tmc_hip_run_static) against tmc_hip_run, through the C ABI.

Every device buffer of a static run comes from one static per-device array that all calls
share, so a call must clear what it uses and read nothing a previous call left: two static runs
in a row, the second on a smaller problem (fewer batches, a smaller grid, other tallies), each
give bit for bit the integer tallies of the allocating entry point on the same input.
"""

import numpy as np
import pytest

from conftest import load_spec

pytestmark = pytest.mark.gpu

KEYS = ("rd", "tt", "absorbed", "az", "rr", "tr", "arz", "ra", "ta", "rra", "tra", "rt", "tt_t")


def assert_same(a: dict, b: dict) -> None:
    for key in KEYS:
        x, y = a["fixed"].get(key), b["fixed"].get(key)
        assert (x is None) == (y is None), key
        if x is not None:
            assert np.array_equal(x, y), key
    for k in ("events", "anomalies", "boundary_steps"):
        assert [bb[k] for bb in a["batches"]] == [bb[k] for bb in b["batches"]], k
    assert a["specular"] == b["specular"]
    assert a["launch"] == b["launch"]


def test_static_runs_match_the_allocating_entry_point(gpu):
    big_spec = load_spec("three_layer_c2")
    big = dict(photons=400_000, batches=40, seed=7, arz=True)
    print("static run 1: three-layer stack, 200 x 200 grid with A(r,z), 40 batches", flush=True)
    s1 = gpu.run(big_spec, no_alloc=True, **big)
    a1 = gpu.run(big_spec, **big)
    assert_same(a1, s1)
    assert int(s1["fixed"]["arz"].sum()) > 0 and s1["events"] > 0

    small_spec = dict(load_spec("v3_three_layer"), na=8)
    small = dict(photons=60_000, batches=3, seed=9, angles=True, time=(16, 0.005))
    print("static run 2: 40 x 50 grid, angles and time bins, 3 batches", flush=True)
    s2 = gpu.run(small_spec, no_alloc=True, **small)
    a2 = gpu.run(small_spec, **small)
    assert_same(a2, s2)
    # nothing of the first run's larger layout leaks into the second's tallies
    f = s2["fixed"]
    for b in range(3):
        assert int(f["az"][b].sum()) == int(f["absorbed"][b].sum())
        assert int(f["rr"][b].sum()) == int(f["rd"][b])
        assert int(f["tr"][b].sum()) == int(f["tt"][b])
    assert int(f["ra"].sum()) == int(f["rd"].sum())


def test_static_run_refuses_a_run_beyond_its_scratch(gpu):
    spec = dict(load_spec("v3_three_layer"), nz=1500, nr=1500)
    with pytest.raises(gpu.ConfigError, match="static buffer holds"):
        gpu.run(spec, photons=1000, batches=2, arz=True, no_alloc=True)
