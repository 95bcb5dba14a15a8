"""Lane aggregation (TMC_FLAG_AGGREGATE) on the GPU, through the C ABI.

With the flag, a wave combines the A_z and A(r,z) adds of lanes that score the same bin (ballot +
DPP sum, one atomic per group) before the LDS atomics. Integer sums do not depend on grouping, so
every tally and counter must equal the plain path's bit for bit: on the three-layer workload, on a
semi-infinite medium, and on coarse grids where most lanes of a wave share a bin (large groups,
the A(r,z) tile's counters wrapping under group sums).
"""

import numpy as np
import pytest

import tissuemc_amd.engine as engine
from conftest import load_spec

pytestmark = pytest.mark.gpu

FIXED = ("rd", "tt", "absorbed", "az", "rr", "tr", "arz", "ra", "ta", "rra", "tra", "rt", "tt_t")


def same(a, b):
    for key in FIXED:
        if key in a["fixed"]:
            assert np.array_equal(a["fixed"][key], b["fixed"][key]), key
    for key in ("rd_a", "t_a", "rd_ra", "t_ra", "rd_t", "t_t"):
        if key in a:
            assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), key
    for key in ("events", "boundary_steps", "anomalies"):
        assert a[key] == b[key], key


@pytest.mark.parametrize("name,grid", [
    ("v3_three_layer", None),
    ("tissue_g09", None),
    ("v3_three_layer", dict(nz=2, nr=2)),       # nearly every lane of a wave in one bin
    ("tissue_g09", dict(nz=1, nr=1, dz=10.0, dr=10.0)),
])
def test_aggregated_tallies_equal_plain(gpu, name, grid):
    spec = dict(load_spec(name), **(grid or {}), na=6)
    kw = dict(photons=200_000, batches=8, seed=21, arz=True, angles=True)
    plain = engine.run(spec, **kw)
    agg = engine.run(spec, aggregate=True, **kw)
    same(plain, agg)
    assert agg["fixed"]["arz"].sum() > 0


@pytest.mark.parametrize("arz", [True, False])
def test_aggregated_with_time_profiles(gpu, arz):
    spec = dict(load_spec("v3_three_layer"), na=4)
    kw = dict(photons=200_000, batches=8, seed=23, arz=arz, angles=True, time=(40, 0.005))
    plain = engine.run(spec, **kw)
    agg = engine.run(spec, aggregate=True, **kw)
    same(plain, agg)
    assert plain["fixed"]["rt"].sum() > 0 and plain["fixed"]["tt_t"].sum() > 0


def test_aggregated_profiles_without_arz(gpu):
    spec = load_spec("v3_three_layer")
    kw = dict(photons=200_000, batches=8, seed=22)
    same(engine.run(spec, **kw), engine.run(spec, aggregate=True, **kw))


def test_batched_plan_aggregate(gpu):
    stacks = engine.generate_stacks(31, 0, 48, engine.default_stack_dist(photons_lo=2000,
                                                                          photons_hi=20000))
    grid = dict(dz=0.05, dr=0.05, nz=4, nr=4)
    outs = []
    for agg in (False, True):
        with engine.BatchedPlan(len(stacks), batches_per_stack=4, aggregate=agg, **grid) as plan:
            outs.append(plan.execute(stacks, seed=6))
    for key in ("rd", "tt", "absorbed", "events", "anomalies", "boundary_steps", "az", "rr", "tr"):
        assert np.array_equal(outs[0][key], outs[1][key]), key


def test_aggregate_needs_profiles(gpu):
    spec = load_spec("v1_slab")
    with pytest.raises(engine.ConfigError, match="aggregation"):
        engine.run(spec, photons=1000, batches=2, profiles=False, aggregate=True)
