"""The oracle against the reference's own outputs.

- Golden fixtures written by the reference's CPU engine (tests/golden/make_golden.py): the
  oracle's float64 and float32 paths must reproduce every per-batch sum, event count and
  profile bin exactly (same operation order, same libm).
- Values recorded in the reference's results/ files (validation-cpu.json "ours", and the event
  count of cpu-scaling.json), which pin the oracle at the reference's full validation sizes.
"""

import json

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_spec

FIXTURES = sorted((GOLDEN / "ref").glob("*.json"))


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_oracle_reproduces_reference_fixture(path):
    fx = json.loads(path.read_text())
    spec = load_spec(fx["case"])
    got = oracle.run(spec, photons=fx["photons"], batches=fx["batch_count"], seed=fx["seed"],
                     precision=fx["precision"], threads=0)
    assert got["specular"] == fx["specular"]
    assert got["events"] == fx["events"]
    assert got["anomalies"] == fx["anomalies"] == 0
    assert len(got["batches"]) == len(fx["batches"])
    for a, b in zip(got["batches"], fx["batches"]):
        assert a["first_photon"] == b["first_photon"]
        assert a["photons"] == b["photons"]
        assert a["rd"] == b["rd"]
        assert a["tt"] == b["tt"]
        assert a["absorbed"] == b["absorbed"]
        assert a["events"] == b["events"]
    assert got["a_z"] == fx["a_z"]
    assert got["rd_r"] == fx["rd_r"]


# (case, photons, batches, seed, recorded R_d or T_t mean) from results/validation-cpu.json
RECORDED = [
    ("v1_slab", "rd", 0.09740625141058865),
    ("v1_slab", "tt", 0.6609797492321345),
    ("v2_semi_infinite", "rd", 0.2198572838428182),
    ("o1_semi_infinite_matched", "rd", 0.4148798081284894),
    ("v3_three_layer", "rd", 0.2379575040124078),
    ("v3_three_layer", "tt", 0.09600239244240832),
]


@pytest.mark.slow
@pytest.mark.parametrize("case", sorted({c for c, _, _ in RECORDED}))
def test_oracle_matches_recorded_validation_runs(case):
    """4e6 photons, 50 batches, seed 1, float64: the reference's validation run of each case.

    The recorded numbers were produced on another platform (Apple clang/libm), so agreement is
    required to 1e-12 relative rather than bitwise.
    """
    run = oracle.run(load_spec(case), photons=4_000_000, batches=50, seed=1, threads=0)
    ph = [b["photons"] for b in run["batches"]]
    for c, q, value in RECORDED:
        if c != case:
            continue
        mean = sum(b[q] for b in run["batches"]) / sum(ph)
        assert mean == pytest.approx(value, rel=1e-12, abs=0)


@pytest.mark.slow
def test_oracle_matches_recorded_event_count_on_tissue_workload():
    """results/cpu-scaling.json: tissue_g09, 300,000 photons, 240 batches, seed 1 -> 464,569,365
    absorption events and R_d = 0.6163458655733034 (recorded on another platform)."""
    run = oracle.run(load_spec("tissue_g09"), photons=300_000, batches=240, seed=1, threads=0)
    assert run["events"] == 464_569_365
    rd = sum(b["rd"] for b in run["batches"]) / 300_000
    assert rd == pytest.approx(0.6163458655733034, rel=1e-12)


def test_thread_count_does_not_change_results():
    spec = load_spec("v3_three_layer")
    ref = oracle.run(spec, photons=30_000, batches=12, seed=9, threads=1)
    for t in (2, 3, 8):
        r = oracle.run(spec, photons=30_000, batches=12, seed=9, threads=t)
        assert r["batches"] == ref["batches"]
        assert r["a_z"] == ref["a_z"] and r["rd_r"] == ref["rd_r"] and r["t_r"] == ref["t_r"]


def test_extra_profiles_are_consistent_per_batch():
    """T_r and A(r,z) have no reference counterpart; they must account exactly (to summation
    rounding) for T_t and A_z in every batch, and A_z for the absorbed weight."""
    spec = load_spec("v3_three_layer")
    r = oracle.run(spec, photons=20_000, batches=5, seed=3, arz=True)
    for b, batch in enumerate(r["batches"]):
        assert np.isclose(r["tr_b"][b].sum(), batch["tt"], rtol=1e-12)
        assert np.isclose(r["rr_b"][b].sum(), batch["rd"], rtol=1e-12)
        assert np.isclose(r["az_b"][b].sum(), sum(batch["absorbed"]), rtol=1e-12)
        assert np.allclose(r["a_rz_b"][b].sum(axis=1), r["az_b"][b], rtol=1e-12, atol=1e-15)
