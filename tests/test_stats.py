"""The parity judge, pinned.

Every verdict of the GPU parity tests rests on the batch-means estimator and the 3-sigma pass
rule (tissuemc_amd.stats) and the CLI's JSON summary (the C++ host layer, run_result.cpp). They
are checked here against:
- known values (the cases of the reference's tests/python/test_stats.py:7-57, stated anew);
- the `summary` block the reference's own engine wrote into every golden fixture: recomputed
  from the fixture's batches by the Python estimator and by the C++ writer, to 1e-12;
- the reference engine's JSON schema: every key it writes is in this engine's document, in the
  same order and of the same JSON type.
"""

import json
import math
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, CONFIGS
from tissuemc_amd import stats

FIXTURES = sorted((GOLDEN / "ref").glob("*.json"))


# ---------------------------------------------------------------- known values

def test_batch_estimate_known_values():
    # per-batch means 0.1, 0.2, 0.3: mean 0.2, sample sd 0.1, se 0.1 / sqrt(3)
    e = stats.batch_estimate([1.0, 2.0, 3.0], [10, 10, 10])
    assert e.mean == pytest.approx(0.2, rel=1e-15)
    assert e.se == pytest.approx(0.1 / math.sqrt(3), rel=1e-14)
    # unequal batches: the mean weights photons, the spread weights batches
    e = stats.batch_estimate([1.0, 4.0], [10, 20])
    assert e.mean == pytest.approx(5.0 / 30.0, rel=1e-15)
    assert e.se == pytest.approx(abs(0.2 - 0.1) / math.sqrt(2) / math.sqrt(2), rel=1e-14)


def test_single_batch_has_no_error_estimate():
    e = stats.batch_estimate([5.0], [100])
    assert e.mean == pytest.approx(0.05) and math.isnan(e.se)


def test_batch_estimate_rejects_bad_shapes():
    with pytest.raises(ValueError):
        stats.batch_estimate([1.0, 2.0], [10])
    with pytest.raises(ValueError):
        stats.batch_estimate([], [])


def test_pass_rule_is_inclusive_at_three_standard_errors():
    assert stats.compare(1.0 + 3.0e-3, 1e-3, 1.0).passed
    assert not stats.compare(1.0 + 3.01e-3, 1e-3, 1.0).passed
    assert not stats.compare(1.0 - 3.01e-3, 1e-3, 1.0).passed


def test_pass_rule_combines_both_standard_errors():
    c = stats.compare(1.0, 3e-4, 1.0 + 1e-3, 4e-4)  # combined 5e-4: z = -2
    assert c.z == pytest.approx(-2.0) and c.passed
    assert not stats.compare(1.0, 3e-4, 1.0 + 1.6e-3, 4e-4).passed  # z = -3.2


def test_rounding_allowance_widens_the_window():
    assert not stats.compare(0.2049, 0.0, 0.20).passed
    assert stats.compare(0.2049, 0.0, 0.20, rounding=0.005).passed
    assert not stats.compare(0.2051, 0.0, 0.20, rounding=0.005).passed


def test_zero_error_needs_an_exact_match():
    c = stats.compare(0.0, 0.0, 0.0)
    assert c.passed and c.z == 0.0
    assert not stats.compare(1.0, 0.0, 0.0).passed
    assert stats.compare(1.0, 0.0, 0.0).z == math.inf


def test_estimated_reference_se_scales_with_photons():
    assert stats.estimated_reference_se(1e-4, 4_000_000, 400_000) == pytest.approx(
        1e-4 * math.sqrt(10))


def test_profile_estimate_matches_the_scalar_estimator_bin_by_bin():
    rng = np.random.default_rng(3)
    pb = rng.random((7, 5))
    ph = [100, 100, 100, 99, 99, 99, 99]
    mean, se = stats.profile_estimate(pb, ph)
    for i in range(5):
        e = stats.batch_estimate(list(pb[:, i]), ph)
        assert mean[i] == pytest.approx(e.mean, rel=1e-14)
        assert se[i] == pytest.approx(e.se, rel=1e-12)
    z = stats.profile_z(np.array([1.0, 0.0, 2.0]), np.array([0.3, 0.0, 0.0]),
                        np.array([1.4, 0.0, 1.0]), np.array([0.4, 0.0, 0.0]))
    assert z[0] == pytest.approx(-0.8) and z[1] == 0.0 and z[2] == math.inf


# ---------------------------------------------------------------- the reference's summaries

def summary_of(fx: dict, estimate) -> dict:
    """rd, tt, absorbed and the conservation residual from a fixture's batches."""
    ph = [b["photons"] for b in fx["batches"]]
    absorbed = [sum(b["absorbed"]) for b in fx["batches"]]
    total = [b["rd"] + b["tt"] + a for b, a in zip(fx["batches"], absorbed)]
    out = {"rd": estimate([b["rd"] for b in fx["batches"]], ph),
           "tt": estimate([b["tt"] for b in fx["batches"]], ph),
           "absorbed": estimate(absorbed, ph)}
    m, se = estimate(total, ph)
    out["conservation_residual"] = (m + fx["specular"] - 1.0, se)
    return out


def assert_summary(got: dict, want: dict) -> None:
    for key, w in want.items():
        m, se = got[key]
        if key == "conservation_residual":
            # a difference of nearly equal numbers: compared absolutely
            assert m == pytest.approx(w["mean"], abs=1e-15), key
        else:
            assert m == pytest.approx(w["mean"], rel=1e-12, abs=0), key
        if w["se"] is None:
            assert math.isnan(se), key
        elif key == "conservation_residual":
            # per-batch totals equal up to rounding: their spread is rounding noise (~1e-17)
            assert se == pytest.approx(w["se"], abs=1e-15), key
        else:
            assert se == pytest.approx(w["se"], rel=1e-12, abs=1e-18), key


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_python_estimator_reproduces_reference_summary(path):
    fx = json.loads(path.read_text())

    def est(sums, ph):
        e = stats.batch_estimate(sums, ph)
        return e.mean, e.se

    assert_summary(summary_of(fx, est), fx["summary"])


@pytest.fixture(scope="module")
def summary_driver(tmp_path_factory):
    """tests/cpp/summary_driver.cpp linked with the C++ host layer (run_result.cpp)."""
    lib = ROOT / "tissuemc_amd" / "lib"
    if not (lib / "libtissuemc_hip.so").exists():
        pytest.fail("the engine library is not built")
    exe = tmp_path_factory.mktemp("driver") / "summary_driver"
    csrc = ROOT / "tissuemc_amd" / "csrc"
    subprocess.run(["g++", "-std=c++17", "-O2", f"-I{ROOT / 'include'}", f"-I{csrc}",
                    str(ROOT / "tests" / "cpp" / "summary_driver.cpp"), str(csrc / "run_result.cpp"),
                    f"-L{lib}", "-ltissuemc_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    return exe


def driver_doc(exe, fx: dict) -> dict:
    L = len(fx["batches"][0]["absorbed"])
    lines = [f"{fx['specular']!r} {len(fx['batches'])} {L}"]
    for b in fx["batches"]:
        vals = [b["first_photon"], b["photons"], float(b["rd"]), float(b["tt"]),
                *[float(a) for a in b["absorbed"]], b["events"], b["anomalies"]]
        lines.append(" ".join(repr(v) for v in vals))
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out)


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_cpp_summary_reproduces_reference_summary(summary_driver, path):
    fx = json.loads(path.read_text())
    doc = driver_doc(summary_driver, fx)
    got = {k: (v["mean"], math.nan if v["se"] is None else v["se"])
           for k, v in doc["summary"].items()}
    assert_summary(got, fx["summary"])
    assert doc["photons"] == fx["photons"] and doc["events"] == fx["events"]


def json_type(v) -> str:
    if isinstance(v, bool):
        return "bool"
    if v is None or isinstance(v, (int, float)):
        return "number"
    return type(v).__name__


def assert_schema_contains(ref, ours, where="run"):
    """Every key of `ref` is in `ours`, in the same relative order, with the same JSON type;
    lists are compared through their first element."""
    assert json_type(ours) == json_type(ref), where
    if isinstance(ref, dict):
        keys = [k for k in ours if k in ref]
        assert keys == list(ref), (where, [k for k in ref if k not in ours])
        for k in ref:
            assert_schema_contains(ref[k], ours[k], f"{where}.{k}")
    elif isinstance(ref, list) and ref and ours:
        assert_schema_contains(ref[0], ours[0], f"{where}[0]")


def test_json_schema_contains_the_reference_schema(summary_driver, tmp_path):
    import oracle

    if not oracle.ref_available():
        pytest.skip("the reference engine is not built here")
    ref = oracle.run_ref(CONFIGS / "v3_three_layer.mci", photons=400, batches=4, seed=2,
                         out_json=tmp_path / "ref.json")
    fx = {"specular": ref["specular"], "batches": ref["batches"]}
    ours = driver_doc(summary_driver, fx)
    assert_schema_contains(ref, ours)
    # same numbers where both engines define them from the same batches
    assert ours["summary"]["rd"]["mean"] == pytest.approx(ref["summary"]["rd"]["mean"], rel=1e-15)
    assert ours["summary"]["tt"]["se"] == pytest.approx(ref["summary"]["tt"]["se"], rel=1e-12)
    assert ours["batches"][1]["rd"] == ref["batches"][1]["rd"]
