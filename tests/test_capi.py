"""The C-ABI library on the CPU: it loads, exports every declared symbol, parses .mci input like
the reference (tests/cpp/test_mci.cpp), and rejects bad runs before touching a device."""

import re
import subprocess

import pytest

import tissuemc_amd.engine as engine
from conftest import CONFIGS, ROOT

HEADER = ROOT / "include" / "tissuemc_hip.h"
CLI = ROOT / "tissuemc_amd" / "bin" / "tissuemc"

THREE_LAYER = """1.0 # file version
1 # number of runs
### Specify data
comp.mco A # output filename, ASCII/Binary
1000000 # No. of photons
.01 .01 # dz, dr
40 50 1 # No. of dz, dr & da.
3 # No. of layers
# n mua mus g d # One line for each layer
1.0 # n for medium above.
1.37 1 100 0.90 0.1 # layer 1
1.37 1 10 0 0.1 # layer 2
1.37 2 10 0.70 0.2 # layer 3
1.0 # n for medium below.
"""


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|uint64_t|double|void|const char \*)\s*(tmc_\w+)\s*\(",
                                 text, flags=re.M)))


def test_library_exports_every_declared_function():
    L = engine.load()
    names = declared_functions()
    assert len(names) >= 9
    for name in names:
        assert hasattr(L, name), name
    assert set(names) <= set(engine.EXPORTED) | {"tmc_hip_run"}
    assert L.tmc_fixed_point_unit() == 2.0 ** -32


def test_exported_symbols_are_plain_c():
    out = subprocess.run(["nm", "-D", "--defined-only", str(engine.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for name in declared_functions():
        assert name in exported, name


def mutate(old: str, new: str) -> str:
    assert old in THREE_LAYER
    return THREE_LAYER.replace(old, new, 1)


def parse_error(text: str) -> str:
    try:
        engine.parse_mci(text)
    except engine.ConfigError as e:
        return str(e)
    return ""


def test_parses_the_manual_three_layer_example():
    (r,) = engine.parse_mci(THREE_LAYER)
    assert r["output_name"] == "comp.mco" and r["photons"] == 1_000_000
    assert (r["dz"], r["dr"], r["nz"], r["nr"], r["na"]) == (0.01, 0.01, 40, 50, 1)
    assert r["layers"][0] == {"n": 1.37, "mua": 1.0, "mus": 100.0, "g": 0.9, "d": 0.1}
    assert r["layers"][2]["mua"] == 2.0 and r["layers"][2]["d"] == 0.2


def test_parses_several_runs_and_scientific_notation():
    text = "1.0\n2\n"
    text += "a.mco A\n100\n0.1 0.1\n1 1 1\n1\n1.0\n1.0 1 9 0 1e8\n1.0\n"
    text += "b.mco A\n200\n0.1 0.1\n1 1 1\n1\n1.5\n1.4 1 9 0.5 2E-1\n1.5\n"
    a, b = engine.parse_mci(text)
    assert a["layers"][0]["d"] == 1e8 and b["photons"] == 200 and b["n_above"] == 1.5
    assert b["layers"][0]["d"] == pytest.approx(0.2)


def test_errors_name_the_offending_line():
    assert "line 1" in parse_error(mutate("1.0 # file version", "2.0 # file version"))
    assert "line 4" in parse_error(mutate("comp.mco A", "comp.mco B"))
    assert "line 5" in parse_error(mutate("1000000 #", "1000000.5 #"))
    assert "line 12" in parse_error(mutate("1.37 1 10 0 0.1", "1.37 1 10 0"))
    assert "line 11" in parse_error(mutate("1.37 1 100 0.90 0.1", "1.37 1 100 abc 0.1"))


@pytest.mark.parametrize("old,new", [
    ("1.37 1 100 0.90 0.1", "1.37 1 100 1.0 0.1"),   # g = 1
    ("1.37 1 100 0.90 0.1", "1.37 -1 100 0.9 0.1"),  # mua < 0
    ("1.37 1 100 0.90 0.1", "1.37 1 100 0.9 0"),     # d = 0
    ("1.37 1 100 0.90 0.1", "0.9 1 100 0.9 0.1"),    # n < 1
    ("1.37 1 100 0.90 0.1", "1.37 0 0 0 0.1"),       # clear layer
    ("1000000 #", "0 #"),                            # no photons
    ("3 # No. of layers", "17 # No. of layers"),     # too many layers
])
def test_rejects_invalid_and_unsupported_input(old, new):
    assert parse_error(mutate(old, new))


def test_rejects_truncated_and_over_long_files():
    assert parse_error("")
    assert parse_error("1.0\n1\ncomp.mco A\n")
    assert parse_error(THREE_LAYER + "extra line\n")
    assert parse_error(mutate("1 # number of runs", "2 # number of runs"))


def test_missing_file_names_the_file():
    with pytest.raises(engine.ConfigError, match="definitely_missing.mci"):
        engine.parse_mci(path="/nonexistent/definitely_missing.mci")


def test_bundled_inputs_parse():
    for p in sorted(CONFIGS.glob("*.mci")):
        assert len(engine.parse_mci(path=p)) == 1


def test_bad_runs_are_rejected_before_any_device_work():
    (spec,) = engine.parse_mci(THREE_LAYER)
    with pytest.raises(engine.ConfigError, match="batches must be >= 1"):
        engine.run(spec, photons=10, batches=0)
    with pytest.raises(engine.ConfigError, match="cannot exceed"):
        engine.run(spec, photons=10, batches=11)
    with pytest.raises(engine.ConfigError, match="layer 1: g must be"):
        engine.run(dict(spec, layers=[dict(spec["layers"][0], g=-1.0)]), photons=10, batches=1)
    with pytest.raises(engine.ConfigError, match="ambient"):
        engine.validate_spec(dict(spec, n_above=0.5))


def test_no_device_is_an_error_not_a_fallback():
    if engine.available():
        pytest.skip("a GPU is present")
    (spec,) = engine.parse_mci(THREE_LAYER)
    with pytest.raises(engine.DeviceError, match="no usable HIP device"):
        engine.run(spec, photons=100, batches=1)


def cli(*args):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True)


def test_cli_surface():
    assert cli("--version").returncode == 0
    assert cli().returncode == 2
    assert cli("run").returncode == 2
    r = cli("run", str(CONFIGS / "v1_slab.mci"), "--bogus")
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = cli("run", str(CONFIGS / "v1_slab.mci"), "--batches", "-3")
    assert r.returncode == 2
    r = cli("run", str(CONFIGS / "v1_slab.mci"), "--engine", "cpu")
    assert r.returncode == 2 and "oracle" in r.stderr
    for bad in ("0", "1.5", "x"):
        r = cli("run", str(CONFIGS / "v1_slab.mci"), "--autotune", bad)
        assert r.returncode == 2 and "fraction" in r.stderr
    r = cli("run", "/nonexistent/x.mci")
    assert r.returncode == 1 and "tissuemc: error:" in r.stderr and "x.mci" in r.stderr


def test_autotune_rejects_a_bad_pilot_fraction_before_any_device_work():
    (spec,) = engine.parse_mci(THREE_LAYER)
    for f in (0.0, -1.0, 1.5, float("nan")):
        with pytest.raises(engine.ConfigError, match="pilot_fraction"):
            engine.autotune(spec, photons=1000, batches=10, pilot_fraction=f)


def test_angle_profiles_are_checked_before_any_device_work():
    (spec,) = engine.parse_mci(THREE_LAYER)
    with pytest.raises(engine.ConfigError, match="no-profiles"):
        engine.run(spec, photons=100, batches=1, angles=True, profiles=False)
    L = engine.load()
    o = engine.TmcBatchedOpts(seed=1, batches_per_stack=1, dz=0.01, dr=0.01, nz=10, nr=10,
                              flags=engine.FLAG_ANGLES)
    plan = engine.C.c_void_p()
    assert L.tmc_batched_create(4, engine.C.byref(o), engine.C.byref(plan)) == engine.TMC_ERR_CONFIG
    assert "single-run options" in L.tmc_last_error().decode()


def test_time_and_aggregation_options_are_checked_before_any_device_work():
    (spec,) = engine.parse_mci(THREE_LAYER)
    bad = [(dict(time=(0, 0.01)), "nt >= 1"), (dict(time=(10, 0.0)), "dt"),
           (dict(time=(10, float("nan"))), "dt"),
           (dict(time=(10, 0.01), profiles=False), "no-profiles"),
           (dict(aggregate=True, profiles=False), "no-profiles")]
    for kw, msg in bad:
        with pytest.raises(engine.ConfigError, match=msg):
            engine.run(spec, photons=100, batches=1, **kw)
    L = engine.load()
    o = engine.TmcBatchedOpts(seed=1, batches_per_stack=1, dz=0.01, dr=0.01, nz=10, nr=10,
                              flags=engine.FLAG_TIME)
    plan = engine.C.c_void_p()
    assert L.tmc_batched_create(4, engine.C.byref(o), engine.C.byref(plan)) == engine.TMC_ERR_CONFIG
    assert "single-run options" in L.tmc_last_error().decode()


# Inputs the reader must reject, each with the message the reference's own parser gives.
BAD_INPUTS = [
    "",
    "# only a comment\n\n",
    "2.0\n1\n",
    "1.0\n0\n",
    "1.0\n1.5\n",
    "1.0\n1\ncomp.mco\n",
    "1.0\n1\ncomp.mco B\n",
    "1.0\n1\ncomp.mco A\n",
    "1.0\n1\ncomp.mco A\n100.0\n",
    "1.0\n1\ncomp.mco A\n0\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 x\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 inf\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 1e999\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 0.1\n10 10\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 0.1\n10 0 1\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 0.1\n10 10 2000000\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 0.1\n10 10 1\n17\n",
    "1.0\n1\ncomp.mco A\n100\n0.1 0.1\n10 10 1\n1\n1.0\n1.4 1 10 0.9\n",
]


@pytest.mark.parametrize("mutation", [
    ("1.0 # file version", "2.0 # file version"), ("comp.mco A", "comp.mco B"),
    ("1000000 #", "1000000.5 #"), ("1.37 1 10 0 0.1", "1.37 1 10 0"),
    ("1.37 1 100 0.90 0.1", "1.37 1 100 abc 0.1"), ("1.37 1 100 0.90 0.1", "1.37 1 100 1.0 0.1"),
    ("1.37 1 100 0.90 0.1", "1.37 -1 100 0.9 0.1"), ("1.37 1 100 0.90 0.1", "1.37 1 100 0.9 0"),
    ("1.37 1 100 0.90 0.1", "0.9 1 100 0.9 0.1"), ("1.37 1 100 0.90 0.1", "1.37 0 0 0 0.1"),
    ("1.0 # n for medium above.", "0.5 # n for medium above."),
    ("3 # No. of layers", "17 # No. of layers"), ("1 # number of runs", "2 # number of runs"),
    ("1.0 # n for medium below.\n", "1.0 # n for medium below.\nextra line\n"),
] + [(None, t) for t in BAD_INPUTS])
def test_error_messages_match_the_reference_parser(tmp_path, mutation):
    """Same input, same message: this reader's error text against the reference engine's own
    parser (its CLI prints "tissuemc: error: <message>"), built from its sources where they are
    present."""
    import oracle

    if not oracle.ref_available():
        pytest.skip("the reference engine is not built here")
    old, new = mutation
    text = new if old is None else mutate(old, new)
    path = tmp_path / "input.mci"
    path.write_text(text)
    ref = subprocess.run([str(oracle.REF_BIN), "run", str(path), "--photons", "1"],
                         capture_output=True, text=True)
    assert ref.returncode == 1, ref.stderr
    want = ref.stderr.strip().removeprefix("tissuemc: error: ")
    with pytest.raises(engine.ConfigError) as err:
        engine.parse_mci(path=path)
    assert str(err.value) == want
    ours = cli("run", str(path))
    assert ours.returncode == 1 and ours.stderr.strip() == ref.stderr.strip()


@pytest.mark.parametrize("text", [
    THREE_LAYER,
    THREE_LAYER.replace("\n", "\r\n"),
    THREE_LAYER.replace(" ", "\t").rstrip("\n"),
    "1.0\n1\nx.mco A\n+50\n1E-2 .5e-1\n7 9 3\n1\n+1.0\n1.33 0x1p-2 1e2 -0.5 3.5 # hex mua\n1.2",
])
def test_parsed_values_match_the_reference_parser(tmp_path, text):
    """Valid inputs in unusual spellings (CRLF, tabs, signs, exponents, hex, no final newline)
    read to the same run as the reference engine reads, per the input echo of its JSON."""
    import json

    import oracle

    if not oracle.ref_available():
        pytest.skip("the reference engine is not built here")
    path = tmp_path / "input.mci"
    path.write_bytes(text.encode())
    out = tmp_path / "ref.json"
    subprocess.run([str(oracle.REF_BIN), "run", str(path), "--photons", "2", "--batches", "1",
                    "--output", str(out)], check=True, capture_output=True)
    ref = json.loads(out.read_text())["runs"][0]
    (ours,) = engine.parse_mci(path=path)
    assert ours["output_name"] == ref["output_name"]
    assert ours["layers"] == ref["layers"]
    assert (ours["dz"], ours["dr"], ours["nz"], ours["nr"]) == tuple(
        ref["grid"][k] for k in ("dz", "dr", "nz", "nr"))
    assert (ours["n_above"], ours["n_below"]) == (ref["n_above"], ref["n_below"])


def test_static_run_rejects_a_run_beyond_its_scratch_before_any_device_work():
    """tmc_hip_run_static takes every device buffer from a fixed 16 MiB array: a run whose
    tallies need more (here A(r,z) on 1500 x 1500 bins, 18 MB) is refused with a clear error
    before any device call, on any machine."""
    L = engine.load()
    assert L.tmc_static_scratch_bytes() == 16 << 20
    (spec,) = engine.parse_mci(THREE_LAYER)
    big = dict(spec, nz=1500, nr=1500)
    with pytest.raises(engine.ConfigError, match="bytes of device scratch; the static buffer holds 16777216"):
        engine.run(big, photons=100, batches=2, arz=True, no_alloc=True)
    with pytest.raises(engine.ConfigError, match="batches must be >= 1"):
        engine.run(spec, photons=100, batches=0, no_alloc=True)
