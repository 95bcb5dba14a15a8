"""The rocprofv3 profile tool (tissuemc_amd/profile.py) on the CPU: command construction (the
program directly after --, counters never combined with tracing, every pass within one run's
counter limits), parsing of rocprofv3's CSV files, derived figures and the side-by-side digest."""

import csv
from pathlib import Path

from tissuemc_amd import profile as P

TRACE_FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id",
                "Kernel_Id", "Kernel_Name", "Correlation_Id", "Start_Timestamp", "End_Timestamp"]
COUNTER_FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id",
                  "Thread_Id", "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size",
                  "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
                  "SGPR_Count", "Counter_Name", "Counter_Value"]
K = "void tmc::transport_kernel<true, true>(tmc::KernelArgs)"


def write_csv(path: Path, fields, rows):
    with path.open("w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow({k: r.get(k, "") for k in fields})


def test_commands_put_the_program_right_after_the_separator():
    prog = ["/x/tissuemc", "run", "a.mci"]
    t = P.rocprof_command(Path("/o"), "trace", prog, None)
    assert t[-4:] == ["--", *prog[:1], *prog[1:]][-4:] and t[t.index("--") + 1:] == prog
    assert "--kernel-trace" in t and "--pmc" not in t
    c = P.rocprof_command(Path("/o"), "p", prog, ["FETCH_SIZE"])
    assert "--pmc" in c and "--kernel-trace" not in c and "--sys-trace" not in c
    assert c[c.index("--") + 1:] == prog


def test_passes_stay_within_one_runs_counter_limits():
    for p in P.PASSES:
        sq = [c for c in p if c.startswith("SQ_")]
        assert len(sq) <= 8
        tcc = 3 * ("FETCH_SIZE" in p) + 2 * ("WRITE_SIZE" in p)
        assert tcc <= 4


def test_parses_trace_and_counters_of_the_matching_kernel(tmp_path):
    write_csv(tmp_path / "t.csv", TRACE_FIELDS, [
        {"Kernel_Name": K, "Start_Timestamp": 1000, "End_Timestamp": 251000},
        {"Kernel_Name": K, "Start_Timestamp": 300000, "End_Timestamp": 550000},
        {"Kernel_Name": "__amd_rocclr_copyBuffer", "Start_Timestamp": 0, "End_Timestamp": 5000},
    ])
    tr = P.read_trace(tmp_path / "t.csv", "transport_kernel")
    assert tr["launches"] == 2 and tr["duration_us"] == 250.0 and tr["kernels"] == [K]
    write_csv(tmp_path / "c.csv", COUNTER_FIELDS, [
        {"Kernel_Name": K, "Counter_Name": "SQ_WAVES", "Counter_Value": 4096},
        {"Kernel_Name": K, "Counter_Name": "SQ_INSTS_VALU", "Counter_Value": 4096 * 300},
        {"Kernel_Name": K, "Counter_Name": "FETCH_SIZE", "Counter_Value": 100},
        {"Kernel_Name": K, "Counter_Name": "WRITE_SIZE", "Counter_Value": 50},
        {"Kernel_Name": "other", "Counter_Name": "SQ_WAVES", "Counter_Value": 1},
    ])
    c = P.read_counters(tmp_path / "c.csv", "transport_kernel")
    assert c["SQ_WAVES"] == 4096
    d = P.derive(c, fetch_scale=2.0, write_scale=1.0)
    assert d["valu_per_wave"] == 300
    assert d["hbm_bytes"] == (100 * 2.0 + 50) * 1024


def test_digest_prints_ratios():
    a = {"duration_us": 200.0, "counters": {"SQ_WAVES": 4096.0}, "derived": {"valu_per_wave": 300.0}}
    b = {"duration_us": 100.0, "counters": {"SQ_WAVES": 4096.0}, "derived": {"valu_per_wave": 150.0}}
    text = "\n".join(P.digest(a, b))
    assert "0.500" in text and "SQ_WAVES" in text and "1.000" in text


def test_run_options_reach_the_engine_command():
    import argparse

    args = argparse.Namespace(mci=Path("tissuemc_amd/configs/three_layer_c2.mci"), photons=1000,
                              batches=10, seed=3, arz=True, angles=True, aggregate=True,
                              time=["200", "0.005"], no_profiles=False)
    prog = P.engine_program(args)
    assert prog[1] == "run" and prog[2].endswith("three_layer_c2.mci")
    assert prog[3:9] == ["--photons", "1000", "--batches", "10", "--seed", "3"]
    assert {"--arz", "--angles", "--aggregate"} <= set(prog)
    i = prog.index("--time")
    assert prog[i + 1:i + 3] == ["200", "0.005"]
    args.time, args.arz, args.angles, args.aggregate, args.no_profiles = None, False, False, False, True
    prog = P.engine_program(args)
    assert "--time" not in prog and "--arz" not in prog and prog[-1] == "--no-profiles"
