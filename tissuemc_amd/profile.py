"""Hardware-counter profile of the transport kernel with rocprofv3, and a side-by-side digest.

    python -m tissuemc_amd.profile run --output results/c2.json [--mci FILE] [--photons N]
        [--batches B] [--arz] [--no-profiles] [--angles] [--aggregate] [--time NT DT]
    python -m tissuemc_amd.profile digest results/c2.json results/c2-no-profiles.json

The counterpart of the reference's Nsight Compute tool (python/tissuemc_analysis/ncu_profile.py:
a profile of one configuration, and ``--digest`` printing two profiles side by side with their
ratio) for the ablation loop on the MI355X. ``run`` starts the engine's command-line host under
rocprofv3 once for a kernel trace and once per counter pass (rocprofv3 does not split counters
over passes, so each pass stays within one run's hardware limits), then keeps, for the kernels
whose names match ``--kernel``, the launch durations, every counter, and the derived figures that
say what limits the kernel: VALU instructions per wave, the fraction of lanes active per VALU
instruction, VALU/SALU/MFMA busy, occupancy, LDS utilisation and bank conflicts, LDS atomics, and
HBM bytes (FETCH_SIZE and WRITE_SIZE in KiB, scaled by ``--fetch-scale``/``--write-scale``, the
factors a plain device copy of known size reads for them).

Run it on a machine with the GPU, from any directory; it sets TMPDIR=/tmp for rocprofv3 and runs
the host program directly after ``--``.
"""

from __future__ import annotations

import argparse
import csv
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent
CLI = ROOT / "bin" / "tissuemc"

# Counter passes, each within one run's limits (at most 8 SQ, 4 TCC of which FETCH_SIZE takes 3
# and WRITE_SIZE 2, 4 TCP, 2 TA, 2 TD, 2 GRBM counters).
PASSES: list[list[str]] = [
    ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_MFMA",
     "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU"],
    ["VALUBusy", "VALUUtilization", "SALUBusy", "OccupancyPercent"],
    ["LDSBankConflict", "LdsUtil", "SQ_INSTS_LDS_ATOMIC", "SQ_LDS_ATOMIC_RETURN",
     "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_INT32"],
    ["MfmaUtil"],
    ["FETCH_SIZE"],
    ["WRITE_SIZE"],
]

# (name, what it says) of the figures the digest prints first
KEY = [
    ("duration_us", "mean launch duration"),
    ("valu_per_wave", "VALU instructions per wave"),
    ("VALUUtilization", "% of lanes active per VALU instruction (divergence)"),
    ("VALUBusy", "% VALU busy"),
    ("SALUBusy", "% SALU busy"),
    ("MfmaUtil", "% MFMA busy"),
    ("OccupancyPercent", "% of wave slots occupied"),
    ("LdsUtil", "% LDS busy"),
    ("LDSBankConflict", "% of LDS cycles lost to bank conflicts"),
    ("lds_atomics_per_wave", "LDS atomic instructions per wave"),
    ("hbm_bytes", "HBM bytes (calibrated FETCH_SIZE + WRITE_SIZE)"),
]


def rocprof_command(out_dir: Path, name: str, program: list[str], pmc: list[str] | None) -> list[str]:
    cmd = [shutil.which("rocprofv3") or "rocprofv3"]
    if pmc:
        cmd += ["--pmc", *pmc]
    else:
        cmd += ["--kernel-trace", "--stats"]
    return cmd + ["--output-format", "csv", "-d", str(out_dir / name), "-o", "run", "--", *program]


def read_counters(path: Path, pattern: str) -> dict[str, float]:
    """Counters of the matching kernels in a rocprofv3 counter_collection.csv, averaged over
    dispatches (counts per dispatch)."""
    sums: dict[str, list[float]] = {}
    with path.open() as f:
        for row in csv.DictReader(f):
            if not re.search(pattern, row["Kernel_Name"]):
                continue
            sums.setdefault(row["Counter_Name"], []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in sums.items()}


def read_trace(path: Path, pattern: str) -> dict:
    durs, names = [], set()
    with path.open() as f:
        for row in csv.DictReader(f):
            if re.search(pattern, row["Kernel_Name"]):
                durs.append((int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3)
                names.add(row["Kernel_Name"])
    if not durs:
        return {}
    return {"kernels": sorted(names), "launches": len(durs), "duration_us": sum(durs) / len(durs),
            "duration_min_us": min(durs), "duration_max_us": max(durs)}


def derive(c: dict[str, float], fetch_scale: float, write_scale: float) -> dict[str, float]:
    d: dict[str, float] = {}
    waves = c.get("SQ_WAVES")
    if waves:
        for key, name in (("SQ_INSTS_VALU", "valu_per_wave"), ("SQ_INSTS_SALU", "salu_per_wave"),
                          ("SQ_INSTS_LDS", "lds_per_wave"), ("SQ_INSTS_MFMA", "mfma_per_wave"),
                          ("SQ_INSTS_LDS_ATOMIC", "lds_atomics_per_wave")):
            if key in c:
                d[name] = c[key] / waves
    if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
        d["hbm_bytes"] = (c["FETCH_SIZE"] * fetch_scale + c["WRITE_SIZE"] * write_scale) * 1024.0
    return d


def profile(program: list[str], kernel: str, work: Path, fetch_scale: float = 2.0,
            write_scale: float = 1.0, timeout: int = 180) -> dict:
    """Trace and counter passes of `program`; figures for the kernels matching `kernel`."""
    env = {**os.environ, "TMPDIR": "/tmp"}
    runs = [("trace", None)] + [(f"pass{i}", p) for i, p in enumerate(PASSES)]
    result: dict = {"program": program, "kernel": kernel, "counters": {}}
    for name, pmc in runs:
        cmd = rocprof_command(work, name, program, pmc)
        print(f"profile: {name}", file=sys.stderr, flush=True)
        proc = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0:
            raise SystemExit(f"rocprofv3 {name} failed ({proc.returncode}):\n{proc.stderr[-2000:]}")
        if pmc is None:
            trace = next((work / name).rglob("*kernel_trace.csv"), None)
            if trace is None:
                raise SystemExit("rocprofv3 wrote no kernel trace")
            result.update(read_trace(trace, kernel))
        else:
            coll = next((work / name).rglob("*counter_collection.csv"), None)
            if coll is None:
                raise SystemExit(f"rocprofv3 {name} wrote no counters")
            result["counters"].update(read_counters(coll, kernel))
    if "launches" not in result:
        raise SystemExit(f"no kernel matched {kernel!r}")
    result["derived"] = derive(result["counters"], fetch_scale, write_scale)
    result["scales"] = {"fetch": fetch_scale, "write": write_scale}
    return result


def figure(p: dict, name: str) -> float | None:
    for part in (p, p.get("derived", {}), p.get("counters", {})):
        if name in part and isinstance(part[name], (int, float)):
            return float(part[name])
    return None


def digest(a: dict, b: dict) -> list[str]:
    """Side-by-side lines: the key figures first, then every counter either profile holds."""
    lines = [f"{'':<44} {'A':>14} {'B':>14} {'B/A':>8}"]

    def line(label: str, va: float | None, vb: float | None) -> str:
        fa = f"{va:.4g}" if va is not None else "-"
        fb = f"{vb:.4g}" if vb is not None else "-"
        ratio = f"{vb / va:.3f}" if va and vb is not None else "-"
        return f"{label:<44.44} {fa:>14} {fb:>14} {ratio:>8}"

    for name, what in KEY:
        lines.append(line(f"{name} ({what})", figure(a, name), figure(b, name)))
    lines.append("")
    for name in sorted(set(a.get("counters", {})) | set(b.get("counters", {}))):
        lines.append(line(name, a.get("counters", {}).get(name), b.get("counters", {}).get(name)))
    return lines


def engine_program(args) -> list[str]:
    """The `tissuemc run` command line that a `run` subcommand profiles."""
    program = [str(CLI), "run", str(args.mci.resolve()), "--photons", str(args.photons),
               "--batches", str(args.batches), "--seed", str(args.seed)]
    program += ["--arz"] * args.arz + ["--angles"] * args.angles + ["--aggregate"] * args.aggregate
    if args.time:
        program += ["--time", *args.time]
    program += ["--no-profiles"] * args.no_profiles
    return program


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m tissuemc_amd.profile",
                                 description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="command", required=True)
    r = sub.add_parser("run", help="profile one configuration of the engine")
    r.add_argument("--mci", type=Path, default=ROOT / "configs" / "three_layer_c2.mci")
    r.add_argument("--photons", type=int, default=20_000_000)
    r.add_argument("--batches", type=int, default=100)
    r.add_argument("--seed", type=int, default=1)
    r.add_argument("--arz", action="store_true")
    r.add_argument("--angles", action="store_true")
    r.add_argument("--aggregate", action="store_true")
    r.add_argument("--time", nargs=2, metavar=("NT", "DT"), help="time-of-flight bins and width (ns)")
    r.add_argument("--no-profiles", action="store_true")
    r.add_argument("--kernel", default="transport_kernel")
    r.add_argument("--fetch-scale", type=float, default=2.0)
    r.add_argument("--write-scale", type=float, default=1.0)
    r.add_argument("--output", type=Path, required=True)
    d = sub.add_parser("digest", help="print two profiles side by side")
    d.add_argument("a", type=Path)
    d.add_argument("b", type=Path)
    args = ap.parse_args(argv)

    if args.command == "digest":
        a, b = (json.loads(p.read_text()) for p in (args.a, args.b))
        print(f"A: {' '.join(a.get('program', []))}\nB: {' '.join(b.get('program', []))}")
        print("\n".join(digest(a, b)))
        return 0

    if not CLI.exists():
        raise SystemExit(f"the engine's host program is not built ({CLI}); run make -C tissuemc_amd")
    program = engine_program(args)
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        res = profile(program, args.kernel, Path(tmp), args.fetch_scale, args.write_scale)
    args.output.parent.mkdir(parents=True, exist_ok=True)
    args.output.write_text(json.dumps(res, indent=1) + "\n")
    print(f"profile: wrote {args.output}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
