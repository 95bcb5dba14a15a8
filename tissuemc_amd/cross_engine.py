"""Compare two engines on identical photons.

Both engines give photon i the stream (seed, i) and split the photons into the same batches, so
a correct implementation agrees far more closely than statistical noise does. The limits are the
reference's, fixed before any GPU result existed (python/tissuemc_analysis/cross_engine.py:43-51),
and so are the quantities compared (compare_runs :95-121, _profile_gaps :75-92): the batch-means
estimate of each total in standard errors, the worst per-batch relative gap, the event count,
anomalies, and for each profile its total and its cumulative shape. The profiles here are the
reference's a_z and rd_r plus T_r (compared like rd_r) and A(r,z), compared twice: flattened row
by row (depth-major, so the cumulative shape runs through every cell) and as its radial marginal
sum over depth.

One command runs both engines: the CPU engine is any binary with the reference CLI's ``run``
flags (the reference's own ``tissuemc`` build), the other is this package's HIP engine.

    python -m tissuemc_amd.cross_engine --case v3_three_layer --cpu-engine /path/to/tissuemc
    python -m tissuemc_amd.cross_engine --case v3_three_layer --a-json cpu_result.json
"""

from __future__ import annotations

import argparse
import json
import math
import subprocess
import sys
import tempfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import stats

MAX_TOTAL_SE = 0.25          # gap between totals, in standard errors
MAX_BATCH_REL = 5e-3         # relative gap in any one batch's sum
MAX_EVENTS_REL = 1e-3        # relative gap in absorption events
MAX_PROFILE_TOTAL_REL = 1e-4  # relative gap between the totals of a profile
MAX_PROFILE_SHAPE = 1e-3     # largest gap between cumulative profiles, as a fraction of the total

PROFILE_KEYS = ("a_z", "rd_r", "t_r", "a_rz", "a_rz_r")


@dataclass(frozen=True)
class Gap:
    name: str
    value: float
    limit: float

    @property
    def passed(self) -> bool:
        return self.value <= self.limit

    def line(self) -> str:
        return (f"  [{'PASS' if self.passed else 'FAIL'}] {self.name:<34} {self.value:.3e}"
                f"  (limit {self.limit:.1e})")


def batch_sums(run: dict, key: str) -> np.ndarray:
    """Per-batch sums of ``rd``, ``tt`` or ``absorbed`` (all layers) as float64."""
    if key == "absorbed":
        return np.array([math.fsum(b["absorbed"]) for b in run["batches"]])
    return np.array([float(b[key]) for b in run["batches"]])


def profile(run: dict, key: str) -> np.ndarray | None:
    """A profile of a result as a 1-D float64 array, or None if the result lacks it. ``a_rz``
    is A(r,z) flattened depth-major; ``a_rz_r`` its sum over depth."""
    base = "a_rz" if key in ("a_rz", "a_rz_r") else key
    v = run.get(base)
    if v is None:
        return None
    arr = np.asarray(v, dtype=np.float64)
    if key == "a_rz_r":
        return arr.sum(axis=0) if arr.ndim == 2 else None
    return arr.reshape(-1)


def profile_gaps(a: dict, b: dict, key: str) -> list[Gap]:
    pa, pb = profile(a, key), profile(b, key)
    if pa is None or pb is None or pa.size == 0 or pa.shape != pb.shape:
        return []
    total = math.fsum(pa)
    if total == 0:
        return []
    shape = float(np.max(np.abs(np.cumsum(pa) - np.cumsum(pb)))) / abs(total)
    return [Gap(f"{key}: total relative gap", abs(total - math.fsum(pb)) / abs(total),
                MAX_PROFILE_TOTAL_REL),
            Gap(f"{key}: worst cumulative gap / total", shape, MAX_PROFILE_SHAPE)]


def _total_gaps(key: str, sa: np.ndarray, sb: np.ndarray, photons: list[int]) -> list[Gap]:
    ea, eb = stats.batch_estimate(sa, photons), stats.batch_estimate(sb, photons)
    se = max(ea.se, eb.se)
    if se > 0:
        in_se = abs(ea.mean - eb.mean) / se
    else:
        in_se = 0.0 if ea.mean == eb.mean else math.inf
    nz = sa != 0
    worst = float(np.max(np.abs(sa[nz] - sb[nz]) / np.abs(sa[nz]))) if nz.any() else 0.0
    return [Gap(f"{key}: total gap / standard error", in_se, MAX_TOTAL_SE),
            Gap(f"{key}: worst per-batch relative gap", worst, MAX_BATCH_REL)]


def compare_runs(a: dict, b: dict, profiles: bool = True,
                 profile_keys: tuple[str, ...] = PROFILE_KEYS) -> list[Gap]:
    """Gaps between two results that simulated the same photons in the same batches. Profiles
    that either result lacks are skipped."""
    photons = [x["photons"] for x in a["batches"]]
    if photons != [x["photons"] for x in b["batches"]]:
        raise ValueError("the two runs do not use the same batches")
    gaps: list[Gap] = []
    for key in ("rd", "tt", "absorbed"):
        gaps += _total_gaps(key, batch_sums(a, key), batch_sums(b, key), photons)
    ev_a, ev_b = a["events"], b["events"]
    gaps.append(Gap("events: relative gap", abs(ev_a - ev_b) / ev_a if ev_a else 0.0,
                    MAX_EVENTS_REL))
    gaps.append(Gap("anomalies (both runs)", float(a["anomalies"] + b["anomalies"]), 0.0))
    if profiles:
        for key in profile_keys:
            gaps += profile_gaps(a, b, key)
    return gaps


def run_cpu_engine(binary: Path, mci: Path, *, photons: int, batches: int, seed: int,
                   precision: str) -> dict:
    """First run of ``binary run`` (the reference CLI's flags: src/cli/main.cpp:53-117) as a
    result dict."""
    with tempfile.TemporaryDirectory() as tmp:
        out = Path(tmp) / "cpu.json"
        cmd = [str(binary), "run", str(mci), "--photons", str(photons), "--batches",
               str(batches), "--seed", str(seed), "--threads", "0", "--precision", precision,
               "--output", str(out)]
        proc = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
        if proc.returncode != 0:
            raise SystemExit(f"CPU engine failed ({proc.returncode}): {proc.stderr.strip()}")
        doc = json.loads(out.read_text())
    return doc["runs"][0] if "runs" in doc else doc


def main(argv: list[str] | None = None) -> int:
    from . import engine

    ap = argparse.ArgumentParser(prog="tissuemc-cross-engine", description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", default="v3_three_layer", help="input under tissuemc_amd/configs")
    ap.add_argument("--mci", type=Path, help="an .mci file instead of --case (its first run)")
    ap.add_argument("--photons", type=int, default=2_000_000)
    ap.add_argument("--batches", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--cpu-engine", type=Path,
                     help="CPU engine binary with the reference CLI (run --photons --batches "
                          "--seed --threads --precision --output)")
    src.add_argument("--a-json", type=Path,
                     help="engine A's result (a JSON file in the reference schema, same photons)")
    ap.add_argument("--a-precision", default="float32", choices=["float64", "float32"])
    ap.add_argument("--no-profiles", action="store_true",
                    help="run the HIP engine as the profiles-off diagnostic; totals must agree")
    ap.add_argument("--report", type=Path, help="write the gaps to this JSON file")
    args = ap.parse_args(argv)

    mci = args.mci or Path(__file__).resolve().parent / "configs" / f"{args.case}.mci"
    spec = engine.parse_mci(path=mci)[0]
    if args.cpu_engine:
        a = run_cpu_engine(args.cpu_engine, mci, photons=args.photons, batches=args.batches,
                           seed=args.seed, precision=args.a_precision)
    else:
        doc = json.loads(args.a_json.read_text())
        a = doc["runs"][0] if "runs" in doc else doc
    b = engine.run(spec, photons=args.photons, batches=args.batches, seed=args.seed,
                   profiles=not args.no_profiles)
    print(f"{mci.stem}: {args.photons} photons, {args.batches} batches, seed {args.seed}")
    print(f"  A = {'CPU engine ' + str(args.cpu_engine) if args.cpu_engine else args.a_json}"
          f"    B = HIP engine on {b['device']}")
    gaps = compare_runs(a, b, profiles=not args.no_profiles)
    for g in gaps:
        print(g.line())
    ok = all(g.passed for g in gaps)
    print("\nengines agree on identical photons" if ok else "\nENGINES DISAGREE: investigate")
    if args.report:
        args.report.parent.mkdir(parents=True, exist_ok=True)
        args.report.write_text(json.dumps({
            "input": str(mci), "photons": args.photons, "batches": args.batches,
            "seed": args.seed, "device": b["device"], "profiles": not args.no_profiles,
            "gaps": [g.__dict__ | {"passed": g.passed} for g in gaps], "agree": ok}, indent=1))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
