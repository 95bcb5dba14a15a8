"""Compare two engines on identical photons.

Both engines give photon i the stream (seed, i) and split photons into the same batches, so a
correct implementation agrees far more closely than statistical noise. The limits are the
reference's, fixed before any GPU result existed (python/tissuemc_analysis/cross_engine.py:43-51);
the comparison rules follow compare_runs (:95-121) and _profile_gaps (:75-92). T_r is compared
like the reference's rd_r.

    python -m tissuemc_amd.cross_engine --case v3_three_layer --photons 2000000
"""

from __future__ import annotations

import argparse
import itertools
import json
import math
import sys
from dataclasses import dataclass
from pathlib import Path

from . import stats

MAX_TOTAL_SE = 0.25          # gap between totals, in standard errors
MAX_BATCH_REL = 5e-3         # relative gap in any one batch's sum
MAX_EVENTS_REL = 1e-3        # relative gap in absorption events
MAX_PROFILE_TOTAL_REL = 1e-4  # relative gap between the totals of a profile
MAX_PROFILE_SHAPE = 1e-3     # largest gap between cumulative profiles, as a fraction of the total


@dataclass(frozen=True)
class Gap:
    name: str
    value: float
    limit: float

    @property
    def passed(self) -> bool:
        return self.value <= self.limit

    def line(self) -> str:
        flag = "PASS" if self.passed else "FAIL"
        return f"  [{flag}] {self.name:<34} {self.value:.3e}  (limit {self.limit:.1e})"


def _sums(run: dict, key: str) -> list[float]:
    if key == "absorbed":
        return [math.fsum(b["absorbed"]) for b in run["batches"]]
    return [b[key] for b in run["batches"]]


def profile_gaps(a: dict, b: dict, key: str) -> list[Gap]:
    pa = [float(x) for x in (a.get(key) if a.get(key) is not None else [])]
    pb = [float(x) for x in (b.get(key) if b.get(key) is not None else [])]
    if not pa or not pb or len(pa) != len(pb):
        return []
    total = math.fsum(pa)
    if total == 0:
        return []
    cum_a, cum_b = list(itertools.accumulate(pa)), list(itertools.accumulate(pb))
    shape = max(abs(x - y) for x, y in zip(cum_a, cum_b, strict=True)) / abs(total)
    return [
        Gap(f"{key}: total relative gap", abs(total - math.fsum(pb)) / abs(total),
            MAX_PROFILE_TOTAL_REL),
        Gap(f"{key}: worst cumulative gap / total", shape, MAX_PROFILE_SHAPE),
    ]


def compare_runs(a: dict, b: dict, profiles: bool = True,
                 profile_keys: tuple[str, ...] = ("a_z", "rd_r", "t_r")) -> list[Gap]:
    """Gaps between two results that simulated the same photons in the same batches."""
    if [x["photons"] for x in a["batches"]] != [x["photons"] for x in b["batches"]]:
        raise ValueError("the two runs do not use the same batches")
    gaps: list[Gap] = []
    photons = [x["photons"] for x in a["batches"]]
    for key in ("rd", "tt", "absorbed"):
        sa, sb = _sums(a, key), _sums(b, key)
        ea = stats.batch_estimate(sa, photons)
        eb = stats.batch_estimate(sb, photons)
        se = max(ea.se, eb.se)
        gap = abs(ea.mean - eb.mean) / se if se > 0 else (0.0 if ea.mean == eb.mean else math.inf)
        gaps.append(Gap(f"{key}: total gap / standard error", gap, MAX_TOTAL_SE))
        rel = [abs(x - y) / abs(x) for x, y in zip(sa, sb, strict=True) if x != 0]
        gaps.append(Gap(f"{key}: worst per-batch relative gap", max(rel) if rel else 0.0,
                        MAX_BATCH_REL))
    ev_a, ev_b = a["events"], b["events"]
    gaps.append(Gap("events: relative gap", abs(ev_a - ev_b) / ev_a if ev_a else 0.0,
                    MAX_EVENTS_REL))
    gaps.append(Gap("anomalies (both runs)", float(a["anomalies"] + b["anomalies"]), 0.0))
    if profiles:
        for key in profile_keys:
            gaps += profile_gaps(a, b, key)
    return gaps


def main(argv: list[str] | None = None) -> int:
    from . import engine

    ap = argparse.ArgumentParser(prog="tissuemc-cross-engine", description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", default="v3_three_layer")
    ap.add_argument("--photons", type=int, default=2_000_000)
    ap.add_argument("--batches", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--a-json", type=Path, required=True,
                    help="engine A's result (a JSON file in the reference schema, same photons)")
    ap.add_argument("--report", type=Path)
    args = ap.parse_args(argv)
    a = json.loads(args.a_json.read_text())
    a = a["runs"][0] if "runs" in a else a
    cfg = Path(__file__).resolve().parent / "configs" / f"{args.case}.mci"
    spec = engine.parse_mci(path=cfg)[0]
    b = engine.run(spec, photons=args.photons, batches=args.batches, seed=args.seed)
    gaps = compare_runs(a, b)
    for g in gaps:
        print(g.line())
    ok = all(g.passed for g in gaps)
    print("\nengines agree on identical photons" if ok else "\nENGINES DISAGREE: investigate")
    if args.report:
        args.report.write_text(json.dumps([g.__dict__ | {"passed": g.passed} for g in gaps], indent=1))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
