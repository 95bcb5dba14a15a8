"""Validate the GPU engine against published values and the H-function, with run invariants.

    python -m tissuemc_amd.validate [--cases v1_slab,v3_three_layer] [--photons N] [--seeds 1,2]

For every case in configs/references.json the engine runs the case's input, and each published
value is checked with the 3-sigma rule of tissuemc_amd.stats.compare. Every run must also conserve
energy (rd + tt + absorbed + specular = 1 within 3 standard errors plus a float32 rounding floor)
and report no anomalies. The flow and the rules restate python/tissuemc_analysis/validate.py
(evaluate_case :111-131, invariant_results :134-153, CONSERVATION_FLOOR :23) of the reference;
the engine is called in-process through the C ABI instead of through a subprocess.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
from dataclasses import dataclass
from pathlib import Path

from . import hfunction, stats

CONFIGS = Path(__file__).resolve().parent / "configs"
CONSERVATION_FLOOR = {"float64": 1e-12, "float32": 1e-6}


@dataclass(frozen=True)
class Check:
    case: str
    seed: int
    quantity: str
    source: str
    comparison: stats.Comparison

    def line(self) -> str:
        c = self.comparison
        flag = "PASS" if c.passed else "FAIL"
        return (f"  [{flag}] {self.quantity:<17} ours {c.ours:.6f} +/- {c.se_ours:.6f}   "
                f"ref {c.reference:.6f}   z = {c.z:+.2f}   {self.source}")


def references() -> dict:
    return json.loads((CONFIGS / "references.json").read_text())["cases"]


def estimates(run: dict) -> dict[str, stats.Estimate]:
    """Estimates recomputed from the raw batch sums, independently of the engine's summary."""
    b = run["batches"]
    ph = [x["photons"] for x in b]
    rd = stats.batch_estimate([x["rd"] for x in b], ph)
    tt = stats.batch_estimate([x["tt"] for x in b], ph)
    return {"rd": rd, "tt": tt, "rd_plus_specular": stats.Estimate(rd.mean + run["specular"], rd.se)}


def hfunction_value(run: dict, quantity: str) -> float:
    layers = run["layers"]
    if len(layers) != 1 or layers[0]["g"] != 0:
        raise ValueError("the H-function reference needs one isotropic layer")
    L = layers[0]
    r = hfunction.semi_infinite_isotropic(L["mua"], L["mus"], L["n"], run["n_above"])
    return {"rd": r.diffuse, "rd_plus_specular": r.total}[quantity]


def evaluate_case(name: str, case: dict, run: dict, seed: int) -> list[Check]:
    est = estimates(run)
    out = []
    for chk in case["checks"]:
        q = chk["q"]
        ours = est[q]
        if chk.get("hfunction"):
            ref, se_ref, rounding = hfunction_value(run, q), 0.0, 0.0
        else:
            ref = chk["value"]
            rounding = 0.5 * 10 ** (-chk["decimals"]) if "decimals" in chk else 0.0
            if "se" in chk:
                se_ref = chk["se"]
            elif "photons" in chk:
                se_ref = stats.estimated_reference_se(ours.se, run["photons"], chk["photons"])
            else:
                se_ref = 0.0
        out.append(Check(name, seed, q, chk["ref"], stats.compare(ours.mean, ours.se, ref, se_ref,
                                                                 rounding)))
    return out


def invariants(name: str, run: dict, seed: int) -> list[Check]:
    floor = CONSERVATION_FLOOR[run["precision"]]
    b = run["batches"]
    ph = [x["photons"] for x in b]
    tot = stats.batch_estimate([x["rd"] + x["tt"] + math.fsum(x["absorbed"]) for x in b], ph)
    resid = tot.mean + run["specular"] - 1.0
    out = [Check(name, seed, "conservation", "rd + tt + absorbed + rsp = 1",
                 stats.compare(resid, tot.se, 0.0, 0.0, rounding=floor)),
           Check(name, seed, "anomalies", "non-finite weights or non-unit directions",
                 stats.compare(float(run["anomalies"]), 0.0, 0.0))]
    return out


def validate_run(name: str, case: dict, run: dict, seed: int) -> list[Check]:
    return evaluate_case(name, case, run, seed) + invariants(name, run, seed)


def main(argv: list[str] | None = None) -> int:
    from . import engine

    ap = argparse.ArgumentParser(prog="tissuemc-validate", description=__doc__.split("\n\n")[0])
    ap.add_argument("--cases", help="comma-separated case names (default: all)")
    ap.add_argument("--photons", type=int, help="photons per case (default: the case's)")
    ap.add_argument("--batches", type=int, default=50)
    ap.add_argument("--seeds", default="1")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--report", type=Path)
    args = ap.parse_args(argv)
    refs = references()
    names = args.cases.split(",") if args.cases else list(refs)
    unknown = [n for n in names if n not in refs]
    if unknown:
        raise SystemExit(f"unknown case(s): {', '.join(unknown)}; known: {', '.join(refs)}")
    results: list[Check] = []
    for name in names:
        case = refs[name]
        spec = engine.parse_mci(path=CONFIGS / case["input"])[0]
        print(name)
        for seed in (int(s) for s in args.seeds.split(",")):
            run = engine.run(spec, photons=args.photons or case["photons"], batches=args.batches,
                             seed=seed, device=args.device)
            checks = validate_run(name, case, run, seed)
            for c in checks:
                print(c.line())
            results += checks
    failed = [c for c in results if not c.comparison.passed]
    print(f"\n{len(results) - len(failed)}/{len(results)} checks passed")
    if args.report:
        args.report.write_text(json.dumps(
            [{"case": c.case, "seed": c.seed, "quantity": c.quantity, "source": c.source,
              **c.comparison.__dict__} for c in results], indent=1) + "\n")
    return 0 if not failed else 1


if __name__ == "__main__":
    sys.exit(main())
