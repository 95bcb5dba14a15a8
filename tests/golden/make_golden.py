"""Write the golden fixtures: outputs of the reference's own CPU engine.

The engine is built from the reference's sources by ``make -C oracle ref`` (oracle/Makefile)
into oracle/_ref/tissuemc_ref; this script runs it on the workloads in tissuemc_amd/configs and
keeps the per-batch sums, event counts and profiles of each run as JSON. The fixtures are data:
inputs (case, photons, batches, seed, precision) and the reference's outputs.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "ref"
CONFIGS = ROOT / "tissuemc_amd" / "configs"
REF_BIN = ROOT / "oracle" / "_ref" / "tissuemc_ref"

# (case, photons, batches, seed, precision)
RUNS = [
    (case, photons, 10, seed, prec)
    for case, photons in [("v1_slab", 20000), ("v2_semi_infinite", 20000),
                          ("o1_semi_infinite_matched", 20000), ("v3_three_layer", 20000),
                          ("tissue_g09", 4000), ("slab_c1", 20000), ("three_layer_c2", 10000)]
    for seed, prec in [(1, "float64"), (2, "float64"), (1, "float32")]
] + [
    ("v3_three_layer", 20003, 7, 3, "float64"),   # ragged batches
    ("v3_three_layer", 5000, 1, 4, "float64"),    # one batch
    ("v3_three_layer", 40, 40, 5, "float64"),     # one photon per batch
    ("slab_c1", 20003, 7, 3, "float32"),
]

KEEP = ("engine", "precision", "seed", "photons", "specular", "events", "anomalies", "summary",
        "batches", "a_z", "rd_r", "grid")


def main() -> int:
    if not REF_BIN.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True)
    OUT.mkdir(parents=True, exist_ok=True)
    for case, photons, batches, seed, prec in RUNS:
        with tempfile.TemporaryDirectory() as tmp:
            out = Path(tmp) / "r.json"
            subprocess.run([str(REF_BIN), "run", str(CONFIGS / f"{case}.mci"), "--photons",
                            str(photons), "--batches", str(batches), "--seed", str(seed),
                            "--threads", "0", "--precision", prec, "--output", str(out)],
                           check=True, stdout=subprocess.DEVNULL)
            run = json.loads(out.read_text())["runs"][0]
        fixture = {"case": case, "batch_count": batches, **{k: run[k] for k in KEEP}}
        name = f"{case}_n{photons}_b{batches}_s{seed}_{prec}.json"
        (OUT / name).write_text(json.dumps(fixture, separators=(",", ":")) + "\n")
        print(name, run["events"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
