"""The ctypes mirrors of the C ABI's structs have the header's layout.

tissuemc_amd/engine.py and tissuemc_amd/mlp.py declare every struct of include/tissuemc_hip.h as a
ctypes.Structure. A field missing, misordered or of the wrong width would not fail loudly: the
library would read the wrong bytes. This compiles a small C program against the header (gcc, no
GPU needed) that prints sizeof of each struct and offsetof of each field named in the ctypes
mirror, and compares them with ctypes' own numbers.
"""

import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

import tissuemc_amd.engine as engine
import tissuemc_amd.mlp as mlp

ROOT = Path(__file__).resolve().parents[1]

MIRRORS = {
    "tmc_layer": engine.TmcLayer,
    "tmc_spec": engine.TmcSpec,
    "tmc_opts": engine.TmcOpts,
    "tmc_out": engine.TmcOut,
    "tmc_stack": engine.TmcStack,
    "tmc_batched_opts": engine.TmcBatchedOpts,
    "tmc_batched_out": engine.TmcBatchedOut,
    "tmc_stack_dist": engine.TmcStackDist,
    "tmc_tune_entry": engine.TmcTuneEntry,
    "tmc_tune_report": engine.TmcTuneReport,
    "tmc_mlp_config": mlp.MlpConfig,
    "tmc_surrogate_gate": mlp.SurrogateGate,
    "tmc_surrogate_answers": mlp.SurrogateAnswers,
}


def c_layout(tmp_path: Path) -> dict[str, dict[str, int]]:
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "tissuemc_hip.h"',
             "int main(void) {"]
    for cname, cls in MIRRORS.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, *_ in cls._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout: dict[str, dict[str, int]] = {}
    for line in out.splitlines():
        cname, key, value = line.split()
        layout.setdefault(cname, {})[key] = int(value)
    return layout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_the_header(tmp_path):
    layout = c_layout(tmp_path)
    for cname, cls in MIRRORS.items():
        got = layout[cname]
        assert ctypes.sizeof(cls) == got["sizeof"], cname
        for field, *_ in cls._fields_:
            assert getattr(cls, field).offset == got[field], f"{cname}.{field}"
