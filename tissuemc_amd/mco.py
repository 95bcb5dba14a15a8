"""Reader for MCML's ASCII output files (.mco, format "A1"), as `tissuemc run --mco` writes them.

The echoed input and the RAT block are read by the reference's rules
(python/tissuemc_analysis/mcml_compare.py:212-238: the first lines starting with "InParm" and
"RAT"; the echo is the data rows between them with comments dropped; RAT is the first field of
the four lines after it). The arrays are read by section name: A_l, A_z, Rd_r, Rd_a, Tt_r, Tt_a
(one value per line, up to the next blank line) and A_rz, Rd_ra, Tt_ra (whitespace-separated
values after the section line, up to the next blank line), reshaped to [nr, nz] and [nr, na].
"""

from __future__ import annotations

import numpy as np

ONE_D = ("A_l", "A_z", "Rd_r", "Rd_a", "Tt_r", "Tt_a")
TWO_D = ("A_rz", "Rd_ra", "Tt_ra")


def _data_rows(lines: list[str]) -> list[list[str]]:
    rows = []
    for raw in lines:
        body = raw.split("#", 1)[0].split()
        if body:
            rows.append(body)
    return rows


def _block(lines: list[str], start: int) -> list[float]:
    """Numbers on the lines after `start` up to the first blank line."""
    out: list[float] = []
    for ln in lines[start + 1:]:
        if not ln.strip():
            break
        out += [float(x) for x in ln.split("#", 1)[0].split()]
    return out


def read_mco(text: str) -> dict:
    lines = text.splitlines()
    try:
        rat = next(i for i, ln in enumerate(lines) if ln.startswith("RAT"))
        inparm = next(i for i, ln in enumerate(lines) if ln.startswith("InParm"))
    except StopIteration:
        raise ValueError("not an MCML output file: no InParm/RAT section") from None
    if not lines or not lines[0].startswith("A1"):
        raise ValueError("not an MCML output file of format A1")
    r = _data_rows(lines[inparm + 1:rat])
    n_layers = int(r[4][0])
    echo = {
        "output_name": r[0][0],
        "photons": int(r[1][0]),
        "dz": float(r[2][0]), "dr": float(r[2][1]),
        "nz": int(r[3][0]), "nr": int(r[3][1]), "na": int(r[3][2]),
        "n_above": float(r[5][0]),
        "layers": [[float(x) for x in r[6 + i]] for i in range(n_layers)],
        "n_below": float(r[6 + n_layers][0]),
    }
    rsp, rd, absorbed, tt = (float(lines[rat + 1 + k].split()[0]) for k in range(4))
    out = {"echo": echo, "rsp": rsp, "rd": rd, "absorbed": absorbed, "tt": tt}
    heads = {ln.split()[0]: i for i, ln in enumerate(lines) if ln.strip() and not ln.startswith("#")}
    nz, nr, na = echo["nz"], echo["nr"], echo["na"]
    for name in ONE_D:
        if name in heads:
            out[name] = np.array(_block(lines, heads[name]))
    shapes = {"A_rz": (nr, nz), "Rd_ra": (nr, na), "Tt_ra": (nr, na)}
    for name in TWO_D:
        if name in heads:
            out[name] = np.array(_block(lines, heads[name])).reshape(shapes[name])
    return out
