"""Host model of the weight-gradient product's LDS stage and its transposing fragment reads.

`gemm_tn_group` (tissuemc_amd/csrc/mlp.hip) copies 32 batch rows x 128 columns of each operand into LDS
row by row, with the 16-byte chunks of each row XOR-swizzled, and takes the MFMA fragments (8
consecutive batch rows of one column per lane) through gfx950's `ds_read_b64_tr_b16`. This file
models both halves on the host:

* the instruction, from its lane layout as measured on the device (tests/golden/
  ds_read_tr_probe.txt, written by ds_read_tr_probe.hip beside it): within each 16 lanes, lane i
  receives element i % 4 of the 8 bytes that lane 4 j + i / 4 addresses, as its element j;
* the kernel's stage copy and fragment addresses, with the swizzle read from mlp.hip itself.

Checks: the model reproduces the device's three probe patterns; every lane of every wave
receives exactly the MFMA operand layout (row lane % 16, batch rows 8 (lane / 16) .. + 7) for all
four fragments of both halves of the wave tile; and each read keeps the 32 lanes the LDS serves
together on 64 distinct banks. Planted slips (the earlier swizzle 2 (r & 7), a wrong column or
row offset) must fail one of the checks.
"""

import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tissuemc_amd" / "csrc" / "mlp.hip"
PROBE = ROOT / "tests" / "golden" / "ds_read_tr_probe.txt"


def kernel_swizzle():
    """tn_swz as written in mlp.hip, as a Python function of the row."""
    m = re.search(r"int tn_swz\(int r\) \{ return ([^;]+); \}", SRC.read_text())
    assert m, "tn_swz not found"
    expr = m.group(1)
    assert re.fullmatch(r"[0-9r&>()*+ ]+", expr), expr
    return lambda r: eval(expr, {}, {"r": r})  # noqa: S307 - digits, r and operators only


def tr_b16(lds: np.ndarray, addr: list[int]) -> np.ndarray:
    """ds_read_b64_tr_b16 for one wave: addr[l] is lane l's element offset (a multiple of 4);
    returns the 4 values each lane receives."""
    src = np.array([lds[a:a + 4] for a in addr])
    out = np.zeros((64, 4), lds.dtype)
    for lane in range(64):
        g, i = divmod(lane, 16)
        for j in range(4):
            out[lane, j] = src[16 * g + 4 * j + i // 4, i % 4]
    return out


def read_probe():
    pats, cur = [], None
    for line in PROBE.read_text().splitlines():
        if line.startswith("pattern"):
            cur = []
            pats.append(cur)
        else:
            cur.append([int(v) for v in line.split(":")[1].split()])
    return [np.array(p) for p in pats]


def test_model_matches_the_device_probe():
    lds = np.arange(1024, dtype=np.int64)
    patterns = [[4 * l for l in range(64)],
                [64 * (l % 16) + 4 * (l // 16) for l in range(64)],
                [16 * l for l in range(64)]]
    measured = read_probe()
    assert len(measured) == 3
    for addr, got in zip(patterns, measured):
        assert np.array_equal(tr_b16(lds, addr), got)


def stage(g: np.ndarray, swz) -> np.ndarray:
    """LDS image of one operand's stage: the kernel's global_load_lds copy (wave w, load q, lane
    l -> row 8 w + 4 q + l / 16, stored chunk l % 16, fetched chunk (l % 16) ^ swz(row))."""
    lds = np.zeros(32 * 128, g.dtype)
    for w in range(4):
        for q in range(2):
            base = (8 * w + 4 * q) * 128
            for lane in range(64):
                r = 8 * w + 4 * q + lane // 16
                p = lane % 16
                c = p ^ swz(r)
                dst = base + 8 * lane
                assert dst == r * 128 + 8 * p
                lds[dst:dst + 8] = g[r, 8 * c:8 * c + 8]
    return lds


def fragment_addresses(swz, c0: int, upper: bool, col_slip=False, row_slip=False) -> list[int]:
    """The kernel's transposing-read address of every lane for the fragment at columns c0..c0+15,
    batch rows 8 g .. 8 g + 3 of its group g (upper: 8 g + 4 .. 8 g + 7)."""
    addr = []
    for lane in range(64):
        gq, L = lane >> 4, lane & 15
        row = 8 * gq + ((L & 3) if row_slip else (L >> 2)) + (4 if upper else 0)
        col = c0 + (4 * (L >> 2) if col_slip else 4 * (L & 3))
        addr.append(row * 128 + (((col >> 3) ^ swz(row)) << 3) + (col & 7))
    return addr


def fragments_ok(swz, read_swz=None, **slip) -> bool:
    """Every lane's fragments after the stage copy with swz and the reads with read_swz (the
    same by default)."""
    g = np.arange(32 * 128, dtype=np.int64).reshape(32, 128)  # batch row x column, all distinct
    lds = stage(g, swz)
    swz = read_swz or swz
    for wm in (0, 64):
        for i in range(4):
            c0 = wm + 16 * i
            lo = tr_b16(lds, fragment_addresses(swz, c0, False, **slip))
            hi = tr_b16(lds, fragment_addresses(swz, c0, True, **slip))
            frag = np.concatenate([lo, hi], axis=1)  # lane x 8 batch rows
            for lane in range(64):
                want = g[8 * (lane >> 4):8 * (lane >> 4) + 8, c0 + (lane & 15)]
                if not np.array_equal(frag[lane], want):
                    return False
    return True


def conflict_free(swz) -> bool:
    """Each read's two 32-lane halves each touch 64 distinct 4-byte banks (8 bytes a lane)."""
    for wm in (0, 64):
        for i in range(4):
            for upper in (False, True):
                addr = fragment_addresses(swz, wm + 16 * i, upper)
                for half in (range(0, 32), range(32, 64)):
                    banks = []
                    for lane in half:
                        word = addr[lane] * 2 // 4  # element offset -> 4-byte word
                        banks += [word % 64, (word + 1) % 64]
                    if len(set(banks)) != 64:
                        return False
    return True


def test_kernel_fragments_are_the_mfma_layout():
    assert fragments_ok(kernel_swizzle())


def test_kernel_reads_are_free_of_bank_conflicts():
    assert conflict_free(kernel_swizzle())


def test_planted_slips_fail():
    old = lambda r: 2 * (r & 7)  # noqa: E731 - the swizzle that measured 50% bank conflicts
    assert fragments_ok(old) and not conflict_free(old)
    swz = kernel_swizzle()
    assert not fragments_ok(swz, col_slip=True)
    assert not fragments_ok(swz, row_slip=True)
    # the reads must undo the copy's swizzle exactly; any swizzle does, but only some avoid
    # conflicts
    assert not fragments_ok(swz, read_swz=lambda r: 2 * (r & 7))
    odd = lambda r: 2 * (r & 3) + 1  # noqa: E731
    assert fragments_ok(odd) and not conflict_free(odd)


@pytest.mark.parametrize("r", range(32))
def test_swizzle_range(r):
    s = kernel_swizzle()(r)
    assert 0 <= s < 16 and s % 2 == 0
