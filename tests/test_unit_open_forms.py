"""The transport kernel's uniform deviate and roulette test equal the reference's float mapping.

The reference maps a 32-bit Philox word x to ((x >> 9) + 0.5) * 2^-23 in float
(include/tissuemc/philox.hpp:133-136). The kernel (tissuemc_amd/csrc/device_rng.h, unit_open)
forms it as the float with bits 0x7F << 23 | x >> 9 (one v_alignbit_b32), minus 1 - 2^-24, and
makes the roulette draw unit_open(x) < 0.1f (physics.hpp:250-259) an integer comparison of x with
kUnitBelowTenth. The constants are read from the kernel source, and both forms are checked
against the exact value over every one of the 2^23 deviates (the low 9 bits of x do not take
part). numpy, no GPU.
"""

import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SRC = (ROOT / "tissuemc_amd" / "csrc" / "device_rng.h").read_text()


def _constant(name: str) -> str:
    m = re.search(r"constexpr \w+ " + name + r" = ([0-9a-fA-Fx.p+-]+?)[uf]?;", SRC)
    assert m, f"{name} not found in device_rng.h"
    return m.group(1)


EXPONENT = int(_constant("kUnitExponent"), 0)
OFFSET = np.float32(float.fromhex(_constant("kUnitOffset")))
BELOW_TENTH = int(_constant("kUnitBelowTenth"), 0)
M = np.arange(1 << 23, dtype=np.uint32)  # x >> 9 for every x


def exact() -> np.ndarray:
    """(m + 0.5) 2^-23, exactly (float64), which float32 also holds exactly."""
    return (M.astype(np.float64) + 0.5) * 2.0 ** -23


def kernel_form(exponent: int, offset: np.float32) -> np.ndarray:
    f = (M | np.uint32(exponent << 23)).view(np.float32)
    return (f - offset).astype(np.float32)


def test_alignbit_form_matches_reference_mapping_everywhere():
    assert re.search(r"__builtin_amdgcn_alignbit\(expo, x, 9u\)", SRC)
    got = kernel_form(EXPONENT, OFFSET)
    assert np.array_equal(got.astype(np.float64), exact())
    assert got.min() > 0 and got.max() < 1


def test_a_slipped_constant_fails():
    # one ulp off in the offset, or the exponent of [2, 4), changes deviates
    assert not np.array_equal(kernel_form(EXPONENT, np.nextafter(OFFSET, np.float32(0))).astype(np.float64), exact())
    assert not np.array_equal(kernel_form(EXPONENT + 1, OFFSET).astype(np.float64), exact())


def test_roulette_word_limit_is_the_float_comparison():
    ref = exact().astype(np.float32) < np.float32(0.1)
    assert BELOW_TENTH % 512 == 0
    for low in (0, 511):  # the extremes of the bits unit_open drops
        x = (M.astype(np.uint64) << 9 | low).astype(np.uint32)
        assert np.array_equal(x < np.uint32(BELOW_TENTH), ref)
    # a limit one deviate lower or higher would not be
    for wrong in (BELOW_TENTH - 512, BELOW_TENTH + 512):
        assert not np.array_equal((M.astype(np.uint64) << 9) < wrong, ref)
    src = (ROOT / "tissuemc_amd" / "csrc" / "transport.hip").read_text()
    assert "rw < kUnitBelowTenth" in src
