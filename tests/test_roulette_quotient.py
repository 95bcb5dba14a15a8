"""The transport kernel's roulette quotient w / 0.1f is the IEEE quotient.

A photon that survives the roulette (physics.hpp:250-259) has its weight divided by CHANCE =
0.1f. The kernel (tissuemc_amd/csrc/transport.hip, roulette) forms it as q = w * 10 followed by
one correction with the exact residual, fma(fma(-q, 0.1f, w), 10, q), for weights at or above a
threshold, and divides below it. This checks, exhaustively over every float in [threshold, 1e-4)
(the weights that reach the roulette), that the two-fma form equals the correctly rounded
quotient, and that below 2^-125 it would not, so the threshold is needed. The threshold is read
from the kernel source, so the two cannot drift apart. gcc, no GPU.
"""

import re
import shutil
import struct
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

CHECK = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float quotient(float w) {
    const float q = w * 10.0f;
    return fmaf(fmaf(-q, 0.1f, w), 10.0f, q);
}

int main(int argc, char **argv) {
    const uint32_t lo = (uint32_t)strtoul(argv[1], 0, 0), hi = (uint32_t)strtoul(argv[2], 0, 0);
    unsigned long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static, 1 << 20)
    for (long long u = lo; u < (long long)hi; ++u) {
        const uint32_t b = (uint32_t)u;
        float w, ref, got;
        memcpy(&w, &b, 4);
        volatile float d = 0.1f;
        ref = w / d;
        got = quotient(w);
        if (memcmp(&ref, &got, 4) != 0) ++bad;
    }
    printf("%llu\n", bad);
    return 0;
}
"""


def bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def kernel_threshold() -> float:
    src = (ROOT / "tissuemc_amd" / "csrc" / "transport.hip").read_text()
    m = re.search(r"if \(w >= (0x1p-\d+)f\) \{\s*const float q10 = w \* 10\.0f;", src)
    assert m, "roulette quotient not found in the kernel"
    return float.fromhex(m.group(1))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("quotient")
    (d / "q.c").write_text(CHECK)
    flags = ["-O2", "-ffp-contract=off", "-fno-fast-math"]
    try:
        subprocess.run(["gcc", *flags, "-fopenmp", str(d / "q.c"), "-o", str(d / "q"), "-lm"],
                       check=True, capture_output=True)
    except subprocess.CalledProcessError:
        subprocess.run(["gcc", *flags, str(d / "q.c"), "-o", str(d / "q"), "-lm"],
                       check=True, capture_output=True)

    def count(lo: int, hi: int) -> int:
        out = subprocess.run([str(d / "q"), str(lo), str(hi)], check=True, capture_output=True,
                             text=True, timeout=600)
        return int(out.stdout)

    return count


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_two_fma_quotient_is_exact_over_every_roulette_weight(checker):
    t = kernel_threshold()
    assert t <= 2.0 ** -60
    # every float w with threshold <= w < 1e-4f
    assert checker(bits(t), bits(1e-4)) == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_two_fma_quotient_fails_where_the_residual_underflows(checker):
    # w in [2^-149, 2^-125): the residual falls below the normal range and the form is off
    assert checker(1, bits(2.0 ** -125)) > 0
