"""Host model of a 64-lane wave running the transport kernel's lane aggregation.

The kernel's TMC_FLAG_AGGREGATE path (tissuemc_amd/csrc/transport.hip: wave_sum_u32,
wave_sum_fixed, aggregate_lds_add, aggregate_tile_add) is the wave64 counterpart of the
reference's aggregated_add (include/tissuemc/gpu_tally.hpp:200-234), which the reference tests on
the host with a 32-lane warp model (tests/cpp/test_gpu_emulation.cpp:670-792). This model follows
the same plan for a 64-lane wave:

* DPP is modelled from the GCN/CDNA definition of the controls the kernel uses: row_shr:n (lane l
  reads lane l-n of its 16-lane row; a source before the row start is out of bounds), row_bcast:15
  (every lane of row r reads lane 15 of row r-1) and row_bcast:31 (rows 2 and 3 read lane 31);
  row_mask / bank_mask select which rows / 4-lane banks are written; an out-of-bounds source writes
  0 with bound_ctrl, and leaves the lane unwritten otherwise.
* The DPP steps and the number of peeled groups are read from transport.hip itself, so the test
  checks the kernel's constants rather than a copy of them.

Checks: the reduction leaves the exact 64-lane sum in lane 63 (and a plausible slip in a mask or a
shift does not); the fixed-point split is exact up to weights of 2^32; one add per peeled group,
by the group's lowest lane, carrying the group's whole weight; lanes outside the score, even with
the same bin, add nothing; the A(r,z) tile's 32-bit counters with their carries into 64-bit bins
give exact totals.
"""

import re
from pathlib import Path

import numpy as np
import pytest

SRC = Path(__file__).resolve().parents[1] / "tissuemc_amd" / "csrc" / "transport.hip"
M32 = 0xFFFFFFFF


def kernel_dpp_steps():
    """The (source, ctrl, row_mask, bank_mask, bound_ctrl) steps of wave_sum_u32, in order."""
    text = SRC.read_text()
    body = text[text.index("uint32_t wave_sum_u32("):]
    body = body[:body.index("\n}\n")]
    steps = re.findall(r"t \+= __builtin_amdgcn_update_dpp\(0, (x|t), (0x[0-9a-f]+), (0x[0-9a-f]+), "
                       r"(0x[0-9a-f]+), (true|false)\)", body)
    assert steps, "wave_sum_u32's DPP steps not found"
    assert "readlane(t, 63)" in body
    return [(src, int(c, 16), int(rm, 16), int(bm, 16), bc == "true") for src, c, rm, bm, bc in steps]


def kernel_peels():
    return int(re.search(r"constexpr int kAggPeels = (\d+);", SRC.read_text()).group(1))


def dpp_read(src, ctrl, row_mask, bank_mask, bound_ctrl, old=0):
    """update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) on a full 64-lane wave."""
    out = [old] * 64
    for lane in range(64):
        row, pos = divmod(lane, 16)
        if not (row_mask >> row) & 1 or not (bank_mask >> (pos // 4)) & 1:
            continue  # lane not written: keeps old
        if 0x111 <= ctrl <= 0x11F:  # row_shr:n
            n = ctrl - 0x110
            source = lane - n if pos >= n else None
        elif ctrl == 0x142:  # row_bcast:15
            source = 16 * row - 1 if row >= 1 else None
        elif ctrl == 0x143:  # row_bcast:31
            source = 31 if row >= 2 else None
        else:
            raise ValueError(f"DPP control {ctrl:#x} is not modelled")
        if source is None:
            if bound_ctrl:
                out[lane] = 0
            continue
        out[lane] = src[source]
    return out


def wave_sum_u32(v, steps=None):
    """The kernel's wave_sum_u32: the value lane 63 holds after its DPP steps."""
    steps = kernel_dpp_steps() if steps is None else steps
    x = [int(a) & M32 for a in v]
    t = list(x)
    for src, ctrl, rm, bm, bc in steps:
        d = dpp_read(x if src == "x" else t, ctrl, rm, bm, bc)
        t = [(a + b) & M32 for a, b in zip(t, d)]
    return t[63]


def wave_sum_fixed(q):
    lo = wave_sum_u32([int(a) & 0xFFFF for a in q])
    hi = wave_sum_u32([int(a) >> 16 for a in q])
    return (hi << 16) + lo


def aggregate_adds(score, keys, q, peels=None):
    """The kernel's aggregate_lds_add for one wave: returns the adds as (bin, lane, value)."""
    peels = kernel_peels() if peels is None else peels
    pending = list(score)
    rest = [l for l in range(64) if pending[l]]
    adds = []
    for _ in range(peels):
        if not rest:
            break
        leader = rest[0]  # ctz of the ballot
        kb = keys[leader]  # readlane
        inside = [pending[l] and keys[l] == kb for l in range(64)]
        total = wave_sum_fixed([q[l] if inside[l] else 0 for l in range(64)])
        adds.append((kb, leader, total))
        pending = [p and not i for p, i in zip(pending, inside)]
        rest = [l for l in rest if not inside[l]]
    adds += [(keys[l], l, q[l]) for l in range(64) if pending[l]]
    return adds


def skewed_wave(rng, bins=40, max_q=1 << 32):
    score = [bool(b) for b in rng.integers(0, 2, 64)]
    if not any(score):
        score[int(rng.integers(0, 64))] = True
    r = rng.integers(0, 100, 64)
    keys = [0 if x < 40 else (1 if x < 60 else int(rng.integers(0, bins))) for x in r]
    q = [int(x) for x in rng.integers(1, max_q + 1, 64, dtype=np.uint64)]
    return score, keys, q


def check_aggregation(score, keys, q, bins=40):
    adds = aggregate_adds(score, keys, q)
    table = [0] * bins
    for b, _, v in adds:
        table[b] += v
    expected = [0] * bins
    for l in range(64):
        if score[l]:
            expected[keys[l]] += q[l]
    assert table == expected
    # the peeled groups: each is the lowest remaining scoring lane's bin, added once by that lane
    pending = [l for l in range(64) if score[l]]
    n_groups = 0
    for b, lane, v in adds[:kernel_peels()]:
        if not pending:
            break
        assert lane == pending[0] and b == keys[lane]
        members = [l for l in pending if keys[l] == b]
        assert v == sum(q[l] for l in members)
        pending = [l for l in pending if keys[l] != b]
        n_groups += 1
    assert len(adds) == n_groups + len(pending)
    return adds


def test_kernel_reduction_is_the_modelled_sequence():
    steps = kernel_dpp_steps()
    assert [c for _, c, *_ in steps] == [0x111, 0x112, 0x113, 0x114, 0x118, 0x142, 0x143]
    assert kernel_peels() >= 1


def test_wave_sum_is_exact_in_lane_63():
    rng = np.random.default_rng(5)
    for _ in range(200):
        v = [int(x) for x in rng.integers(0, 1 << 26, 64)]
        assert wave_sum_u32(v) == sum(v) & M32
    for lane in range(64):  # a single non-zero lane anywhere reaches lane 63
        v = [0] * 64
        v[lane] = 12345
        assert wave_sum_u32(v) == 12345
    assert wave_sum_u32([1] * 64) == 64


def sums_differ(steps, trials=50):
    rng = np.random.default_rng(9)
    return sum(wave_sum_u32(v, steps) != sum(v)
               for v in ([int(x) for x in rng.integers(0, 1000, 64)] for _ in range(trials)))


@pytest.mark.parametrize("slip", ["shift", "bank_mask", "row_mask", "dropped_step"])
def test_a_slip_in_the_sequence_is_caught(slip):
    # Planted mistakes in the modelled sequence must change some sums: the model tells the
    # kernel's sequence from a plausible wrong one.
    steps = [list(s) for s in kernel_dpp_steps()]
    if slip == "shift":
        steps[2][1] = 0x112  # row_shr:2 twice instead of row_shr:3
    elif slip == "bank_mask":
        steps[4][3] = 0x3  # row_shr:8 written into lanes 0-7 instead of 8-15
    elif slip == "row_mask":
        steps[5][2] = 0x5  # row_bcast:15 into rows 0 and 2 instead of 1 and 3
    else:
        steps = steps[:-1]  # no row_bcast:31: lane 63 holds only rows 2 and 3
    assert sums_differ([tuple(s) for s in steps]) > 0


def test_masks_that_only_guard_out_of_row_lanes_are_immaterial():
    # With old = 0, lanes whose source lies outside the row add 0 whether masked or not, so the
    # bank mask of row_shr:4, the row mask of row_bcast:15 and every bound_ctrl could be
    # dropped without changing lane 63.
    base = [list(s) for s in kernel_dpp_steps()]
    for change in ((3, 3, 0xF), (5, 2, 0xF)):
        steps = [list(s) for s in base]
        steps[change[0]][change[1]] = change[2]
        assert sums_differ([tuple(s) for s in steps]) == 0
    assert sums_differ([tuple(s[:4] + [False]) for s in base]) == 0


def test_fixed_point_sum_is_exact_to_full_weights():
    assert wave_sum_fixed([1 << 32] * 64) == 64 << 32
    assert wave_sum_fixed([(1 << 32) - 1] * 64) == 64 * ((1 << 32) - 1)
    rng = np.random.default_rng(2)
    for _ in range(100):
        q = [int(x) for x in rng.integers(0, (1 << 32) + 1, 64, dtype=np.uint64)]
        assert wave_sum_fixed(q) == sum(q)


def test_one_add_per_peeled_group_carrying_its_whole_weight():
    rng = np.random.default_rng(11)
    for _ in range(300):
        check_aggregation(*skewed_wave(rng))


def test_edge_cases():
    q = [1 << 30] * 64
    adds = check_aggregation([True] * 64, [7] * 64, q)  # every lane in one bin
    assert len(adds) == 1 and adds[0] == (7, 0, 64 << 30)
    adds = check_aggregation([True] * 64, list(range(64)), q, bins=64)  # every lane its own bin
    assert len(adds) == 64
    score = [False] * 64
    score[19] = True
    adds = check_aggregation(score, [3] * 64, q)  # one scoring lane; the others share its bin
    assert adds == [(3, 19, 1 << 30)]
    score = [bool((0x0000F0F0F0F00000 >> l) & 1) for l in range(64)]
    adds = check_aggregation(score, [5] * 64, [1 << 32] * 64)
    assert len(adds) == 1 and adds[0][2] == sum(score) << 32


def test_tile_counters_with_carries_are_exact():
    # aggregate_tile_add: the group's sum (< 2^38) goes into a 32-bit counter; its high word and
    # the counter's wrap go to the 64-bit bin, so counter + bin is the exact total.
    rng = np.random.default_rng(4)
    for _ in range(100):
        score, keys, q = skewed_wave(rng, bins=8)
        counters = [int(x) for x in rng.integers(0, 1 << 32, 8, dtype=np.uint64)]
        start = list(counters)
        bins = [0] * 8
        for b, _, v in aggregate_adds(score, keys, q):
            lo = v & M32
            old = counters[b]
            counters[b] = (old + lo) & M32
            bins[b] += ((v >> 32) + (1 if (old + lo) & M32 < old else 0)) << 32
        for b in range(8):
            total = sum(q[l] for l in range(64) if score[l] and keys[l] == b)
            assert counters[b] + bins[b] - start[b] == total
