"""The oracle's building blocks against known answers and closed forms.

Mirrors the reference's tests/cpp/test_philox.cpp (Random123 known-answer vectors :24-30) and
tests/cpp/test_physics.cpp (Fresnel :64-89, HG moments :91-112, spin :114-159, steps :161-173,
absorbing slab :175-214), so that a wrong constant, a transposed operand or a dropped term
fails here independently of the golden data.
"""

import ctypes as C
import math

import numpy as np
import pytest

import oracle


def test_philox_known_answer_vectors():
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E,
                                                                 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                       0x24126EA1]


def test_stream_layout_is_counter_id_block_and_key_seed():
    seed, pid = 0xDEADBEEFCAFEF00D, 0x1_0000_0005
    words = oracle.stream_u32(seed, pid, 12)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for blk in range(3):
        ctr = [pid & 0xFFFFFFFF, pid >> 32, blk, 0]
        assert list(words[4 * blk:4 * blk + 4]) == oracle.philox(ctr, key)
    # A stream word in counter word 3 gives an unrelated stream.
    other = oracle.stream_u32(seed, pid, 12, word3=7)
    assert not np.any(other == words)


def test_unit_interval_conversions_are_open_and_exact():
    L = oracle.lib()
    for x in (0, 1, 511, 512, 0x7FFFFFFF, 0xFFFFFFFE, 0xFFFFFFFF):
        d = L.orc_u32_to_double(x)
        f = L.orc_u32_to_float(x)
        assert 0.0 < d < 1.0 and 0.0 < f < 1.0
        assert d == (x + 0.5) * 2.0 ** -32
        assert f == ((x >> 9) + 0.5) * 2.0 ** -23
        assert abs(d - f) < 2.5e-7


def test_streams_are_distinct_and_uniform():
    a = oracle.stream_u32(1234, 99, 64)
    assert np.array_equal(a, oracle.stream_u32(1234, 99, 64))
    assert not set(a) & set(oracle.stream_u32(1234, 100, 64))
    assert not set(a) & set(oracle.stream_u32(1235, 99, 64))
    u = (oracle.stream_u32(2026, 0, 1_000_000).astype(np.float64) + 0.5) * 2.0 ** -32
    assert abs(u.mean() - 0.5) < 1.5e-3
    assert abs(u.var() - 1 / 12) < 5e-4


def _textbook_fresnel(n1, n2, ci):
    st = n1 / n2 * math.sqrt(1 - ci * ci)
    if st >= 1:
        return 1.0
    ct = math.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.37, 1.0), (1.33, 1.37)])
def test_fresnel_equals_textbook_and_snell(n1, n2):
    L = oracle.lib()
    for ci in np.arange(0.05, 0.999, 0.05):
        ca2 = C.c_double(0)
        r = L.orc_fresnel(n1, n2, float(ci), C.byref(ca2), 64)
        assert r == pytest.approx(_textbook_fresnel(n1, n2, ci), rel=1e-10)
        if r < 1:
            assert n1 * math.sqrt(1 - ci * ci) == pytest.approx(n2 * math.sqrt(1 - ca2.value ** 2),
                                                                rel=1e-10)


def test_fresnel_edge_cases():
    L = oracle.lib()
    ca2 = C.c_double(0)
    assert L.orc_fresnel(1.37, 1.37, 0.3, C.byref(ca2), 64) == 0.0 and ca2.value == 0.3
    assert L.orc_fresnel(1.0, 1.5, 1.0, C.byref(ca2), 64) == pytest.approx(0.04)
    assert L.orc_fresnel(1.5, 1.0, 0.70, C.byref(ca2), 64) == 1.0 and ca2.value == 0.0
    assert L.orc_critical_cos(1.5, 1.0) == pytest.approx(math.sqrt(1 - 1 / 2.25))
    assert L.orc_critical_cos(1.0, 1.5) == 0.0


@pytest.mark.parametrize("g", [0.0, 0.5, 0.9, -0.3])
@pytest.mark.parametrize("precision", [64, 32])
def test_hg_sampler_moments(g, precision):
    n = 1_000_000
    out = np.zeros(n)
    oracle.lib().orc_hg_samples(g, 7, int(g * 1000 + 1000), n,
                                out.ctypes.data_as(C.POINTER(C.c_double)), precision)
    assert out.min() >= -1 and out.max() <= 1
    assert abs(out.mean() - g) < 3e-3                     # E[cos] = g
    assert abs((out ** 2).mean() - (1 + 2 * g * g) / 3) < 3e-3  # E[cos^2] = (1 + 2g^2)/3


def _walk(g, n, u0, restart, precision):
    worst, mean = C.c_double(0), C.c_double(0)
    u = (C.c_double * 3)(*u0)
    oracle.lib().orc_spin_walk(g, 11 if restart else 5, 0 if restart else 1, n, u, restart,
                               precision, C.byref(worst), C.byref(mean))
    return worst.value, mean.value


def test_spin_keeps_unit_direction_and_mean_cosine():
    worst, mean = _walk(0.75, 400_000, (0, 0, 1), 1000, 64)
    assert worst < 1e-9
    assert abs(mean - 0.75) < 5e-3


@pytest.mark.parametrize("g", [0.0, 0.99])
def test_direction_norm_does_not_drift_over_long_walks(g):
    assert _walk(g, 4_000_000, (0.6, 0.0, 0.8), 0, 64)[0] < 1e-13
    assert _walk(g, 4_000_000, (0.6, 0.0, 0.8), 0, 32)[0] < 5e-7


def test_step_lengths_are_exponential_with_mean_one_over_mut():
    spec = {"photons": 1, "dz": 0.01, "dr": 0.01, "nz": 1, "nr": 1, "na": 1, "n_above": 1.0,
            "n_below": 1.0, "layers": [{"n": 1.0, "mua": 10, "mus": 90, "g": 0.75, "d": 1e6}]}
    n = 1_000_000
    out = np.zeros(n)
    oracle.lib().orc_step_lengths(C.byref(oracle.to_spec(spec)), 3, 0, n,
                                  out.ctypes.data_as(C.POINTER(C.c_double)))
    assert abs(out.mean() - 0.01) < 5e-5
    assert abs(np.median(out) - 0.01 * math.log(2)) < 1e-4


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_purely_absorbing_slab_matches_multiple_reflection_series(precision):
    """mus = 0: a packet goes straight down and back, absorbed entirely at its first step, so
    reflectance and transmittance are binomial with closed-form means (Rsp = Rt = Rb = 0.04)."""
    n, mua, d = 1.5, 5.0, 0.1
    tau = mua * d
    spec = {"photons": 1_000_000, "dz": 0.01, "dr": 0.01, "nz": 10, "nr": 10, "na": 1,
            "n_above": 1.0, "n_below": 1.0, "layers": [{"n": n, "mua": mua, "mus": 0.0, "g": 0.0,
                                                        "d": d}]}
    rsp = rt = rb = 0.04
    loop = 1.0 / (1.0 - rt * rb * math.exp(-2 * tau))
    t_exact = (1 - rsp) * math.exp(-tau) * (1 - rb) * loop
    rd_exact = (1 - rsp) * math.exp(-2 * tau) * rb * (1 - rt) * loop
    r = oracle.run(spec, batches=10, seed=99, precision=precision)
    N = 1_000_000
    rd = sum(b["rd"] for b in r["batches"]) / N
    tt = sum(b["tt"] for b in r["batches"]) / N
    ab = sum(sum(b["absorbed"]) for b in r["batches"]) / N

    def se(x):
        p = x / (1 - rsp)
        return (1 - rsp) * math.sqrt(p * (1 - p) / N)

    assert abs(tt - t_exact) < 5 * se(t_exact)
    assert abs(rd - rd_exact) < 5 * se(rd_exact)
    tol = 1e-13 if precision == "float64" else 1e-5
    assert abs(rsp + rd + tt + ab - 1.0) < tol


def test_batch_plan_partitions_ids():
    L = oracle.lib()
    photons, batches = 1000, 7
    counts = [0] * batches
    for i in range(photons):
        counts[L.orc_batch_of(photons, batches, i)] += 1
    assert counts == [143] * 6 + [142]
    r = oracle.run({"photons": 1000, "dz": 0.01, "dr": 0.01, "nz": 2, "nr": 2, "na": 1,
                    "n_above": 1.0, "n_below": 1.0,
                    "layers": [{"n": 1.5, "mua": 5.0, "mus": 0.0, "g": 0.0, "d": 0.1}]},
                   batches=7, seed=1)
    firsts = [b["first_photon"] for b in r["batches"]]
    assert firsts == [0, 143, 286, 429, 572, 715, 858]


def test_oracle_rejects_bad_options():
    spec = {"photons": 10, "dz": 0.01, "dr": 0.01, "nz": 2, "nr": 2, "na": 1, "n_above": 1.0,
            "n_below": 1.0, "layers": [{"n": 1.5, "mua": 5.0, "mus": 0.0, "g": 0.0, "d": 0.1}]}
    for kw in ({"batches": 0}, {"batches": 11}, {"threads": -1}, {"threads": 5000}):
        with pytest.raises(ValueError):
            oracle.run(spec, **{"batches": 2, **kw})
    bad = dict(spec, layers=[dict(spec["layers"][0], g=1.0)])
    with pytest.raises(ValueError, match="g must be"):
        oracle.run(bad, batches=2)
