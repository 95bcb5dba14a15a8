"""The stack generator of the batched workload (tmc_generate_stacks), on the CPU.

Its definition (include/tissuemc_hip.h) is checked against the oracle's independent Philox: each
stack field is a fixed word of a fixed Philox block keyed by (seed, stack id)."""

import math

import numpy as np

import oracle
import tissuemc_amd.engine as engine

TAG = 0x53544B31


def unit(x: int) -> float:
    return (x + 0.5) * 2.0 ** -32


def test_generator_follows_its_philox_definition():
    seed, first, n = 0xC0FFEE, 1_000_000_007, 40
    arr = engine.generate_stacks(seed, first, n)
    d = engine.default_stack_dist()
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for i in range(n):
        k = first + i
        ctr = lambda blk: [k & 0xFFFFFFFF, k >> 32, blk, TAG]  # noqa: E731
        h = oracle.philox(ctr(0), key)
        s = arr[i]
        assert s.num_layers == 1 + min(4, int(unit(h[0]) * 5))
        budget = math.exp(math.log(1e4) + unit(h[1]) * (math.log(1e6) - math.log(1e4)))
        assert s.photons == round(budget)
        assert s.n_above == s.n_below == 1.0
        for l in range(s.num_layers):
            p, q = oracle.philox(ctr(1 + l), key), oracle.philox(ctr(17 + l), key)
            L = s.layers[l]
            assert math.isclose(L.mua, math.exp(math.log(d.mua_lo) + unit(p[0]) * math.log(d.mua_hi / d.mua_lo)), rel_tol=1e-12)
            assert math.isclose(L.mus, math.exp(math.log(d.mus_lo) + unit(p[1]) * math.log(d.mus_hi / d.mus_lo)), rel_tol=1e-12)
            assert math.isclose(L.g, d.g_lo + unit(p[2]) * (d.g_hi - d.g_lo), rel_tol=1e-12)
            assert math.isclose(L.n, d.n_lo + unit(p[3]) * (d.n_hi - d.n_lo), rel_tol=1e-12)
            assert math.isclose(L.d, math.exp(math.log(d.d_lo) + unit(q[0]) * math.log(d.d_hi / d.d_lo)), rel_tol=1e-12)


def test_ranges_are_consistent_and_deterministic():
    a = engine.generate_stacks(5, 0, 200)
    b = engine.generate_stacks(5, 150, 50)
    for i in range(50):
        assert engine.stack_to_dict(a[150 + i]) == engine.stack_to_dict(b[i])
    c = engine.generate_stacks(6, 0, 200)
    assert any(engine.stack_to_dict(a[i]) != engine.stack_to_dict(c[i]) for i in range(200))


def test_distribution_moments():
    n = 20000
    arr = engine.generate_stacks(11, 0, n)
    layers = np.array([arr[i].num_layers for i in range(n)])
    assert set(np.unique(layers)) == {1, 2, 3, 4, 5}
    assert abs(layers.mean() - 3.0) < 0.05
    lp = np.log10([arr[i].photons for i in range(n)])
    assert abs(lp.mean() - 5.0) < 0.03 and lp.min() >= 4 - 1e-9 and lp.max() <= 6 + 1e-9
    mua = np.log10([arr[i].layers[0].mua for i in range(n)])
    assert abs(mua.mean() - (-0.5)) < 0.03
    g = np.array([arr[i].layers[0].g for i in range(n)])
    assert 0.6 <= g.min() and g.max() <= 0.95 and abs(g.mean() - 0.775) < 0.005
    for i in range(0, n, 97):
        engine.validate_spec({"photons": arr[i].photons, "dz": 0.01, "dr": 0.01, "nz": 1, "nr": 1,
                              "na": 1, "n_above": 1.0, "n_below": 1.0,
                              "layers": engine.stack_to_dict(arr[i])["layers"]})


def test_out_of_range_distribution():
    d = engine.default_stack_dist(mus_hi=1000.0, d_hi=1.0)
    arr = engine.generate_stacks(3, 0, 5000, d)
    mus = [arr[i].layers[0].mus for i in range(5000)]
    assert max(mus) > 500.0
