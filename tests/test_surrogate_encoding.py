"""The surrogate's input encoding and domain gate on the host (tissuemc_amd.surrogate), on the
CPU: the vectorised encoder over the stack array against a plain per-stack loop written from the
encoding's definition, to float32 rounding (5 slots x (log10 mua, log10 mus, g, n, log10 d) mapped to [-1, 1] over the
training box, empty slots 0, then the slot masks; in the domain when the layer count is in range
and every present layer's encoded parameters lie in the box to 1e-9)."""

import math

import numpy as np

import tissuemc_amd.engine as engine
import tissuemc_amd.surrogate as S


def loop_encoding(stacks, dist):
    los = [math.log10(dist.mua_lo), math.log10(dist.mus_lo), dist.g_lo, dist.n_lo, math.log10(dist.d_lo)]
    his = [math.log10(dist.mua_hi), math.log10(dist.mus_hi), dist.g_hi, dist.n_hi, math.log10(dist.d_hi)]
    X, dom = [], []
    for s in stacks:
        row, mask = [], []
        ok = dist.min_layers <= s.num_layers <= dist.max_layers
        for i in range(5):
            if i < s.num_layers:
                L = s.layers[i]
                v = [math.log10(L.mua) if L.mua > 0 else -99.0, math.log10(L.mus) if L.mus > 0 else -99.0,
                     L.g, L.n, math.log10(L.d)]
                row += [(v[j] - (los[j] + his[j]) / 2) / ((his[j] - los[j]) / 2) for j in range(5)]
                ok = ok and all(los[j] - 1e-9 <= v[j] <= his[j] + 1e-9 for j in range(5))
                mask.append(1.0)
            else:
                row += [0.0] * 5
                mask.append(0.0)
        X.append(row + mask)
        dom.append(ok)
    return np.array(X, np.float32), np.array(dom)


def test_vectorised_encoding_matches_the_definition():
    dist = engine.default_stack_dist()
    ins = engine.generate_stacks(3, 0, 300)
    ood = engine.generate_stacks(4, 0, 300, engine.default_stack_dist(mus_hi=1000.0, d_hi=1.0,
                                                                      max_layers=7, mua_lo=1e-3))
    stacks = [ins[i] for i in range(300)] + [ood[i] for i in range(300)]
    # a zero absorption coefficient and an eight-layer stack
    z = engine.stacks_array(stacks[:1])[0]
    z.layers[0].mua = 0.0
    stacks.append(z)
    X, dom = S.features(stacks, dist), S.in_domain(stacks, dist)
    Xr, dr = loop_encoding(stacks, dist)
    # numpy's and the C library's log10 may differ in the last bit of a double: the float32
    # features agree to a float32 rounding step
    assert np.allclose(X, Xr, rtol=0, atol=1e-7)
    assert np.array_equal(dom, dr)
    assert dom[:300].all() and 0 < (~dom[300:]).sum() < 300 and not dom[-1]


def test_gate_and_threshold_choice():
    """Per-output gate: a stack is routed when it is out of domain or either relative spread
    exceeds its threshold; the chosen pair meets the target at the 90th percentile of both
    outputs' held-out errors and accepts at least as many stacks as any single threshold that
    does."""
    rng = np.random.default_rng(1)
    n = 4000
    rd = rng.uniform(0.01, 0.5, n)
    tt = rng.uniform(0.0, 0.8, n)
    s_rd = rng.uniform(0.0, 0.02, n) * rd           # relative spreads up to 2% ...
    s_tt = rng.uniform(0.0, 0.05, n) * np.maximum(tt, 1e-2)  # ... and 5%
    pred = {"rd": rd, "tt": tt, "rd_spread": s_rd, "tt_spread": s_tt}
    dom = rng.random(n) > 0.05
    rel_rd, rel_tt = S.relative_spreads(pred)
    # errors that grow with the spreads: the gate can reach the target
    e_rd, e_tt = 0.2 * rel_rd, 0.1 * rel_tt
    route = S.gate(pred, dom, (0.01, 0.02))
    assert np.array_equal(route, ~dom | (rel_rd > 0.01) | (rel_tt > 0.02))
    a, b = S.choose_gate(pred, dom, e_rd, e_tt, target=0.003, max_route=0.1, min_accept=0.3)
    acc = ~S.gate(pred, dom, (a, b))
    assert np.quantile(e_rd[acc], 0.9) <= 0.003 and np.quantile(e_tt[acc], 0.9) <= 0.003
    assert acc.sum() >= 0.3 * n
    assert a < b  # T_t's spreads are wider: it needs the looser threshold
    # a single threshold for both outputs meeting the target accepts fewer stacks
    grid = np.geomspace(1e-4, 1.0, 61)
    single = []
    for t in grid:
        keep = ~S.gate(pred, dom, t)
        if keep.any() and np.quantile(e_rd[keep], 0.9) <= 0.003 and np.quantile(e_tt[keep], 0.9) <= 0.003:
            single.append(int(keep.sum()))
    assert max(single) < acc.sum()
    # unreachable target: one threshold routing at most max_route
    a, b = S.choose_gate(pred, dom, e_rd + 1.0, e_tt, target=0.004, max_route=0.2)
    assert a == b and np.mean(S.gate(pred, dom, a)) <= 0.2


def test_values_invert_targets():
    """values() undoes targets() for every output (R_d, T_t with their own floors, the profile
    bins with theirs), including exact zeros, to float32 rounding of the stored targets."""
    rng = np.random.default_rng(5)
    n = 400
    rd = np.concatenate([rng.uniform(0.0, 0.6, n - 4), [0.0, 1e-6, 0.5, 1.0]])
    tt = np.concatenate([10.0 ** rng.uniform(-6, 0, n - 4), [0.0, 1e-4, 0.01, 0.9]])
    az = rng.uniform(0.0, 0.05, (n, S.NZ)) * (rng.uniform(size=(n, S.NZ)) > 0.2)
    rr = rng.uniform(0.0, 0.05, (n, S.NR)) * (rng.uniform(size=(n, S.NR)) > 0.2)
    mc = S.McResult(rd, tt, np.zeros(n), np.zeros(n), az, rr, np.full(n, 1e5), 0.0, 0.0)
    v = S.values(S.targets(mc))
    # log10 stored as float32: about 2.7e-7 relative in value + floor
    for key, ref, floor in (("rd", rd, S.FLOOR_RD), ("tt", tt, S.FLOOR_TT), ("az", az, S.FLOOR),
                            ("rr", rr, S.FLOOR)):
        assert np.all(np.abs(v[key] - ref) <= 1e-6 * (ref + floor)), key
    # each output has its own floor: R_d's and T_t's are the larger ones
    assert S.FLOOR < S.FLOOR_RD < S.FLOOR_TT
    y = S.targets(mc)
    assert np.allclose(y[:, 0], np.log10(rd + S.FLOOR_RD), atol=1e-6)
    assert np.allclose(y[:, 1], np.log10(tt + S.FLOOR_TT), atol=1e-6)
