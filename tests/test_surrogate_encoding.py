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
