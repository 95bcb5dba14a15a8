"""Surrogate answers end to end on the device (tmc_mlp_answer) against the host path.

The host path encodes features and the domain flag with numpy (tissuemc_amd.surrogate.features,
in_domain), runs the ensemble (tmc_mlp_predict), then de-standardizes and gates in numpy
(Surrogate.predict, gate). The device path does all of it on the GPU from the stack array. The
domain flags must agree exactly. The network sees bf16 inputs: where both encoders round a
feature to the same bf16 (all but a rare row, when the two log10 implementations differ in the
last bit at a bf16 rounding boundary) the answers agree to double rounding and the routing
decisions are identical.
"""

import numpy as np
import pytest

import tissuemc_amd.engine as engine
import tissuemc_amd.surrogate as S

pytestmark = pytest.mark.gpu


def make_model(hidden: int):
    m = S.Surrogate(heads=4, hidden=hidden, depth=2, batch=512, seed=5)
    rng = np.random.default_rng(7)
    m.y_mean = rng.uniform(-2.0, -0.5, S.N_OUT).astype(np.float32)
    m.y_std = rng.uniform(0.2, 0.6, S.N_OUT).astype(np.float32)
    return m


@pytest.fixture(scope="module")
def model(gpu):
    m = make_model(128)  # the layer-by-layer inference kernels, chunks of 16384 stacks
    yield m
    m.close()


@pytest.fixture(scope="module")
def fused_model(gpu):
    m = make_model(512)  # the fused inference kernel, chunks of 128 x the CU count
    yield m
    m.close()


def queries(n_in: int, n_ood: int):
    a = engine.generate_stacks(11, 0, n_in)
    b = engine.generate_stacks(12, 0, n_ood, engine.default_stack_dist(mus_hi=1000.0, d_hi=1.0,
                                                                       max_layers=7))
    out = (engine.TmcStack * (n_in + n_ood))()
    for i in range(n_in):
        out[i] = a[i]
    for i in range(n_ood):
        out[n_in + i] = b[i]
    return out


@pytest.mark.parametrize("which", ["model", "fused_model"])
def test_device_answers_match_the_host_path(request, which):
    model = request.getfixturevalue(which)
    # three chunks of 16384 rows for the layered model, the last one partial; several chunks
    # and packing threads in both cases (above half a chunk)
    n_in, n_ood = (36_000, 4_000) if which == "model" else (70_000, 6_000)
    qs = queries(n_in, n_ood)
    dom = S.in_domain(qs)
    p = model.predict(S.features(qs))
    rel_rd, rel_tt = S.relative_spreads(p)
    tau = (float(np.quantile(rel_rd[dom], 0.8)), float(np.quantile(rel_tt[dom], 0.7)))
    route = S.gate(p, dom, tau)
    assert 0.2 < np.mean(route[dom]) < 0.6
    d = model.answer(qs, tau, profiles=True)
    assert np.array_equal(d["in_domain"], dom)
    assert 0 < (~dom).sum() <= n_ood  # the wider box's stacks are partly inside the domain
    same = np.ones(len(qs), bool)
    for key in ("rd", "tt", "rd_spread", "tt_spread"):
        close = np.isclose(d[key], p[key], rtol=1e-12, atol=0)
        same &= close
    assert np.mean(same) >= 0.999, np.mean(same)
    assert np.array_equal(d["route"][same], route[same])
    assert np.all(d["route"][~dom])
    prof = np.concatenate([p["az"], p["rr"]], axis=1)
    assert np.allclose(d["profiles"][same], prof[same], rtol=2e-7, atol=0)
    assert d["wall_seconds"] >= d["device_seconds"] > 0


def test_device_answers_edge_sizes(model):
    qs = queries(1, 0)
    d = model.answer(qs, 0.1)
    p = model.predict(S.features(qs))
    assert d["rd"].shape == (1,) and np.isclose(d["rd"][0], p["rd"][0], rtol=1e-12)
    e = model.answer((engine.TmcStack * 0)(), 0.1)
    assert e["rd"].shape == (0,)
    # a chunk boundary exactly
    qs = queries(16384, 0)
    d = model.answer(qs, 0.1)
    p = model.predict(S.features(qs))
    assert np.mean(np.isclose(d["tt"], p["tt"], rtol=1e-12)) >= 0.999


@pytest.mark.parametrize("n", [1, 127, 32768, 32769])
def test_fused_answers_edge_sizes(fused_model, n):
    """One query, less than a workgroup's 128 rows, a chunk of 128 x 256 CUs exactly and one
    more: the answers of the fused path equal the host path's (which runs the same kernel)."""
    qs = queries(n, 0)
    d = fused_model.answer(qs, 0.1)
    p = fused_model.predict(S.features(qs))
    assert d["rd"].shape == (n,)
    assert np.mean(np.isclose(d["rd"], p["rd"], rtol=1e-12)) >= 0.999
    assert np.mean(np.isclose(d["tt"], p["tt"], rtol=1e-12)) >= 0.999


def test_answer_rejects_bad_inputs(model, gpu):
    with pytest.raises(gpu.ConfigError, match="lo < hi"):
        model.net.answer(queries(4, 0), y_mean=model.y_mean, y_std=model.y_std, tau=0.1,
                         box=engine.default_stack_dist(mua_lo=0.0))
