"""The surrogate network's BF16 MFMA kernels against the float64 oracle (oracle/mlp_oracle.py).

The GPU rounds weights, inputs and hidden activations to bfloat16 and accumulates in float32; the
oracle emulates the bfloat16 rounding and accumulates in float64. Tolerances are written per
check below."""

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

CFG = dict(heads=2, inputs=30, hidden=128, depth=3, outputs=202, batch=256, seed=7)
DIMS = (30, 128, 3, 202)


@pytest.fixture()
def model(gpu):
    from tissuemc_amd.mlp import Mlp

    m = Mlp(**CFG)
    yield m
    m.close()


def data(n=512, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 30)).astype(np.float32)
    y = rng.normal(0, 1, (n, 202)).astype(np.float32)
    return x, y


def test_param_count_and_flops(model):
    assert model.param_count == M.param_count(*DIMS)
    per_step, per_query = model.flops()
    # algorithmic: the real widths (30 inputs, 202 outputs), not the padded ones
    q = 2 * (30 * 128 + 128 * 128 * 2 + 202 * 128) * 2
    assert per_query == q
    # a step: forward and weight gradient of every layer, input gradient of all but the first
    per_head = 2 * CFG["batch"] * (30 * 128 * 2 + 128 * 128 * 2 * 3 + 202 * 128 * 3)
    assert per_step == per_head * 2


def test_forward_matches_oracle(model):
    x, _ = data(300)
    mean, spread, ph, _ = model.predict(x, per_head=True)
    params = model.get_params()
    for h in range(2):
        ref_bf, _ = M.forward(params[h].astype(np.float64), x.astype(np.float64), DIMS, bf16=True)
        ref64, _ = M.forward(params[h].astype(np.float64), x.astype(np.float64), DIMS, bf16=False)
        scale = np.abs(ref64).max()
        # bf16 emulation: only float32 accumulation order and bf16 ties differ
        assert np.max(np.abs(ph[h] - ref_bf)) <= 2e-2 * scale
        assert np.linalg.norm(ph[h] - ref_bf) <= 2e-3 * np.linalg.norm(ref_bf)
        # against exact float64: the bf16 operand rounding (2^-9 relative per operand)
        assert np.linalg.norm(ph[h] - ref64) <= 2e-2 * np.linalg.norm(ref64)
    assert np.allclose(mean, ph.mean(axis=0), rtol=1e-5, atol=1e-6)
    assert np.allclose(spread, ph[:, :, :2].std(axis=0, ddof=1), rtol=1e-3, atol=1e-6)


def test_gradients_match_oracle(model):
    x, y = data(512)
    w = np.linspace(0.5, 2.0, 202).astype(np.float32)
    model.set_data(x, y, w)
    rows = np.random.default_rng(5).integers(0, 512, 256)
    g, loss = model.gradients(rows)
    params = model.get_params()
    for h in range(2):
        ref_loss, ref = M.loss_and_grad(params[h].astype(np.float64), x[rows].astype(np.float64),
                                        y[rows].astype(np.float64), w.astype(np.float64), DIMS,
                                        bf16=True)
        assert loss[h] == pytest.approx(ref_loss, rel=2e-3)
        cos = float(g[h] @ ref / (np.linalg.norm(g[h]) * np.linalg.norm(ref)))
        assert cos > 0.9995
        assert np.linalg.norm(g[h] - ref) <= 3e-2 * np.linalg.norm(ref)
        # per layer, weights and biases separately
        k = 0
        for o, i in M.shapes(*DIMS):
            for n_el in (o * i, o):
                a, b = g[h][k:k + n_el], ref[k:k + n_el]
                assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-9
                k += n_el


def test_adam_steps_follow_closed_form(model):
    """Every row of the set is the same sample, so any sampled batch has the same gradient; the
    first Adam step is then p - lr * g / (|g| + eps), and the second follows the moment
    recursion with the gradient at the new point."""
    x, y = data(1)
    model.set_data(np.repeat(x, 64, 0), np.repeat(y, 64, 0))
    rows = np.zeros(256, np.int64)
    p0 = model.get_params().astype(np.float64)
    g0, _ = model.gradients(rows)
    model.train(1)
    p1 = model.get_params().astype(np.float64)
    c = model.config
    exp1, m, v = M.adam_step(p0, g0.astype(np.float64), 0.0, 0.0, 1, c.lr, c.beta1, c.beta2, c.eps)
    assert np.allclose(p1, exp1, rtol=0, atol=2e-7 + 1e-6 * np.abs(p0).max())
    g1, _ = model.gradients(rows)
    model.train(1)
    p2 = model.get_params().astype(np.float64)
    exp2, _, _ = M.adam_step(p1, g1.astype(np.float64), m, v, 2, c.lr, c.beta1, c.beta2, c.eps)
    assert np.allclose(p2, exp2, rtol=0, atol=2e-7 + 1e-6 * np.abs(p0).max())


def test_training_learns_a_smooth_map_and_is_deterministic(gpu):
    from tissuemc_amd.mlp import Mlp

    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (8192, 30)).astype(np.float32)
    a = rng.normal(0, 1, (30, 202)).astype(np.float32) / 6
    y = np.sin(x @ a).astype(np.float32)
    runs = []
    for _ in range(2):
        with Mlp(**dict(CFG, batch=256, seed=11)) as m:
            m.set_data(x, y)
            losses, _ = m.train(600)
            runs.append((losses, m.get_params()))
    losses = runs[0][0]
    assert losses[-50:].mean() < 0.1 * losses[:10].mean()
    assert np.array_equal(runs[0][1], runs[1][1])  # same seed and data: bit-identical training


def test_set_params_round_trip(model):
    p = np.random.default_rng(1).normal(0, 0.1, (2, model.param_count)).astype(np.float32)
    model.set_params(p)
    assert np.array_equal(model.get_params(), p)


def test_gradients_match_oracle_at_full_size(gpu):
    """The training configuration itself (4 heads, 4 x 512 hidden, batch 4096): the weight
    gradients then run split four ways along the batch, with the tile order dealt over the XCDs,
    which the small network above does not reach. Heads 0 and 3 against the float64 oracle, to
    the tolerances of the small case."""
    from tissuemc_amd.mlp import Mlp

    dims = (30, 512, 4, 202)
    x, y = data(8192, seed=11)
    w = np.linspace(0.5, 2.0, 202).astype(np.float32)
    rows = np.random.default_rng(12).integers(0, 8192, 4096)
    with Mlp(heads=4, inputs=30, hidden=512, depth=4, outputs=202, batch=4096, seed=5) as m:
        m.set_data(x, y, w)
        g, loss = m.gradients(rows)
        params = m.get_params()
    for h in (0, 3):
        ref_loss, ref = M.loss_and_grad(params[h].astype(np.float64), x[rows].astype(np.float64),
                                        y[rows].astype(np.float64), w.astype(np.float64), dims,
                                        bf16=True)
        assert loss[h] == pytest.approx(ref_loss, rel=2e-3)
        cos = float(g[h] @ ref / (np.linalg.norm(g[h]) * np.linalg.norm(ref)))
        assert cos > 0.9995
        assert np.linalg.norm(g[h] - ref) <= 3e-2 * np.linalg.norm(ref)
        k = 0
        for o, i in M.shapes(*dims):
            for n_el in (o * i, o):
                a, b = g[h][k:k + n_el], ref[k:k + n_el]
                assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-9
                k += n_el


def test_forward_matches_oracle_at_full_width(gpu):
    """Inference at the training configuration's width (4 x 512 hidden), where the hidden layers
    run on 128 x 256 tiles (gemm_infer_wide), against the float64 oracle for heads 0 and 3, to
    the small network's tolerances; and the ensemble mean and spread against the per-head
    outputs."""
    from tissuemc_amd.mlp import Mlp

    dims = (30, 512, 4, 202)
    x, _ = data(1000, seed=13)
    with Mlp(heads=4, inputs=30, hidden=512, depth=4, outputs=202, batch=4096, seed=6) as m:
        mean, spread, ph, _ = m.predict(x, per_head=True)
        mean2, spread2, _, _ = m.predict(x)
        params = m.get_params()
    for h in (0, 3):
        ref_bf, _ = M.forward(params[h].astype(np.float64), x.astype(np.float64), dims, bf16=True)
        ref64, _ = M.forward(params[h].astype(np.float64), x.astype(np.float64), dims, bf16=False)
        scale = np.abs(ref64).max()
        assert np.max(np.abs(ph[h] - ref_bf)) <= 2e-2 * scale
        assert np.linalg.norm(ph[h] - ref_bf) <= 2e-3 * np.linalg.norm(ref_bf)
        assert np.linalg.norm(ph[h] - ref64) <= 2e-2 * np.linalg.norm(ref64)
    assert np.allclose(mean, ph.mean(axis=0), rtol=1e-5, atol=1e-6)
    assert np.allclose(spread, ph[:, :, :2].std(axis=0, ddof=1), rtol=1e-3, atol=1e-6)
    assert np.array_equal(mean, mean2) and np.array_equal(spread, spread2)


@pytest.mark.parametrize("heads,depth,n", [(4, 4, 40_000), (3, 1, 1000), (1, 2, 129)])
def test_fused_inference_matches_the_layered_kernels(gpu, monkeypatch, heads, depth, n):
    """Hidden width 512 with 30 inputs and 202 outputs goes through the fused whole-network
    kernel; TMC_MLP_INFER=layered runs the layer-by-layer products. Same MFMA, same operand
    order, same bias seeding and GELU: the outputs agree bit for bit. Row counts cross a chunk
    (one workgroup of 128 rows per CU) and end mid-tile."""
    from tissuemc_amd.mlp import Mlp

    cfg = dict(heads=heads, inputs=30, hidden=512, depth=depth, outputs=202, batch=1024, seed=5)
    rng = np.random.default_rng(heads * 10 + depth)
    x = rng.normal(0, 1, (n, 30)).astype(np.float32)
    with Mlp(**cfg) as fused:
        # trained a few steps, so the weights are not the initial draw
        fused.set_data(rng.normal(0, 1, (1024, 30)).astype(np.float32),
                       rng.normal(0, 1, (1024, 202)).astype(np.float32))
        fused.train(3)
        params = fused.get_params()
        mean_f, spread_f, ph_f, _ = fused.predict(x, per_head=True)
    monkeypatch.setenv("TMC_MLP_INFER", "layered")
    with Mlp(**cfg) as layered:
        layered.set_params(params)
        mean_l, spread_l, ph_l, _ = layered.predict(x, per_head=True)
    assert np.array_equal(ph_f, ph_l)
    assert np.array_equal(mean_f, mean_l) and np.array_equal(spread_f, spread_l)
    assert np.all(np.isfinite(ph_f)) and np.abs(ph_f).max() > 0
