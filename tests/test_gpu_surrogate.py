"""The surrogate pipeline end to end at a small scale (tissuemc_amd.surrogate).

- Training targets are the batched engine's tallies; the trained ensemble is judged against Monte
  Carlo at 10x the training photons.
- What the network gains over the method without it: its R_d/T_t errors against those of the
  best predictor that uses no network (the training-set mean), on the same judged stacks.
- Out-of-range queries are always routed back to Monte Carlo, and routed answers are Monte Carlo.
"""

import argparse

import numpy as np
import pytest

import tissuemc_amd.engine as engine
import tissuemc_amd.surrogate as S

pytestmark = pytest.mark.gpu


def test_surrogate_pipeline_small(gpu, tmp_path):
    stacks = engine.generate_stacks(5, 0, 8192)
    mc = S.simulate(stacks, 20_000, seed=3)
    X, Y = S.features(stacks), S.targets(mc)
    assert np.all(np.isfinite(Y)) and X.shape == (8192, S.N_IN) and Y.shape == (8192, S.N_OUT)
    # targets are the engine's tallies: R_d + T_t + absorbed + specular = 1 within noise
    m = S.Surrogate(heads=4, hidden=128, depth=3, batch=512, seed=2)
    try:
        fit = m.fit(X[512:], Y[512:], 3000, lr=2e-3, lr_final=5e-5)
        assert fit["loss_last"] < 0.2 * fit["loss_first"]
        held = [stacks[i] for i in range(512)]
        judge = S.simulate(held, 200_000, seed=4)
        pred = m.predict(X[:512])
        err_net = np.maximum(S.rel_err(pred["rd"], judge.rd, 1e-3), S.rel_err(pred["tt"], judge.tt, 1e-2))
        base_rd = np.full(512, mc.rd[512:].mean())
        base_tt = np.full(512, mc.tt[512:].mean())
        err_base = np.maximum(S.rel_err(base_rd, judge.rd, 1e-3), S.rel_err(base_tt, judge.tt, 1e-2))
        assert np.median(err_net) < 0.25 * np.median(err_base), (np.median(err_net), np.median(err_base))
        # gate: out-of-range stacks always go back to Monte Carlo
        ood = engine.generate_stacks(9, 0, 400, engine.default_stack_dist(mus_hi=1000.0, d_hi=1.0))
        dom = S.in_domain(ood)
        assert 0 < (~dom).sum() < 400
        pq = m.predict(S.features(ood))
        route = S.gate(pq, dom, tau=1e9)
        assert np.array_equal(route, ~dom)
        # routed answers are Monte Carlo: the same stacks simulated twice agree exactly
        idx = np.nonzero(route)[0][:16]
        sub = [ood[i] for i in idx]
        a = S.simulate(sub, 10_000, seed=8)
        b = S.simulate(sub, 10_000, seed=8)
        assert np.array_equal(a.rd, b.rd) and np.array_equal(a.az, b.az)
        # save / load round trip
        m.save(tmp_path / "m.npz")
        m2 = S.Surrogate(heads=4, hidden=128, depth=3, batch=512, seed=99)
        try:
            m2.load(tmp_path / "m.npz")
            p2 = m2.predict(X[:512])
            assert np.array_equal(p2["rd"], pred["rd"])
        finally:
            m2.close()
    finally:
        m.close()


def test_pipeline_command_small(gpu, tmp_path):
    args = argparse.Namespace(out=str(tmp_path), stacks=4096, photons=10_000, steps=400, scalar_weight=500.0, batch=512,
                              heads=4, hidden=128, depth=2, lr=2e-3, seed=3, val=512, val_photons=40_000, tau=0.0,
                              target_err=0.05, max_route=0.1, queries=2000, ood_frac=0.05, fallback_photons=20_000,
                              judge=200, judge_photons=100_000)
    rep = S.pipeline(args)
    assert rep["inference"]["out_of_domain"] <= rep["inference"]["routed"]
    assert rep["judge"]["stacks"] == 200
    assert rep["matched"]["mc_stacks_per_s_matched"] > 0
    assert (tmp_path / "report.json").exists()
