"""Learned surrogate for layered-tissue Monte Carlo, with Monte Carlo as fallback and judge.

No reference counterpart. A stack of up to 5 layers (mua, mus, g, n, thickness) maps to R_d,
T_t, the depth profile A_z (100 bins of 0.01 cm) and the radial reflectance R_r (100 bins of
0.01 cm), learned by an ensemble of MLPs (tissuemc_amd.mlp; BF16 MFMA kernels) from tallies of
the batched Monte Carlo engine. Queries outside the training box, or on which the ensemble
disagrees, are routed back to Monte Carlo; Monte Carlo at many more photons judges the error.

    python -m tissuemc_amd.surrogate pipeline --out DIR [--stacks 1048576 --photons 100000 ...]

Encoding (30 inputs): per layer slot i < 5 the five parameters log10(mua), log10(mus), g, n,
log10(d), each mapped to [-1, 1] over the training box; empty slots are 0; then a 5-entry layer
mask. Targets (202 outputs): log10(R_d + 3e-4), log10(T_t + 3e-3), and log10 of A_z (100 bins)
and R_r (100 bins) plus 1e-7, where A_z and R_r are the per-photon weight fractions in each bin;
outputs are standardized for training. R_d and T_t span orders of magnitude over the stack
distribution, so their errors are learned (and gated) in relative terms; their floors are a few
tenths of the floors their errors are judged against (R_d relative to max(R_d, 1e-3), T_t to
max(T_t, 1e-2)), so that stacks whose T_t is far below 1e-2, whose labels at 1e5 photons are
mostly Monte Carlo noise in log terms, do not dominate the fit.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import engine
from .mlp import Mlp

SLOTS = 5
N_IN = SLOTS * 5 + SLOTS
NZ = NR = 100
DZ = DR = 0.01
N_OUT = 2 + NZ + NR
FLOOR = 1e-7       # profiles: log10(value + FLOOR)
FLOOR_RD = 3e-4    # log10(R_d + FLOOR_RD)
FLOOR_TT = 3e-3    # log10(T_t + FLOOR_TT)
GRID = dict(dz=DZ, dr=DR, nz=NZ, nr=NR)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ encoding

def box(dist) -> np.ndarray:
    """[5 params x (lo, hi)] of the encoded parameters (log10 for mua, mus, d)."""
    return np.array([[math.log10(dist.mua_lo), math.log10(dist.mua_hi)],
                     [math.log10(dist.mus_lo), math.log10(dist.mus_hi)],
                     [dist.g_lo, dist.g_hi], [dist.n_lo, dist.n_hi],
                     [math.log10(dist.d_lo), math.log10(dist.d_hi)]])


_LAYER_DT = np.dtype([(k, "<f8") for k in ("n", "mua", "mus", "g", "d")])


def stack_view(stacks) -> np.ndarray:
    """A numpy structured view (no copy) of a ctypes array of engine.TmcStack."""
    arr = engine.stacks_array(stacks)
    T = engine.TmcStack
    dt = np.dtype({"names": ["num_layers", "n_above", "n_below", "layers", "photons"],
                   "formats": ["<i4", "<f8", "<f8", (_LAYER_DT, (engine.MAX_LAYERS,)), "<u8"],
                   "offsets": [T.num_layers.offset, T.n_above.offset, T.n_below.offset,
                               T.layers.offset, T.photons.offset],
                   "itemsize": C.sizeof(T)})
    return np.frombuffer(arr, dtype=dt, count=len(arr))


def stack_params(stacks) -> tuple[np.ndarray, np.ndarray]:
    """Raw parameters [n x 5 slots x 5] (log10 for mua, mus, d; 0 in empty slots) and layer counts."""
    from numpy.lib import recfunctions

    v = stack_view(stacks)
    nl = v["num_layers"].astype(np.int64)
    # (n, mua, mus, g, d) of the first SLOTS layers as one dense [n x SLOTS x 5] array
    L = recfunctions.structured_to_unstructured(v["layers"][:, :SLOTS])
    present = np.arange(SLOTS)[None, :] < nl[:, None]
    raw = np.zeros((len(nl), SLOTS, 5))
    with np.errstate(divide="ignore", invalid="ignore"):
        for j, (col, log) in enumerate(((1, True), (2, True), (3, False), (0, False), (4, True))):
            x = L[:, :, col]
            if log:
                absent = -99.0 if j < 2 else 0.0
                x = np.where(x > 0, np.log10(np.where(x > 0, x, 1.0)), absent)
            raw[:, :, j] = x
    raw[~present] = 0.0
    return raw, nl


def features(stacks, dist=None) -> np.ndarray:
    dist = dist or engine.default_stack_dist()
    raw, nl = stack_params(stacks)
    b = box(dist)
    center, half = b.mean(axis=1), (b[:, 1] - b[:, 0]) / 2
    mask = (np.arange(SLOTS)[None, :] < nl[:, None]).astype(np.float64)
    enc = (raw - center[None, None, :]) / half[None, None, :] * mask[:, :, None]
    return np.concatenate([enc.reshape(len(nl), SLOTS * 5), mask], axis=1).astype(np.float32)


def in_domain(stacks, dist=None) -> np.ndarray:
    """True where every layer lies in the training box and the layer count is in range."""
    dist = dist or engine.default_stack_dist()
    raw, nl = stack_params(stacks)
    b = box(dist)
    ok = (nl >= dist.min_layers) & (nl <= dist.max_layers)
    tol = 1e-9
    for i in range(SLOTS):
        present = nl > i
        inside = np.all((raw[:, i] >= b[:, 0] - tol) & (raw[:, i] <= b[:, 1] + tol), axis=1)
        ok &= ~present | inside
    return ok


# ------------------------------------------------------------------ Monte Carlo

@dataclass
class McResult:
    rd: np.ndarray      # [n] per-photon diffuse reflectance
    tt: np.ndarray      # [n] per-photon transmittance
    rd_se: np.ndarray   # batch-means standard errors
    tt_se: np.ndarray
    az: np.ndarray      # [n x NZ] per-photon absorbed fraction per depth bin
    rr: np.ndarray      # [n x NR] per-photon reflectance fraction per radial bin
    photons: np.ndarray
    kernel_seconds: float
    wall_seconds: float


def simulate(stacks, photons: int | None, *, seed: int, batches: int = 10, chunk: int = 8192,
             progress: str = "") -> McResult:
    """Batched Monte Carlo of every stack (with its own budget, or `photons` each)."""
    n = len(stacks)
    arr = (engine.TmcStack * n)()
    for k in range(n):
        arr[k] = stacks[k]
        if photons:
            arr[k].photons = photons
    out = {k: [] for k in ("rd", "tt", "rd_se", "tt_se", "az", "rr", "photons")}
    kern = 0.0
    t0 = time.perf_counter()
    plans: dict[int, engine.BatchedPlan] = {}
    try:
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            sub = (engine.TmcStack * m).from_buffer(arr, c0 * C_SIZE)
            if m not in plans:
                hint = int(sum(sub[i].photons for i in range(m)))
                plans[m] = engine.BatchedPlan(m, batches_per_stack=batches, photons_hint=hint, **GRID)
            r = plans[m].execute(sub, seed=seed + c0)
            ph = r["photons"].astype(np.float64)
            rows = np.array([[c for _, c in engine.batch_plan_rows(int(p), batches)] for p in ph], np.float64)
            for key, src in (("rd", r["rd"]), ("tt", r["tt"])):
                per = src.astype(np.float64) * engine.UNIT / rows
                out[key].append(src.astype(np.float64).sum(axis=1) * engine.UNIT / ph)
                out[key + "_se"].append(per.std(axis=1, ddof=1) / math.sqrt(batches))
            out["az"].append(r["az"].astype(np.float64) * engine.UNIT / ph[:, None])
            out["rr"].append(r["rr"].astype(np.float64) * engine.UNIT / ph[:, None])
            out["photons"].append(ph)
            kern += r["kernel_seconds"]
            if progress:
                log(f"{progress}: {c0 + m}/{n} stacks, {time.perf_counter() - t0:.1f} s")
    finally:
        for p in plans.values():
            p.close()
    cat = {k: np.concatenate(v) for k, v in out.items()}
    return McResult(cat["rd"], cat["tt"], cat["rd_se"], cat["tt_se"], cat["az"], cat["rr"],
                    cat["photons"], kern, time.perf_counter() - t0)


C_SIZE = 0  # set below (bytes of one TmcStack)


def targets(mc: McResult) -> np.ndarray:
    return np.concatenate([np.log10(mc.rd[:, None] + FLOOR_RD), np.log10(mc.tt[:, None] + FLOOR_TT),
                           np.log10(mc.az + FLOOR), np.log10(mc.rr + FLOOR)],
                          axis=1).astype(np.float32)


def values(y: np.ndarray) -> dict:
    """The inverse of targets(): R_d, T_t, A_z and R_r from (de-standardized) outputs y [n x
    202], each max(10^y - floor, 0)."""
    y = np.asarray(y, np.float64)
    return {"rd": np.maximum(10.0 ** y[:, 0] - FLOOR_RD, 0.0),
            "tt": np.maximum(10.0 ** y[:, 1] - FLOOR_TT, 0.0),
            "az": np.maximum(10.0 ** y[:, 2:2 + NZ] - FLOOR, 0.0),
            "rr": np.maximum(10.0 ** y[:, 2 + NZ:] - FLOOR, 0.0)}


# ------------------------------------------------------------------ the model

class Surrogate:
    def __init__(self, *, heads=4, hidden=512, depth=4, batch=4096, seed=1, lr=1e-3):
        self.net = Mlp(heads=heads, inputs=N_IN, hidden=hidden, depth=depth, outputs=N_OUT,
                       batch=batch, seed=seed, lr=lr)
        self.y_mean = np.zeros(N_OUT, np.float32)
        self.y_std = np.ones(N_OUT, np.float32)
        self.dist = engine.default_stack_dist()

    def fit(self, X: np.ndarray, Y: np.ndarray, steps: int, *, lr=1e-3, lr_final=1e-5,
            scalar_weight: float = 500.0, segments: int = 20) -> dict:
        self.y_mean = Y.mean(axis=0).astype(np.float32)
        self.y_std = np.maximum(Y.std(axis=0), 1e-6).astype(np.float32)
        w = np.ones(N_OUT, np.float32)
        w[:2] = scalar_weight
        self.net.set_data(X, (Y - self.y_mean) / self.y_std, w)
        losses, seconds = [], 0.0
        per = max(1, steps // segments)
        done = 0
        while done < steps:
            k = min(per, steps - done)
            frac = done / max(steps, 1)
            self.net.set_lr(lr * (lr_final / lr) ** frac)  # geometric decay
            l, s = self.net.train(k)
            losses.append(l)
            seconds += s
            done += k
            log(f"train: {done}/{steps} steps, loss {float(l[-1].mean()):.4g}, {seconds:.1f} s")
        losses = np.concatenate(losses)
        return {"steps": steps, "seconds": seconds, "loss_first": float(losses[:10].mean()),
                "loss_last": float(losses[-100:].mean())}

    def predict(self, X: np.ndarray) -> dict:
        mean, spread, _, sec = self.net.predict(X)
        out = values((mean * self.y_std + self.y_mean).astype(np.float64))
        # spreads come out in log10 units of value + floor: d value / dy = ln(10) (value + floor)
        out["rd_spread"] = (out["rd"] + FLOOR_RD) * math.log(10) * spread[:, 0].astype(np.float64) * self.y_std[0]
        out["tt_spread"] = (out["tt"] + FLOOR_TT) * math.log(10) * spread[:, 1].astype(np.float64) * self.y_std[1]
        out["seconds"] = sec
        return out

    def answer(self, stacks, tau, profiles: bool = False) -> dict:
        """Gated answers end to end on the device (encoding, domain gate, ensemble, spread gate:
        tmc_mlp_answer); the same decisions as predict() + gate() on features()/in_domain()."""
        return self.net.answer(stacks, y_mean=self.y_mean, y_std=self.y_std, tau=tau,
                               floor=FLOOR, floor_rd=FLOOR_RD, floor_tt=FLOOR_TT, box=self.dist,
                               profiles=profiles)

    def save(self, path: Path) -> None:
        np.savez(path, params=self.net.get_params(), y_mean=self.y_mean, y_std=self.y_std)

    def load(self, path: Path) -> None:
        z = np.load(path, allow_pickle=False)
        self.net.set_params(z["params"])
        self.y_mean, self.y_std = z["y_mean"], z["y_std"]

    def close(self) -> None:
        self.net.close()


def relative_spreads(pred: dict) -> tuple[np.ndarray, np.ndarray]:
    """Ensemble spread of R_d relative to max(R_d, 1e-3) and of T_t relative to max(T_t, 1e-2)."""
    return (pred["rd_spread"] / np.maximum(np.abs(pred["rd"]), 1e-3),
            pred["tt_spread"] / np.maximum(np.abs(pred["tt"]), 1e-2))


def gate(pred: dict, domain: np.ndarray, tau) -> np.ndarray:
    """True where the query goes back to Monte Carlo: out of the training box, or a relative
    ensemble spread above its threshold. `tau` is one threshold for both outputs or a pair
    (R_d's, T_t's)."""
    tau_rd, tau_tt = (tau, tau) if np.isscalar(tau) else tau
    rel_rd, rel_tt = relative_spreads(pred)
    return (~domain) | (rel_rd > tau_rd) | (rel_tt > tau_tt)


def choose_gate(pred: dict, domain: np.ndarray, err_rd: np.ndarray, err_tt: np.ndarray,
                target: float, max_route: float, min_accept: float = 0.5) -> tuple[float, float]:
    """Per-output thresholds from held-out stacks: of the (tau_rd, tau_tt) pairs on a grid whose
    accepted stacks have a 90th-percentile R_d error and T_t error both within `target` (and
    accept at least `min_accept` of them), the pair accepting the most; if no pair meets the
    target, the single threshold for both outputs that routes `max_route` of them."""
    rel_rd, rel_tt = relative_spreads(pred)
    grid = np.geomspace(1e-4, 1.0, 61)
    n = len(domain)
    best, best_acc = None, -1
    for a in grid:
        ok_rd = domain & (rel_rd <= a)
        if ok_rd.sum() < min_accept * n:
            continue
        for b in grid:
            acc = ok_rd & (rel_tt <= b)
            k = int(acc.sum())
            if k < min_accept * n or k <= best_acc:
                continue
            if np.quantile(err_rd[acc], 0.9) <= target and np.quantile(err_tt[acc], 0.9) <= target:
                best, best_acc = (float(a), float(b)), k
    if best is not None:
        return best
    t = float(next((t for t in grid if np.mean(gate(pred, domain, t)) <= max_route), 1.0))
    return t, t


def rel_err(pred: np.ndarray, ref: np.ndarray, floor: float) -> np.ndarray:
    return np.abs(pred - ref) / np.maximum(np.abs(ref), floor)


def summarize(err: np.ndarray) -> dict:
    if len(err) == 0:
        return {"n": 0}
    return {"n": int(len(err)), "median": float(np.median(err)), "p90": float(np.quantile(err, 0.9)),
            "mean": float(err.mean()), "within_1pct": float(np.mean(err <= 0.01))}


def query_stacks(n_in: int, ood_frac: float, seed: int):
    """The inference workload's queries: n_in stacks of the training distribution, then
    round(n_in * ood_frac) out-of-range stacks (mus up to 1000 /cm, thickness up to 1 cm)."""
    n_ood = int(round(n_in * ood_frac))
    q_in = engine.generate_stacks(seed + 1, 0, n_in)
    q_ood = engine.generate_stacks(seed + 2, 0, n_ood, engine.default_stack_dist(mus_hi=1000.0, d_hi=1.0))
    out = (engine.TmcStack * (n_in + n_ood))()
    C.memmove(out, q_in, C.sizeof(q_in))
    C.memmove(C.byref(out, C.sizeof(q_in)), q_ood, C.sizeof(q_ood))
    return out


def bench_answers(n_in: int = 1_000_000, ood_frac: float = 0.05, tau: float = 0.02,
                  repeats: int = 3) -> dict:
    """End-to-end answer rate of the full-size network (4 heads, 30 -> 4 x 512 -> 202; random
    weights, which do not change the work) on the inference workload's queries: stacks in host
    memory -> gated R_d / T_t in host memory. The device path (tmc_mlp_answer: packing, transfers,
    encoding, network, gate) beside the host path it replaces (numpy encoding and domain gate,
    tmc_mlp_predict with every output returned, numpy de-standardization and gate)."""
    qs = query_stacks(n_in, ood_frac, seed=1)
    n = len(qs)
    m = Surrogate(seed=3)
    rng = np.random.default_rng(3)
    m.y_mean = rng.uniform(-2.0, -0.5, N_OUT).astype(np.float32)
    m.y_std = rng.uniform(0.2, 0.6, N_OUT).astype(np.float32)
    try:
        m.answer(engine.stacks_array([qs[i] for i in range(2000)]), tau)  # warm-up
        runs = [m.answer(qs, tau) for _ in range(repeats)]
        best = min(runs, key=lambda r: r["wall_seconds"])
        t0 = time.perf_counter()
        X, dom = features(qs, m.dist), in_domain(qs, m.dist)
        t_enc = time.perf_counter() - t0
        p = m.predict(X)
        route = gate(p, dom, tau)
        t_host = time.perf_counter() - t0
        agree = float(np.mean(route == best["route"]))
    finally:
        m.close()
    return {"queries": n, "tau": tau,
            "answers_per_s_end_to_end": n / best["wall_seconds"],
            "answer_wall_seconds": best["wall_seconds"], "answer_device_seconds": best["device_seconds"],
            "routed_frac": float(best["route"].mean()),
            "host_path_seconds": t_host, "host_path_encoding_seconds": t_enc,
            "host_path_answers_per_s": n / t_host, "route_agreement_with_host_path": agree}


# ------------------------------------------------------------------ the pipeline

def pipeline(args) -> dict:
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    report: dict = {"config": vars(args).copy()}
    dist = engine.default_stack_dist()

    # training data: the batched engine at a fixed budget per stack
    stacks = engine.generate_stacks(args.seed, 0, args.stacks)
    mc = simulate(stacks, args.photons, seed=args.seed * 7919, progress="data")
    X, Y = features(stacks), targets(mc)
    report["data"] = {"stacks": args.stacks, "photons_per_stack": args.photons,
                      "mc_kernel_seconds": mc.kernel_seconds, "mc_wall_seconds": mc.wall_seconds,
                      "mc_photons_per_s": args.stacks * args.photons / mc.kernel_seconds}
    log(f"data: {args.stacks} stacks in {mc.wall_seconds:.1f} s")

    n_val = min(args.val, args.stacks // 10)
    model = Surrogate(heads=args.heads, hidden=args.hidden, depth=args.depth, batch=args.batch,
                      seed=args.seed)
    try:
        fit = model.fit(X[n_val:], Y[n_val:], args.steps, lr=args.lr, scalar_weight=args.scalar_weight)
        per_step, per_query = model.net.flops()
        fit["tflops"] = per_step * args.steps / fit["seconds"] / 1e12
        fit["mfma_frac_bf16"] = fit["tflops"] / 2500.0
        report["train"] = fit
        model.save(out / "model.npz")

        # per-output gate thresholds chosen on the held-out stacks, against labels simulated at
        # args.val_photons (the training labels' Monte Carlo noise would otherwise hide the
        # network's own error: at 1e5 photons a mid-range T_t is known to about 1%)
        held = [stacks[i] for i in range(n_val)]
        pv = model.predict(X[:n_val])
        dom = in_domain(held, dist)
        vmc = simulate(held, args.val_photons, seed=args.seed * 7919 + 1, progress="validation")
        e_rd_v, e_tt_v = rel_err(pv["rd"], vmc.rd, 1e-3), rel_err(pv["tt"], vmc.tt, 1e-2)
        if args.tau > 0:
            tau = (args.tau, args.tau)
        else:
            tau = choose_gate(pv, dom, e_rd_v, e_tt_v, args.target_err, args.max_route)
        acc_v = ~gate(pv, dom, tau)
        report["gate"] = {"tau_rd": tau[0], "tau_tt": tau[1], "validation_stacks": n_val,
                          "validation_photons_per_stack": args.val_photons,
                          "validation_accept": float(np.mean(acc_v)),
                          "validation_rd_accepted": summarize(e_rd_v[acc_v]),
                          "validation_tt_accepted": summarize(e_tt_v[acc_v]),
                          "validation_err_all": summarize(np.maximum(e_rd_v, e_tt_v))}
        log(f"gate: tau_rd {tau[0]:.4g}, tau_tt {tau[1]:.4g}, validation accepts "
            f"{report['gate']['validation_accept']:.3f}")

        # queries: the training distribution plus out-of-range stacks, answered end to end on
        # the device (encoding, domain gate, ensemble, spread gate: tmc_mlp_answer)
        queries = query_stacks(args.queries, args.ood_frac, args.seed)
        nq = len(queries)
        model.answer(engine.stacks_array([queries[i] for i in range(2000)]), tau)  # warm-up
        ans = model.answer(queries, tau)
        route, dom_q = ans["route"], ans["in_domain"]
        pq = {"rd": ans["rd"], "tt": ans["tt"]}
        # the host path on the same queries (numpy encoding and gates around tmc_mlp_predict)
        t0 = time.perf_counter()
        host = model.predict(features(queries, dist))
        host_route = gate(host, in_domain(queries, dist), tau)
        t_host = time.perf_counter() - t0
        pq["az"] = host["az"]
        report["inference"] = {"queries": nq, "device_seconds": ans["device_seconds"],
                               "wall_seconds": ans["wall_seconds"],
                               "stacks_per_s_device": nq / ans["device_seconds"],
                               "stacks_per_s_end_to_end": nq / ans["wall_seconds"],
                               "network_device_seconds": host["seconds"],
                               "tflops": per_query * nq / host["seconds"] / 1e12,
                               "mfma_frac_bf16": per_query * nq / host["seconds"] / 2.5e15,
                               "host_path_seconds": t_host,
                               "host_path_route_agreement": float(np.mean(host_route == route)),
                               "routed": int(route.sum()), "routed_frac": float(route.mean()),
                               "out_of_domain": int((~dom_q).sum())}
        log(f"inference: {nq} queries, {ans['wall_seconds'] * 1e3:.1f} ms end to end "
            f"({ans['device_seconds'] * 1e3:.1f} ms on the device), {int(route.sum())} routed to MC")

        # fallback: Monte Carlo at args.fallback_photons for the routed queries
        routed_idx = np.nonzero(route)[0]
        fb_stacks = (engine.TmcStack * len(routed_idx))()
        for j, i in enumerate(routed_idx):
            fb_stacks[j] = queries[i]
        fb = simulate(fb_stacks, args.fallback_photons, seed=77, progress="fallback") if len(routed_idx) else None
        final_rd, final_tt = pq["rd"].copy(), pq["tt"].copy()
        if fb is not None:
            final_rd[routed_idx], final_tt[routed_idx] = fb.rd, fb.tt
        report["fallback"] = {"stacks": int(len(routed_idx)), "photons_per_stack": args.fallback_photons,
                              "kernel_seconds": fb.kernel_seconds if fb else 0.0,
                              "wall_seconds": fb.wall_seconds if fb else 0.0}

        # judge: Monte Carlo at many photons on a random subset of the queries
        rng = np.random.default_rng(args.seed)
        jidx = np.sort(rng.choice(nq, min(args.judge, nq), replace=False))
        js = (engine.TmcStack * len(jidx))()
        for j, i in enumerate(jidx):
            js[j] = queries[i]
        judge = simulate(js, args.judge_photons, seed=99, chunk=1024, progress="judge")
        acc = ~route[jidx]
        e_rd = rel_err(final_rd[jidx], judge.rd, 1e-3)
        e_tt = rel_err(final_tt[jidx], judge.tt, 1e-2)
        e_rd_s = rel_err(pq["rd"][jidx], judge.rd, 1e-3)
        e_tt_s = rel_err(pq["tt"][jidx], judge.tt, 1e-2)
        prof_z = np.abs(np.log10(pq["az"][jidx] + FLOOR) - np.log10(judge.az + FLOOR))
        report["judge"] = {
            "stacks": int(len(jidx)), "photons_per_stack": args.judge_photons,
            "rd_answered": summarize(e_rd), "tt_answered": summarize(e_tt),
            "rd_surrogate_accepted": summarize(e_rd_s[acc]), "tt_surrogate_accepted": summarize(e_tt_s[acc]),
            "rd_surrogate_all": summarize(e_rd_s), "tt_surrogate_all": summarize(e_tt_s),
            "rd_surrogate_routed": summarize(e_rd_s[~acc]), "tt_surrogate_routed": summarize(e_tt_s[~acc]),
            "log10_az_abs_err_accepted": summarize(prof_z[acc].ravel()),
            "judge_rd_se_median": float(np.median(judge.rd_se / np.maximum(judge.rd, 1e-3))),
        }

        # matched-error comparison: photons per stack that batched MC needs for the surrogate's
        # median R_d error on accepted stacks; MC's relative SE at N photons is se_1 / sqrt(N)
        eps = max(report["judge"]["rd_surrogate_accepted"].get("median", 0.01), 1e-4)
        se1 = np.median(mc.rd_se / np.maximum(mc.rd, 1e-3) * np.sqrt(args.photons))
        n_match = float((se1 / eps) ** 2)
        mc_rate = report["data"]["mc_photons_per_s"]
        mc_stacks_per_s = mc_rate / n_match
        fb_cost = (report["fallback"]["kernel_seconds"])
        total_s = ans["wall_seconds"] + fb_cost
        report["matched"] = {
            "surrogate_median_rd_err": eps, "mc_photons_for_matched_err": n_match,
            "mc_stacks_per_s_matched": mc_stacks_per_s,
            "surrogate_stacks_per_s_device": nq / ans["device_seconds"],
            "surrogate_stacks_per_s_end_to_end": nq / ans["wall_seconds"],
            "speedup_device_only": (nq / ans["device_seconds"]) / mc_stacks_per_s,
            "speedup_end_to_end": (nq / ans["wall_seconds"]) / mc_stacks_per_s,
            "answered_stacks_per_s_with_fallback": nq / total_s,
            "speedup_with_fallback": (nq / total_s) / mc_stacks_per_s,
        }
    finally:
        model.close()
    (out / "report.json").write_text(json.dumps(report, indent=1) + "\n")
    return report


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="tissuemc-surrogate", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("pipeline", help="data, training, gated inference, fallback and judge")
    p.add_argument("--out", required=True)
    p.add_argument("--stacks", type=int, default=1 << 20)
    p.add_argument("--photons", type=int, default=100_000)
    p.add_argument("--steps", type=int, default=360_000)
    p.add_argument("--scalar-weight", type=float, default=500.0,
                   help="loss weight of R_d and T_t (each profile bin weighs 1)")
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--heads", type=int, default=4)
    p.add_argument("--hidden", type=int, default=512)
    p.add_argument("--depth", type=int, default=4)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--val", type=int, default=20_000)
    p.add_argument("--val-photons", type=int, default=1_000_000,
                   help="photons per held-out stack for the labels the gate is chosen on")
    p.add_argument("--tau", type=float, default=0.0, help="gate threshold; 0 = chosen on held-out data")
    p.add_argument("--target-err", type=float, default=0.01)
    p.add_argument("--max-route", type=float, default=0.1,
                   help="held-out fraction routed to MC when no threshold meets --target-err")
    p.add_argument("--queries", type=int, default=1_000_000)
    p.add_argument("--ood-frac", type=float, default=0.05)
    p.add_argument("--fallback-photons", type=int, default=1_000_000)
    p.add_argument("--judge", type=int, default=10_000)
    p.add_argument("--judge-photons", type=int, default=10_000_000)
    args = ap.parse_args(argv)
    if args.cmd == "pipeline":
        rep = pipeline(args)
        print(json.dumps(rep))
    return 0


import ctypes as _C  # noqa: E402

C_SIZE = _C.sizeof(engine.TmcStack)

if __name__ == "__main__":
    sys.exit(main())
