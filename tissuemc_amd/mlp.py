"""Python binding of the surrogate network's C ABI (tmc_mlp_*, include/tissuemc_hip.h).

An ensemble of MLPs trained and evaluated by the library's BF16 MFMA kernels. Parameters per
head, per layer: W [out x in] row-major, then b [out], float32.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import engine


class MlpConfig(C.Structure):
    _fields_ = [("heads", C.c_int32), ("inputs", C.c_int32), ("hidden", C.c_int32),
                ("depth", C.c_int32), ("outputs", C.c_int32), ("batch", C.c_int32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("weight_decay", C.c_float), ("seed", C.c_uint64), ("device", C.c_int32)]


class SurrogateGate(C.Structure):
    _fields_ = [("box", engine.TmcStackDist), ("y_mean", C.POINTER(C.c_float)),
                ("y_std", C.POINTER(C.c_float)), ("floor", C.c_double), ("floor_rd", C.c_double),
                ("floor_tt", C.c_double), ("tau_rd", C.c_double), ("tau_tt", C.c_double)]


_PD = C.POINTER(C.c_double)
_PU8 = C.POINTER(C.c_uint8)


class SurrogateAnswers(C.Structure):
    _fields_ = [("rd", _PD), ("tt", _PD), ("rd_spread", _PD), ("tt_spread", _PD),
                ("route", _PU8), ("in_domain", _PU8), ("profiles", C.POINTER(C.c_float)),
                ("device_seconds", C.c_double), ("wall_seconds", C.c_double)]


_PF = C.POINTER(C.c_float)
_bound = False


def _lib():
    global _bound
    L = engine.load()
    if not _bound:
        L.tmc_mlp_default_config.argtypes = [C.POINTER(MlpConfig)]
        L.tmc_mlp_create.argtypes = [C.POINTER(MlpConfig), C.POINTER(C.c_void_p)]
        L.tmc_mlp_create.restype = C.c_int
        L.tmc_mlp_destroy.argtypes = [C.c_void_p]
        L.tmc_mlp_destroy.restype = None
        L.tmc_mlp_param_count.argtypes = [C.c_void_p]
        L.tmc_mlp_param_count.restype = C.c_int64
        L.tmc_mlp_set_data.argtypes = [C.c_void_p, _PF, _PF, C.c_int64, _PF]
        L.tmc_mlp_set_data.restype = C.c_int
        L.tmc_mlp_train.argtypes = [C.c_void_p, C.c_int64, _PF, C.POINTER(C.c_double)]
        L.tmc_mlp_train.restype = C.c_int
        L.tmc_mlp_gradients.argtypes = [C.c_void_p, C.POINTER(C.c_int64), _PF, _PF]
        L.tmc_mlp_gradients.restype = C.c_int
        L.tmc_mlp_predict.argtypes = [C.c_void_p, _PF, C.c_int64, _PF, _PF, _PF,
                                      C.POINTER(C.c_double)]
        L.tmc_mlp_predict.restype = C.c_int
        L.tmc_mlp_get_params.argtypes = [C.c_void_p, _PF]
        L.tmc_mlp_get_params.restype = C.c_int
        L.tmc_mlp_set_params.argtypes = [C.c_void_p, _PF]
        L.tmc_mlp_set_params.restype = C.c_int
        L.tmc_mlp_set_lr.argtypes = [C.c_void_p, C.c_float]
        L.tmc_mlp_set_lr.restype = C.c_int
        L.tmc_mlp_flops.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.tmc_mlp_flops.restype = C.c_int
        L.tmc_mlp_answer.argtypes = [C.c_void_p, C.POINTER(SurrogateGate), C.POINTER(engine.TmcStack),
                                     C.c_int64, C.POINTER(SurrogateAnswers)]
        L.tmc_mlp_answer.restype = C.c_int
        _bound = True
    return L


def _f(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_PF)


def default_config(**overrides) -> MlpConfig:
    c = MlpConfig()
    _lib().tmc_mlp_default_config(C.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class Mlp:
    def __init__(self, **config):
        self.config = default_config(**config)
        self._h = C.c_void_p()
        code = _lib().tmc_mlp_create(C.byref(self.config), C.byref(self._h))
        if code != engine.TMC_OK:
            engine._raise(code)
        self.heads = self.config.heads
        self.inputs, self.outputs = self.config.inputs, self.config.outputs
        self.param_count = int(_lib().tmc_mlp_param_count(self._h))

    def _check(self, code: int) -> None:
        if code != engine.TMC_OK:
            engine._raise(code)

    def set_data(self, x: np.ndarray, y: np.ndarray, loss_weights: np.ndarray | None = None) -> None:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        w = None if loss_weights is None else np.ascontiguousarray(loss_weights, np.float32)
        assert x.shape[1] == self.inputs and y.shape[1] == self.outputs and len(x) == len(y)
        self._check(_lib().tmc_mlp_set_data(self._h, _f(x), _f(y), len(x), _f(w)))

    def train(self, steps: int) -> tuple[np.ndarray, float]:
        """Run Adam steps; returns (losses [steps x heads], device seconds)."""
        losses = np.zeros((steps, self.heads), np.float32)
        sec = C.c_double(0)
        self._check(_lib().tmc_mlp_train(self._h, steps, _f(losses), C.byref(sec)))
        return losses, sec.value

    def gradients(self, rows: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        rows = np.ascontiguousarray(rows, np.int64)
        assert len(rows) == self.config.batch
        g = np.zeros((self.heads, self.param_count), np.float32)
        loss = np.zeros(self.heads, np.float32)
        self._check(_lib().tmc_mlp_gradients(self._h, rows.ctypes.data_as(C.POINTER(C.c_int64)),
                                             _f(g), _f(loss)))
        return g, loss

    def predict(self, x: np.ndarray, per_head: bool = False):
        """(mean [n x outputs], spread [n x 2], per-head [heads x n x outputs] or None, seconds)."""
        x = np.ascontiguousarray(x, np.float32)
        n = len(x)
        mean = np.zeros((n, self.outputs), np.float32)
        spread = np.zeros((n, 2), np.float32)
        ph = np.zeros((self.heads, n, self.outputs), np.float32) if per_head else None
        sec = C.c_double(0)
        self._check(_lib().tmc_mlp_predict(self._h, _f(x), n, _f(mean), _f(spread), _f(ph),
                                           C.byref(sec)))
        return mean, spread, ph, sec.value

    def answer(self, stacks, *, y_mean: np.ndarray, y_std: np.ndarray, tau,
               floor: float = 1e-7, floor_rd: float = 1e-7, floor_tt: float = 1e-7,
               box: engine.TmcStackDist | None = None, profiles: bool = False) -> dict:
        """Gated surrogate answers for stacks, end to end on the device (tmc_mlp_answer): R_d,
        T_t, their ensemble spreads, the route-to-Monte-Carlo and in-domain flags, optionally the
        other outputs' values, and the device and wall-clock seconds of the call. `tau` is the
        spread threshold of both outputs, or a pair (R_d's, T_t's). Output 0 is log10(R_d +
        floor_rd), output 1 log10(T_t + floor_tt), the others log10(value + floor)."""
        tau_rd, tau_tt = (tau, tau) if np.isscalar(tau) else tau
        arr = engine.stacks_array(stacks)
        n = len(arr)
        ym = np.ascontiguousarray(y_mean, np.float32)
        ys = np.ascontiguousarray(y_std, np.float32)
        assert ym.shape == ys.shape == (self.outputs,)
        gate = SurrogateGate(box=box if box is not None else engine.default_stack_dist(),
                             y_mean=_f(ym), y_std=_f(ys), floor=floor, floor_rd=floor_rd,
                             floor_tt=floor_tt, tau_rd=tau_rd, tau_tt=tau_tt)
        res = {k: np.zeros(n) for k in ("rd", "tt", "rd_spread", "tt_spread")}
        route = np.zeros(n, np.uint8)
        dom = np.zeros(n, np.uint8)
        prof = np.zeros((n, self.outputs - 2), np.float32) if profiles else None
        out = SurrogateAnswers(rd=res["rd"].ctypes.data_as(_PD), tt=res["tt"].ctypes.data_as(_PD),
                               rd_spread=res["rd_spread"].ctypes.data_as(_PD),
                               tt_spread=res["tt_spread"].ctypes.data_as(_PD),
                               route=route.ctypes.data_as(_PU8), in_domain=dom.ctypes.data_as(_PU8),
                               profiles=_f(prof))
        self._check(_lib().tmc_mlp_answer(self._h, C.byref(gate), arr, n, C.byref(out)))
        res.update(route=route.astype(bool), in_domain=dom.astype(bool),
                   device_seconds=out.device_seconds, wall_seconds=out.wall_seconds)
        if profiles:
            res["profiles"] = prof
        return res

    def get_params(self) -> np.ndarray:
        p = np.zeros((self.heads, self.param_count), np.float32)
        self._check(_lib().tmc_mlp_get_params(self._h, _f(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.heads, self.param_count)
        self._check(_lib().tmc_mlp_set_params(self._h, _f(p)))

    def set_lr(self, lr: float) -> None:
        self._check(_lib().tmc_mlp_set_lr(self._h, lr))

    def flops(self) -> tuple[int, int]:
        a, b = C.c_uint64(0), C.c_uint64(0)
        self._check(_lib().tmc_mlp_flops(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self) -> None:
        if self._h:
            _lib().tmc_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bench_kernels(queries: int = 1_050_000, train_steps: int = 200, seed: int = 0) -> dict:
    """The surrogate's kernels at the pipeline's full size with random weights and inputs:
    ensemble inference on `queries` rows and `train_steps` graph-replayed Adam steps (4 heads,
    30 -> 4 x 512 GELU -> 202, batch 4096). Device times are HIP events."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (queries, 30)).astype(np.float32)
    with Mlp() as net:
        net.predict(x[:20000])  # warm-up
        _, _, _, sec = net.predict(x)
        per_step, per_query = net.flops()
        n = min(queries, 65536)
        net.set_data(x[:n], rng.normal(0, 1, (n, 202)).astype(np.float32))
        net.train(20)
        _, tsec = net.train(train_steps)
    return {"inference_queries": queries, "inference_device_s": sec,
            "inference_stacks_per_s": queries / sec,
            "inference_tflops": per_query * queries / sec / 1e12,
            "inference_mfma_frac": per_query * queries / sec / 2.5e15,
            "train_step_ms": tsec / train_steps * 1e3,
            "train_tflops": per_step * train_steps / tsec / 1e12,
            "train_mfma_frac": per_step * train_steps / tsec / 2.5e15}


def bench_library(rows=(4096, 16384), reps: int = 50) -> dict:
    """A yardstick for the hidden-layer product: the vendor library (hipBLASLt, through torch's
    bf16 bmm) on the same shape, 4 heads x [rows x 512] . [512 x 512], alone and followed by the
    bias add and tanh-GELU as separate operations (what an unfused framework runs). Device times
    from torch's events on its current stream, microseconds per call."""
    import torch

    def timed(fn):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    g = torch.Generator(device="cuda").manual_seed(0)
    out = {}
    for m in rows:
        a = torch.randn(4, m, 512, device="cuda", dtype=torch.bfloat16, generator=g)
        w = torch.randn(4, 512, 512, device="cuda", dtype=torch.bfloat16, generator=g)
        bias = torch.randn(4, 1, 512, device="cuda", dtype=torch.bfloat16, generator=g)
        c = torch.empty(4, m, 512, device="cuda", dtype=torch.bfloat16)
        flops = 2 * 4 * m * 512 * 512
        t_mm = timed(lambda: torch.bmm(a, w, out=c))
        t_gelu = timed(lambda: torch.nn.functional.gelu(torch.baddbmm(bias, a, w), approximate="tanh"))
        out[f"rows_{m}"] = {"bmm_us": t_mm, "bmm_tflops": flops / t_mm / 1e6,
                            "bmm_bias_gelu_us": t_gelu, "bmm_bias_gelu_tflops": flops / t_gelu / 1e6}
    return out


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python -m tissuemc_amd.mlp",
                                 description="time the surrogate network's kernels")
    ap.add_argument("command", choices=["bench", "library"],
                    help="bench: this package's kernels; library: the vendor GEMM on the "
                         "hidden-layer shape, as a yardstick")
    ap.add_argument("--queries", type=int, default=1_050_000)
    ap.add_argument("--train-steps", type=int, default=200)
    args = ap.parse_args(argv)
    if args.command == "library":
        print(json.dumps(bench_library()))
    else:
        print(json.dumps(bench_kernels(args.queries, args.train_steps)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
