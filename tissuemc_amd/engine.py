"""Python binding of the MI355X engine's C ABI (include/tissuemc_hip.h) through ctypes.

``run()`` returns a dict in the reference's result schema (src/io/json_out.cpp:164-306) with the
GPU's extras. There is no fallback: if the HIP library is missing or no GPU is present, the call
raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libtissuemc_hip.so"
MAX_LAYERS = 16
UNIT = 2.0 ** -32  # weight of one tally unit

TMC_OK, TMC_ERR_CONFIG, TMC_ERR_DEVICE, TMC_ERR_RUNTIME = 0, 1, 2, 3
FLAG_ARZ, FLAG_NO_PROFILES, FLAG_NO_GRAPH, FLAG_ANGLES, FLAG_AGGREGATE, FLAG_TIME = 1, 2, 4, 8, 16, 32


class ConfigError(ValueError):
    """Invalid input or options (the reference's ConfigError)."""


class DeviceError(RuntimeError):
    """No usable GPU, or a HIP call failed."""


class TmcLayer(C.Structure):
    _fields_ = [("n", C.c_double), ("mua", C.c_double), ("mus", C.c_double), ("g", C.c_double),
                ("d", C.c_double)]


class TmcSpec(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("dz", C.c_double), ("dr", C.c_double),
                ("nz", C.c_int32), ("nr", C.c_int32), ("na", C.c_int32),
                ("n_above", C.c_double), ("n_below", C.c_double), ("num_layers", C.c_int32),
                ("layers", TmcLayer * MAX_LAYERS), ("output_name", C.c_char * 128)]


class TmcOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("batches", C.c_int32), ("device", C.c_int32),
                ("photons", C.c_uint64), ("stream", C.c_uint32), ("flags", C.c_uint32),
                ("waves_per_group", C.c_int32), ("groups", C.c_int32), ("chunk", C.c_int32),
                ("arz_tile_z", C.c_int32), ("arz_tile_r", C.c_int32), ("step_limit", C.c_uint64),
                ("dt", C.c_double), ("nt", C.c_int32)]


_PU64 = C.POINTER(C.c_uint64)


class TmcOut(C.Structure):
    _fields_ = [("rd", _PU64), ("tt", _PU64), ("absorbed", _PU64), ("events", _PU64),
                ("anomalies", _PU64), ("boundary_steps", _PU64), ("az", _PU64), ("rr", _PU64),
                ("tr", _PU64), ("arz", _PU64), ("specular", C.c_double),
                ("kernel_seconds", C.c_double), ("wall_seconds", C.c_double),
                ("device", C.c_char * 96), ("groups", C.c_int32), ("waves_per_group", C.c_int32),
                ("lds_bytes", C.c_int32), ("chunk", C.c_int32), ("arz_tile_z", C.c_int32),
                ("arz_tile_r", C.c_int32), ("ra", _PU64), ("ta", _PU64), ("rra", _PU64),
                ("tra", _PU64), ("rt", _PU64), ("tt_t", _PU64)]


class TmcStack(C.Structure):
    _fields_ = [("num_layers", C.c_int32), ("n_above", C.c_double), ("n_below", C.c_double),
                ("layers", TmcLayer * MAX_LAYERS), ("photons", C.c_uint64)]


class TmcBatchedOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("batches_per_stack", C.c_int32), ("device", C.c_int32),
                ("dz", C.c_double), ("dr", C.c_double), ("nz", C.c_int32), ("nr", C.c_int32),
                ("flags", C.c_uint32), ("waves_per_group", C.c_int32), ("groups", C.c_int32),
                ("chunk", C.c_int32), ("step_limit", C.c_uint64), ("photons_hint", C.c_uint64)]


class TmcBatchedOut(C.Structure):
    _fields_ = [("rd", _PU64), ("tt", _PU64), ("absorbed", _PU64), ("events", _PU64),
                ("anomalies", _PU64), ("boundary_steps", _PU64), ("az", _PU64), ("rr", _PU64),
                ("tr", _PU64), ("kernel_seconds", C.c_double), ("wall_seconds", C.c_double),
                ("device", C.c_char * 96), ("groups", C.c_int32), ("waves_per_group", C.c_int32),
                ("lds_bytes", C.c_int32), ("chunk", C.c_int32)]


class TmcStackDist(C.Structure):
    _fields_ = [("min_layers", C.c_int32), ("max_layers", C.c_int32),
                ("mua_lo", C.c_double), ("mua_hi", C.c_double), ("mus_lo", C.c_double),
                ("mus_hi", C.c_double), ("g_lo", C.c_double), ("g_hi", C.c_double),
                ("n_lo", C.c_double), ("n_hi", C.c_double), ("d_lo", C.c_double),
                ("d_hi", C.c_double), ("photons_lo", C.c_double), ("photons_hi", C.c_double),
                ("n_ambient", C.c_double)]


TUNE_MAX = 16


class TmcTuneEntry(C.Structure):
    _fields_ = [("waves_per_group", C.c_int32), ("groups", C.c_int32), ("lds_bytes", C.c_int32),
                ("arz_tile_z", C.c_int32), ("arz_tile_r", C.c_int32), ("kernel_seconds", C.c_double),
                ("confirm_seconds", C.c_double)]


class TmcTuneReport(C.Structure):
    _fields_ = [("pilot_fraction", C.c_double), ("confirm_fraction", C.c_double),
                ("n_tried", C.c_int32), ("best", C.c_int32),
                ("tried", TmcTuneEntry * TUNE_MAX)]


EXPORTED = ("tmc_hip_available", "tmc_hip_run", "tmc_last_error", "tmc_fixed_point_unit",
            "tmc_default_opts", "tmc_validate_spec", "tmc_parse_mci_file", "tmc_parse_mci_text",
            "tmc_hip_philox_stream", "tmc_batched_create", "tmc_batched_execute",
            "tmc_batched_destroy", "tmc_hip_run_batched", "tmc_default_stack_dist",
            "tmc_generate_stacks", "tmc_mlp_default_config", "tmc_mlp_create", "tmc_mlp_destroy",
            "tmc_mlp_param_count", "tmc_mlp_set_data", "tmc_mlp_train", "tmc_mlp_gradients",
            "tmc_mlp_predict", "tmc_mlp_get_params", "tmc_mlp_set_params", "tmc_mlp_set_lr",
            "tmc_mlp_flops", "tmc_autotune", "tmc_batched_autotune", "tmc_hip_run_static",
            "tmc_static_scratch_bytes", "tmc_mlp_answer")

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """The engine library. Raises if it has not been built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise DeviceError(f"the HIP engine library is not built ({LIB_PATH}); run "
                          "__graft_entry__.build() or `make -C tissuemc_amd`")
    L = C.CDLL(str(LIB_PATH))
    L.tmc_hip_available.restype = C.c_int
    L.tmc_hip_run.argtypes = [C.POINTER(TmcSpec), C.POINTER(TmcOpts), C.POINTER(TmcOut)]
    L.tmc_hip_run.restype = C.c_int
    L.tmc_hip_run_static.argtypes = [C.POINTER(TmcSpec), C.POINTER(TmcOpts), C.POINTER(TmcOut)]
    L.tmc_hip_run_static.restype = C.c_int
    L.tmc_static_scratch_bytes.restype = C.c_uint64
    L.tmc_last_error.restype = C.c_char_p
    L.tmc_fixed_point_unit.restype = C.c_double
    L.tmc_default_opts.argtypes = [C.POINTER(TmcOpts)]
    L.tmc_validate_spec.argtypes = [C.POINTER(TmcSpec)]
    L.tmc_validate_spec.restype = C.c_int
    L.tmc_parse_mci_file.argtypes = [C.c_char_p, C.POINTER(TmcSpec), C.c_int32,
                                     C.POINTER(C.c_int32)]
    L.tmc_parse_mci_file.restype = C.c_int
    L.tmc_parse_mci_text.argtypes = [C.c_char_p, C.POINTER(TmcSpec), C.c_int32,
                                     C.POINTER(C.c_int32)]
    L.tmc_parse_mci_text.restype = C.c_int
    L.tmc_hip_philox_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64,
                                        C.POINTER(C.c_uint32), C.c_int32]
    L.tmc_hip_philox_stream.restype = C.c_int
    L.tmc_batched_create.argtypes = [C.c_int32, C.POINTER(TmcBatchedOpts), C.POINTER(C.c_void_p)]
    L.tmc_batched_create.restype = C.c_int
    L.tmc_batched_execute.argtypes = [C.c_void_p, C.POINTER(TmcStack), C.c_int32, C.c_uint64,
                                      C.POINTER(TmcBatchedOut)]
    L.tmc_batched_execute.restype = C.c_int
    L.tmc_batched_destroy.argtypes = [C.c_void_p]
    L.tmc_batched_destroy.restype = None
    L.tmc_hip_run_batched.argtypes = [C.POINTER(TmcStack), C.c_int32, C.POINTER(TmcBatchedOpts),
                                      C.POINTER(TmcBatchedOut)]
    L.tmc_hip_run_batched.restype = C.c_int
    L.tmc_default_stack_dist.argtypes = [C.POINTER(TmcStackDist)]
    L.tmc_generate_stacks.argtypes = [C.c_uint64, C.c_uint64, C.c_int32, C.POINTER(TmcStackDist),
                                      C.POINTER(TmcStack)]
    L.tmc_generate_stacks.restype = C.c_int
    L.tmc_autotune.argtypes = [C.POINTER(TmcSpec), C.POINTER(TmcOpts), C.c_double,
                               C.POINTER(TmcOpts), C.POINTER(TmcTuneReport)]
    L.tmc_autotune.restype = C.c_int
    L.tmc_batched_autotune.argtypes = [C.c_void_p, C.POINTER(TmcStack), C.c_int32, C.c_uint64,
                                       C.c_double, C.POINTER(TmcTuneReport)]
    L.tmc_batched_autotune.restype = C.c_int
    _lib = L
    return L


def _raise(code: int) -> None:
    msg = load().tmc_last_error().decode()
    if code == TMC_ERR_CONFIG:
        raise ConfigError(msg)
    raise DeviceError(msg)


def available() -> bool:
    return bool(load().tmc_hip_available())


# ------------------------------------------------------------------ specs

def spec_to_c(spec: dict) -> TmcSpec:
    s = TmcSpec()
    s.photons = int(spec["photons"])
    s.dz, s.dr = float(spec["dz"]), float(spec["dr"])
    s.nz, s.nr, s.na = int(spec["nz"]), int(spec["nr"]), int(spec.get("na", 1))
    s.n_above, s.n_below = float(spec["n_above"]), float(spec["n_below"])
    layers = spec["layers"]
    if len(layers) > MAX_LAYERS:
        raise ConfigError(f"at most {MAX_LAYERS} layers are supported")
    s.num_layers = len(layers)
    for i, L in enumerate(layers):
        s.layers[i] = TmcLayer(L["n"], L["mua"], L["mus"], L["g"], L["d"])
    s.output_name = str(spec.get("output_name", "")).encode()[:127]
    return s


def spec_from_c(s: TmcSpec) -> dict:
    return {
        "output_name": s.output_name.decode(), "photons": int(s.photons), "dz": s.dz, "dr": s.dr,
        "nz": s.nz, "nr": s.nr, "na": s.na, "n_above": s.n_above, "n_below": s.n_below,
        "layers": [{"n": s.layers[i].n, "mua": s.layers[i].mua, "mus": s.layers[i].mus,
                    "g": s.layers[i].g, "d": s.layers[i].d} for i in range(s.num_layers)],
    }


def parse_mci(text: str | None = None, *, path: str | Path | None = None) -> list[dict]:
    """Runs of an .mci file (by path or text), via the library's reader."""
    L = load()
    runs = (TmcSpec * 64)()
    n = C.c_int32(0)
    if path is not None:
        code = L.tmc_parse_mci_file(str(path).encode(), runs, 64, C.byref(n))
    else:
        code = L.tmc_parse_mci_text(text.encode(), runs, 64, C.byref(n))
    if code != TMC_OK:
        _raise(code)
    return [spec_from_c(runs[i]) for i in range(n.value)]


def validate_spec(spec: dict) -> None:
    code = load().tmc_validate_spec(C.byref(spec_to_c(spec)))
    if code != TMC_OK:
        _raise(code)


# ------------------------------------------------------------------ runs

def _p(a: np.ndarray):
    return a.ctypes.data_as(_PU64)


def batch_estimate(sums, photons) -> tuple[float, float]:
    """(mean, se) of the per-photon value from per-batch sums (json_out.cpp:37-64)."""
    sums = np.asarray(sums, dtype=np.float64)
    photons = np.asarray(photons, dtype=np.float64)
    mean = float(sums.sum() / photons.sum())
    n = len(sums)
    if n < 2:
        return mean, float("nan")
    per = sums / photons
    return mean, float(np.sqrt(((per - per.mean()) ** 2).sum() / (n - 1) / n))


def _opts(*, seed, batches, device, photons, stream, arz, profiles, waves_per_group=0, groups=0,
          chunk=0, arz_tile=None, step_limit=0, angles=False, aggregate=False,
          time=None) -> TmcOpts:
    o = TmcOpts()
    load().tmc_default_opts(C.byref(o))
    o.seed, o.batches, o.device, o.photons, o.stream = seed, int(batches), device, photons, stream
    o.flags = ((FLAG_ARZ if arz else 0) | (0 if profiles else FLAG_NO_PROFILES)
               | (FLAG_ANGLES if angles else 0) | (FLAG_AGGREGATE if aggregate else 0)
               | (FLAG_TIME if time is not None else 0))
    if time is not None:
        o.nt, o.dt = int(time[0]), float(time[1])
    o.waves_per_group, o.groups, o.chunk, o.step_limit = waves_per_group, groups, chunk, step_limit
    if arz_tile is not None:
        o.arz_tile_z, o.arz_tile_r = arz_tile
    return o


def _report(r: TmcTuneReport) -> dict:
    tried = [{"waves_per_group": e.waves_per_group, "groups": e.groups, "lds_bytes": e.lds_bytes,
              "arz_tile": [e.arz_tile_z, e.arz_tile_r], "kernel_seconds": e.kernel_seconds,
              "confirm_seconds": e.confirm_seconds}
             for e in r.tried[:r.n_tried]]
    return {"pilot_fraction": r.pilot_fraction, "confirm_fraction": r.confirm_fraction,
            "best": r.best, "tried": tried}


def autotune(spec: dict, *, photons: int = 0, batches: int = 10, seed: int = 1, device: int = 0,
             arz: bool = False, profiles: bool = True, stream: int = 0, chunk: int = 0,
             step_limit: int = 0, pilot_fraction: float = 1 / 32, aggregate: bool = False) -> dict:
    """Pick the launch shape of a single run by timing pilot runs (tmc_autotune).

    Returns ``{"run_kwargs": ..., "report": ...}``; ``run(spec, **run_kwargs, ...)`` uses the
    shape chosen. Tallies do not depend on it.
    """
    L = load()
    cs = spec_to_c(spec)
    o = _opts(seed=seed, batches=batches, device=device, photons=photons or int(spec["photons"]),
              stream=stream, arz=arz, profiles=profiles, chunk=chunk, step_limit=step_limit,
              aggregate=aggregate)
    tuned, rep = TmcOpts(), TmcTuneReport()
    code = L.tmc_autotune(C.byref(cs), C.byref(o), float(pilot_fraction), C.byref(tuned),
                          C.byref(rep))
    if code != TMC_OK:
        _raise(code)
    kw = {"waves_per_group": tuned.waves_per_group, "groups": tuned.groups}
    if arz:
        kw["arz_tile"] = (tuned.arz_tile_z, tuned.arz_tile_r)
    return {"run_kwargs": kw, "report": _report(rep)}


def run(spec: dict, *, photons: int = 0, batches: int = 10, seed: int = 1, device: int = 0,
        arz: bool = False, profiles: bool = True, stream: int = 0, waves_per_group: int = 0,
        groups: int = 0, chunk: int = 0, arz_tile: tuple[int, int] | None = None,
        step_limit: int = 0, angles: bool = False, aggregate: bool = False,
        time: tuple[int, float] | None = None, no_alloc: bool = False) -> dict:
    """Simulate one run on the GPU; returns the reference's result schema plus extras.

    With ``no_alloc`` the run goes through tmc_hip_run_static: no device allocation, every
    device buffer from the library's static scratch array (same results).

    Extras: ``t_r``, ``a_rz`` (if ``arz``), the exit-angle profiles ``rd_a``/``t_a`` [na] and
    ``rd_ra``/``t_ra`` [nr, na] (if ``angles``), the time-of-flight profiles ``rd_t``/``t_t``
    [nt] (if ``time = (nt, dt_ns)``), ``profile_se`` (per-bin standard errors),
    per-batch profiles ``az_b``/``rr_b``/``tr_b`` and the raw integer tallies under ``fixed``.
    """
    L = load()
    cs = spec_to_c(spec)
    n_ph = photons or int(spec["photons"])
    B, NL, nz, nr = int(batches), cs.num_layers, cs.nz, cs.nr
    o = _opts(seed=seed, batches=B, device=device, photons=n_ph, stream=stream, arz=arz,
              profiles=profiles, waves_per_group=waves_per_group, groups=groups, chunk=chunk,
              arz_tile=arz_tile, step_limit=step_limit, angles=angles, aggregate=aggregate,
              time=time)
    if B < 1:
        raise ConfigError("batches must be >= 1")
    na = max(int(cs.na), 1)
    rd, tt, ev, an, bs = (np.zeros(B, np.uint64) for _ in range(5))
    ab = np.zeros(B * NL, np.uint64)
    az = np.zeros(B * nz, np.uint64)
    rr = np.zeros(B * nr, np.uint64)
    tr = np.zeros(B * nr, np.uint64)
    a_rz = np.zeros(nz * nr, np.uint64) if arz else None
    ang = [np.zeros(na, np.uint64), np.zeros(na, np.uint64), np.zeros(nr * na, np.uint64),
           np.zeros(nr * na, np.uint64)] if angles else None
    out = TmcOut(rd=_p(rd), tt=_p(tt), absorbed=_p(ab), events=_p(ev), anomalies=_p(an),
                 boundary_steps=_p(bs), az=_p(az), rr=_p(rr), tr=_p(tr),
                 arz=_p(a_rz) if arz else None)
    if angles:
        out.ra, out.ta, out.rra, out.tra = (_p(x) for x in ang)
    tof = [np.zeros(max(int(time[0]), 1), np.uint64) for _ in range(2)] if time is not None else None
    if time is not None:
        out.rt, out.tt_t = _p(tof[0]), _p(tof[1])
    entry = L.tmc_hip_run_static if no_alloc else L.tmc_hip_run
    code = entry(C.byref(cs), C.byref(o), C.byref(out))
    if code != TMC_OK:
        _raise(code)

    base, extra = divmod(n_ph, B)
    firsts = [b * base + min(b, extra) for b in range(B)]
    counts = [base + (1 if b < extra else 0) for b in range(B)]
    az_b = az.reshape(B, nz)
    rr_b = rr.reshape(B, nr)
    tr_b = tr.reshape(B, nr)

    def prof(pb: np.ndarray) -> tuple[list[float], list[float]]:
        tot = [float(v) * UNIT for v in pb.sum(axis=0, dtype=np.uint64)]
        per = pb.astype(np.float64) * UNIT / np.asarray(counts, np.float64)[:, None]
        se = (per.std(axis=0, ddof=1) / np.sqrt(B)).tolist() if B > 1 else [float("nan")] * pb.shape[1]
        return tot, se

    a_z, a_z_se = prof(az_b)
    rd_r, rd_r_se = prof(rr_b)
    t_r, t_r_se = prof(tr_b)
    batches_out = [
        {"first_photon": firsts[b], "photons": counts[b], "rd": float(rd[b]) * UNIT,
         "tt": float(tt[b]) * UNIT,
         "absorbed": [float(x) * UNIT for x in ab[b * NL:(b + 1) * NL]],
         "events": int(ev[b]), "anomalies": int(an[b]), "boundary_steps": int(bs[b])}
        for b in range(B)
    ]
    res = {
        "engine": "hip", "precision": "float32", "seed": seed, "threads": 1,
        "photons": int(sum(counts)), "output_name": spec.get("output_name", ""),
        "n_above": cs.n_above, "n_below": cs.n_below,
        "layers": [dict(x) for x in spec["layers"]],
        "grid": {"dz": cs.dz, "dr": cs.dr, "nz": nz, "nr": nr},
        "specular": out.specular, "wall_seconds": out.wall_seconds,
        "kernel_seconds": out.kernel_seconds, "device": out.device.decode(),
        "tally_precision": "int64 fixed point (2^-32)", "rung": 0, "profiles": profiles,
        "events": int(ev.sum()), "anomalies": int(an.sum()), "boundary_steps": int(bs.sum()),
        "launch": {"groups": out.groups, "waves_per_group": out.waves_per_group,
                   "lds_bytes": out.lds_bytes, "chunk": out.chunk,
                   "arz_tile": [out.arz_tile_z, out.arz_tile_r]},
        "batches": batches_out,
        "a_z": a_z, "rd_r": rd_r, "t_r": t_r,
        "profile_se": {"a_z": a_z_se, "rd_r": rd_r_se, "t_r": t_r_se},
        "az_b": az_b, "rr_b": rr_b, "tr_b": tr_b,
        "fixed": {"rd": rd, "tt": tt, "absorbed": ab.reshape(B, NL), "az": az_b, "rr": rr_b,
                  "tr": tr_b, "arz": a_rz.reshape(nz, nr) if arz else None},
    }
    photons_b = [bb["photons"] for bb in batches_out]
    summ = {}
    for key in ("rd", "tt"):
        m, s = batch_estimate([bb[key] for bb in batches_out], photons_b)
        summ[key] = {"mean": m, "se": s}
    m, s = batch_estimate([sum(bb["absorbed"]) for bb in batches_out], photons_b)
    summ["absorbed"] = {"mean": m, "se": s}
    m, s = batch_estimate([bb["rd"] + bb["tt"] + sum(bb["absorbed"]) for bb in batches_out],
                          photons_b)
    summ["conservation_residual"] = {"mean": m + out.specular - 1.0, "se": s}
    res["summary"] = summ
    if arz:
        res["a_rz"] = (a_rz.astype(np.float64) * UNIT).reshape(nz, nr)
    if angles:
        res["fixed"].update(ra=ang[0], ta=ang[1], rra=ang[2].reshape(nr, na),
                            tra=ang[3].reshape(nr, na))
        res["rd_a"] = ang[0].astype(np.float64) * UNIT
        res["t_a"] = ang[1].astype(np.float64) * UNIT
        res["rd_ra"] = (ang[2].astype(np.float64) * UNIT).reshape(nr, na)
        res["t_ra"] = (ang[3].astype(np.float64) * UNIT).reshape(nr, na)
    if time is not None:
        res["fixed"].update(rt=tof[0], tt_t=tof[1])
        res["rd_t"] = tof[0].astype(np.float64) * UNIT
        res["t_t"] = tof[1].astype(np.float64) * UNIT
        res["time_grid"] = {"nt": int(time[0]), "dt_ns": float(time[1])}
    return res


def philox_stream(seed: int, photon_id: int, n: int, word3: int = 0, device: int = 0) -> np.ndarray:
    """The first n words of the device's Philox stream for (seed, photon_id, word3)."""
    out = np.zeros(n, np.uint32)
    code = load().tmc_hip_philox_stream(seed, photon_id, word3, n,
                                        out.ctypes.data_as(C.POINTER(C.c_uint32)), device)
    if code != TMC_OK:
        _raise(code)
    return out


# ------------------------------------------------------------------ batched runs and stacks

def default_stack_dist(**overrides) -> TmcStackDist:
    """The batched workload's stack distribution, with optional field overrides."""
    d = TmcStackDist()
    load().tmc_default_stack_dist(C.byref(d))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def generate_stacks(seed: int, first_id: int, count: int, dist: TmcStackDist | None = None):
    """Stacks first_id .. first_id + count - 1 of the sequence keyed by seed (a ctypes array)."""
    arr = (TmcStack * count)()
    d = dist if dist is not None else default_stack_dist()
    code = load().tmc_generate_stacks(seed, first_id, count, C.byref(d), arr)
    if code != TMC_OK:
        raise ConfigError("invalid stack distribution")
    return arr


def stack_from_dict(st: dict) -> TmcStack:
    s = TmcStack()
    s.num_layers = len(st["layers"])
    s.n_above, s.n_below = float(st["n_above"]), float(st["n_below"])
    for i, L in enumerate(st["layers"]):
        s.layers[i] = TmcLayer(L["n"], L["mua"], L["mus"], L["g"], L["d"])
    s.photons = int(st["photons"])
    return s


def stack_to_dict(s: TmcStack) -> dict:
    return {"n_above": s.n_above, "n_below": s.n_below, "photons": int(s.photons),
            "layers": [{"n": s.layers[i].n, "mua": s.layers[i].mua, "mus": s.layers[i].mus,
                        "g": s.layers[i].g, "d": s.layers[i].d} for i in range(s.num_layers)]}


def stacks_array(stacks) -> C.Array:
    if isinstance(stacks, C.Array):
        return stacks
    arr = (TmcStack * len(stacks))()
    for i, st in enumerate(stacks):
        arr[i] = st if isinstance(st, TmcStack) else stack_from_dict(st)
    return arr


class BatchedPlan:
    """A resident batched plan: device buffers allocated once, clear + transport replayed as a
    HIP graph on every execute()."""

    def __init__(self, n_stacks: int, *, batches_per_stack: int = 10, dz: float = 0.01,
                 dr: float = 0.01, nz: int = 100, nr: int = 100, device: int = 0,
                 profiles: bool = True, graph: bool = True, aggregate: bool = False,
                 waves_per_group: int = 0,
                 groups: int = 0, chunk: int = 0, step_limit: int = 0, photons_hint: int = 0):
        self.n_stacks, self.bs, self.nz, self.nr = n_stacks, batches_per_stack, nz, nr
        self._h = C.c_void_p()
        flags = ((0 if profiles else FLAG_NO_PROFILES) | (0 if graph else FLAG_NO_GRAPH)
                 | (FLAG_AGGREGATE if aggregate else 0))
        o = TmcBatchedOpts(seed=1, batches_per_stack=batches_per_stack, device=device, dz=dz,
                           dr=dr, nz=nz, nr=nr, flags=flags,
                           waves_per_group=waves_per_group, groups=groups, chunk=chunk,
                           step_limit=step_limit, photons_hint=photons_hint)
        code = load().tmc_batched_create(n_stacks, C.byref(o), C.byref(self._h))
        if code != TMC_OK:
            _raise(code)

    def autotune(self, stacks, seed: int = 1, pilot_fraction: float = 1 / 16) -> dict:
        """Time pilot runs of these stacks at each candidate launch shape and keep the fastest
        for later executes (tmc_batched_autotune); returns the report."""
        arr = stacks_array(stacks)
        rep = TmcTuneReport()
        code = load().tmc_batched_autotune(self._h, arr, len(arr), seed, float(pilot_fraction),
                                           C.byref(rep))
        if code != TMC_OK:
            _raise(code)
        return _report(rep)

    def execute(self, stacks, seed: int = 1) -> dict:
        arr = stacks_array(stacks)
        S, Bs, nz, nr = self.n_stacks, self.bs, self.nz, self.nr
        rd, tt, ev, an, bs = (np.zeros(S * Bs, np.uint64) for _ in range(5))
        ab = np.zeros(S * Bs * MAX_LAYERS, np.uint64)
        az = np.zeros(S * nz, np.uint64)
        rr = np.zeros(S * nr, np.uint64)
        tr = np.zeros(S * nr, np.uint64)
        out = TmcBatchedOut(rd=_p(rd), tt=_p(tt), absorbed=_p(ab), events=_p(ev), anomalies=_p(an),
                            boundary_steps=_p(bs), az=_p(az), rr=_p(rr), tr=_p(tr))
        code = load().tmc_batched_execute(self._h, arr, len(arr), seed, C.byref(out))
        if code != TMC_OK:
            _raise(code)
        photons = np.array([arr[s].photons for s in range(len(arr))], np.uint64)
        return {
            "rd": rd.reshape(S, Bs), "tt": tt.reshape(S, Bs),
            "absorbed": ab.reshape(S, Bs, MAX_LAYERS), "events": ev.reshape(S, Bs),
            "anomalies": an.reshape(S, Bs), "boundary_steps": bs.reshape(S, Bs),
            "az": az.reshape(S, nz), "rr": rr.reshape(S, nr), "tr": tr.reshape(S, nr),
            "photons": photons, "kernel_seconds": out.kernel_seconds,
            "wall_seconds": out.wall_seconds, "device": out.device.decode(),
            "launch": {"groups": out.groups, "waves_per_group": out.waves_per_group,
                       "lds_bytes": out.lds_bytes, "chunk": out.chunk},
        }

    def close(self) -> None:
        if self._h:
            load().tmc_batched_destroy(self._h)
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


def run_batched(stacks, *, seed: int = 1, **plan_kw) -> dict:
    """One batched run (plan created and destroyed around a single execute)."""
    arr = stacks_array(stacks)
    with BatchedPlan(len(arr), **plan_kw) as plan:
        return plan.execute(arr, seed)


def batch_plan_rows(photons: int, batches: int) -> list[tuple[int, int]]:
    """(first id, count) of each batch: the reference's BatchPlan (batching.hpp:11-29)."""
    base, extra = divmod(photons, batches)
    return [(b * base + min(b, extra), base + (1 if b < extra else 0)) for b in range(batches)]
