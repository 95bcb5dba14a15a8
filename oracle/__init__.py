"""ORACLE — test infrastructure, not product code.

Python bindings for the C restatement of the reference's CPU engine (mc_oracle.c) and for the
reference engine built from its own sources (oracle/_ref/tissuemc_ref), plus a small .mci reader.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this package, and
only as the checker; the product never imports it.

Result dictionaries follow the reference's JSON schema (src/io/json_out.cpp:164-306): per-batch
``rd``/``tt``/``absorbed``/``events``/``anomalies``, ``a_z`` and ``rd_r`` summed over batches in
batch order (src/cpu/engine.cpp:201-209). Extras with no reference counterpart are returned as
numpy arrays: per-batch ``az_b``/``rr_b``/``tr_b``, ``t_r`` and ``a_rz``.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "bin" / "liboracle.so"
REF_BIN = HERE / "_ref" / "tissuemc_ref"
MAX_LAYERS = 16


class OrcLayer(C.Structure):
    _fields_ = [("n", C.c_double), ("mua", C.c_double), ("mus", C.c_double), ("g", C.c_double),
                ("d", C.c_double)]


class OrcSpec(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("dz", C.c_double), ("dr", C.c_double),
                ("nz", C.c_int32), ("nr", C.c_int32), ("na", C.c_int32),
                ("n_above", C.c_double), ("n_below", C.c_double), ("num_layers", C.c_int32),
                ("layers", OrcLayer * MAX_LAYERS)]


class OrcOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("batches", C.c_int32), ("threads", C.c_int32),
                ("precision", C.c_int32), ("stream", C.c_uint32), ("photons", C.c_uint64),
                ("nt", C.c_int32), ("dt", C.c_double)]


_P64 = C.POINTER(C.c_uint64)
_PD = C.POINTER(C.c_double)


class OrcOut(C.Structure):
    _fields_ = [("first", _P64), ("count", _P64), ("rd", _PD), ("tt", _PD), ("absorbed", _PD),
                ("events", _P64), ("anomalies", _P64), ("boundary_steps", _P64), ("az", _PD),
                ("rr", _PD), ("tr", _PD), ("arz", _PD), ("ra", _PD), ("ta", _PD), ("rra", _PD),
                ("tra", _PD), ("rt", _PD), ("tt_t", _PD), ("specular", C.c_double),
                ("threads_used", C.c_int32), ("wall_seconds", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """Load (building first if needed) the oracle library."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, stdout=subprocess.DEVNULL)
        L = C.CDLL(str(LIB_PATH))
        L.orc_run.argtypes = [C.POINTER(OrcSpec), C.POINTER(OrcOpts), C.POINTER(OrcOut),
                              C.c_char_p, C.c_int]
        L.orc_run.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_stream_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64,
                                     C.POINTER(C.c_uint32)]
        L.orc_u32_to_double.argtypes = [C.c_uint32]
        L.orc_u32_to_double.restype = C.c_double
        L.orc_u32_to_float.argtypes = [C.c_uint32]
        L.orc_u32_to_float.restype = C.c_float
        L.orc_batch_of.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_batch_of.restype = C.c_uint64
        L.orc_critical_cos.argtypes = [C.c_double, C.c_double]
        L.orc_critical_cos.restype = C.c_double
        L.orc_fresnel.argtypes = [C.c_double, C.c_double, C.c_double, _PD, C.c_int]
        L.orc_fresnel.restype = C.c_double
        L.orc_hg_samples.argtypes = [C.c_double, C.c_uint64, C.c_uint64, C.c_int64, _PD, C.c_int]
        L.orc_spin_walk.argtypes = [C.c_double, C.c_uint64, C.c_uint64, C.c_int64, _PD, C.c_int64,
                                    C.c_int, _PD, _PD]
        L.orc_step_lengths.argtypes = [C.POINTER(OrcSpec), C.c_uint64, C.c_uint64, C.c_int64, _PD]
        L.orc_validate.argtypes = [C.POINTER(OrcSpec), C.c_char_p, C.c_int]
        L.orc_validate.restype = C.c_int
        _lib = L
    return _lib


# ------------------------------------------------------------------ .mci reading (src/io/mci.cpp)

def parse_mci_text(text: str) -> list[dict]:
    """Runs of an MCML .mci file as spec dicts (restates the reference parser's format)."""
    lines = []
    for raw in text.splitlines():
        raw = raw.split("#", 1)[0]
        toks = raw.split()
        if toks:
            lines.append(toks)
    pos = 0

    def nxt(count: int) -> list[str]:
        nonlocal pos
        toks = lines[pos]
        pos += 1
        if len(toks) != count:
            raise ValueError(f"expected {count} values, found {len(toks)}")
        return toks

    if float(nxt(1)[0]) != 1.0:
        raise ValueError("unsupported file version")
    runs = []
    for _ in range(int(nxt(1)[0])):
        name, fmt = nxt(2)
        photons = int(nxt(1)[0])
        dz, dr = (float(x) for x in nxt(2))
        nz, nr, na = (int(x) for x in nxt(3))
        nl = int(nxt(1)[0])
        n_above = float(nxt(1)[0])
        layers = [dict(zip(("n", "mua", "mus", "g", "d"), (float(x) for x in nxt(5))))
                  for _ in range(nl)]
        n_below = float(nxt(1)[0])
        runs.append(dict(output_name=name, photons=photons, dz=dz, dr=dr, nz=nz, nr=nr, na=na,
                         n_above=n_above, n_below=n_below, layers=layers))
    return runs


def to_spec(spec: dict) -> OrcSpec:
    s = OrcSpec()
    s.photons = int(spec["photons"])
    s.dz, s.dr = float(spec["dz"]), float(spec["dr"])
    s.nz, s.nr, s.na = int(spec["nz"]), int(spec["nr"]), int(spec.get("na", 1))
    s.n_above, s.n_below = float(spec["n_above"]), float(spec["n_below"])
    s.num_layers = len(spec["layers"])
    for i, L in enumerate(spec["layers"]):
        s.layers[i] = OrcLayer(L["n"], L["mua"], L["mus"], L["g"], L["d"])
    return s


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def run(spec: dict, *, photons: int = 0, batches: int = 10, seed: int = 1, threads: int = 0,
        precision: str = "float64", stream: int = 0, arz: bool = False,
        angles: bool = False, time: tuple[int, float] | None = None) -> dict:
    """Simulate one run with the oracle; returns a result dict in the reference's schema.

    With ``angles`` the exit-angle profiles are added: ``rd_a``/``t_a`` [na] and ``rd_ra``/
    ``t_ra`` [nr, na] (sums over batches), per batch under ``ra_b``/``ta_b``. With ``time =
    (nt, dt)`` the time-of-flight profiles of the escaping weight: ``rd_t``/``t_t`` [nt] in bins
    of dt ns (time = optical path / c), per batch under ``rt_b``/``tt_b``."""
    s = to_spec(spec)
    n_ph = photons or s.photons
    B, L, nz, nr = batches, s.num_layers, s.nz, s.nr
    nt, dt = (int(time[0]), float(time[1])) if time is not None else (0, 0.0)
    o = OrcOpts(seed=seed, batches=batches, threads=threads,
                precision=32 if precision == "float32" else 64, stream=stream, photons=n_ph,
                nt=nt, dt=dt)
    tof = [np.zeros(B * max(nt, 1)), np.zeros(B * max(nt, 1))] if time is not None else [None] * 2
    first = np.zeros(B, np.uint64)
    count = np.zeros(B, np.uint64)
    rd = np.zeros(B)
    tt = np.zeros(B)
    absorbed = np.zeros(B * L)
    events = np.zeros(B, np.uint64)
    anomalies = np.zeros(B, np.uint64)
    bsteps = np.zeros(B, np.uint64)
    az = np.zeros(B * nz)
    rr = np.zeros(B * nr)
    tr = np.zeros(B * nr)
    arz_b = np.zeros(B * nz * nr) if arz else None
    na = max(int(s.na), 1)
    ang = [np.zeros(B * na), np.zeros(B * na), np.zeros(B * nr * na), np.zeros(B * nr * na)] \
        if angles else [None] * 4
    out = OrcOut(first=_ptr(first, C.c_uint64), count=_ptr(count, C.c_uint64),
                 rd=_ptr(rd, C.c_double), tt=_ptr(tt, C.c_double),
                 absorbed=_ptr(absorbed, C.c_double), events=_ptr(events, C.c_uint64),
                 anomalies=_ptr(anomalies, C.c_uint64), boundary_steps=_ptr(bsteps, C.c_uint64),
                 az=_ptr(az, C.c_double), rr=_ptr(rr, C.c_double), tr=_ptr(tr, C.c_double),
                 arz=_ptr(arz_b, C.c_double) if arz else None,
                 ra=_ptr(ang[0], C.c_double) if angles else None,
                 ta=_ptr(ang[1], C.c_double) if angles else None,
                 rra=_ptr(ang[2], C.c_double) if angles else None,
                 tra=_ptr(ang[3], C.c_double) if angles else None,
                 rt=_ptr(tof[0], C.c_double) if time is not None else None,
                 tt_t=_ptr(tof[1], C.c_double) if time is not None else None)
    err = C.create_string_buffer(256)
    if lib().orc_run(C.byref(s), C.byref(o), C.byref(out), err, 256) != 0:
        raise ValueError(err.value.decode())
    az_b, rr_b, tr_b = az.reshape(B, nz), rr.reshape(B, nr), tr.reshape(B, nr)
    # Batch-order merge with plain double adds, as src/cpu/engine.cpp:201-207 does.
    a_z, rd_r, t_r = [0.0] * nz, [0.0] * nr, [0.0] * nr
    for b in range(B):
        for i in range(nz):
            a_z[i] += float(az_b[b, i])
        for i in range(nr):
            rd_r[i] += float(rr_b[b, i])
            t_r[i] += float(tr_b[b, i])
    res = {
        "engine": "oracle", "precision": precision, "seed": seed, "threads": out.threads_used,
        "photons": int(count.sum()), "n_above": s.n_above, "n_below": s.n_below,
        "layers": [dict(L) for L in spec["layers"]],
        "grid": {"dz": s.dz, "dr": s.dr, "nz": nz, "nr": nr},
        "specular": out.specular, "wall_seconds": out.wall_seconds,
        "events": int(events.sum()), "anomalies": int(anomalies.sum()),
        "boundary_steps": int(bsteps.sum()),
        "batches": [
            {"first_photon": int(first[b]), "photons": int(count[b]), "rd": float(rd[b]),
             "tt": float(tt[b]), "absorbed": [float(x) for x in absorbed[b * L:(b + 1) * L]],
             "events": int(events[b]), "anomalies": int(anomalies[b])}
            for b in range(B)
        ],
        "a_z": a_z, "rd_r": rd_r, "t_r": t_r,
        "az_b": az_b, "rr_b": rr_b, "tr_b": tr_b,
    }
    if arz:
        arz3 = arz_b.reshape(B, nz, nr)
        res["a_rz_b"] = arz3
        res["a_rz"] = arz3.sum(axis=0)
    if angles:
        res["ra_b"], res["ta_b"] = ang[0].reshape(B, na), ang[1].reshape(B, na)
        res["rd_a"], res["t_a"] = res["ra_b"].sum(axis=0), res["ta_b"].sum(axis=0)
        res["rd_ra"] = ang[2].reshape(B, nr, na).sum(axis=0)
        res["t_ra"] = ang[3].reshape(B, nr, na).sum(axis=0)
    if time is not None:
        res["rt_b"], res["tt_b"] = tof[0].reshape(B, nt), tof[1].reshape(B, nt)
        res["rd_t"], res["t_t"] = res["rt_b"].sum(axis=0), res["tt_b"].sum(axis=0)
    return res


def stream_u32(seed: int, photon_id: int, n: int, word3: int = 0) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_stream_u32(seed, photon_id, word3, n, _ptr(out, C.c_uint32))
    return out


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


# ------------------------------------------------------------------ the reference engine binary

def ref_available() -> bool:
    return REF_BIN.exists() and os.access(REF_BIN, os.X_OK)


def run_ref(mci_path: str | Path, *, photons: int, batches: int, seed: int = 1, threads: int = 0,
            precision: str = "float64", out_json: str | Path) -> dict:
    """Run the reference CPU engine built from its sources; returns its first run's JSON."""
    import json

    cmd = [str(REF_BIN), "run", str(mci_path), "--photons", str(photons), "--batches",
           str(batches), "--seed", str(seed), "--threads", str(threads), "--precision", precision,
           "--output", str(out_json)]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return json.loads(Path(out_json).read_text())["runs"][0]
