"""A pure-Python float64 photon walker: test infrastructure for pinning the oracle.

It follows the reference's physics (include/tissuemc/physics.hpp) one photon at a time, with
Python floats (IEEE double, same operation order) and the C library's log/cos/sin/sqrt through
``math``, so on this toolchain it reproduces the reference's float64 engine photon for photon.
It is written independently of the oracle's C code and records every event, so its own binning
of the recorded positions checks the oracle's profile binning:

- ``launch`` physics.hpp:58-66, ``specular_reflectance`` :50-55;
- ``fresnel`` :71-98, ``sample_hg_cos`` :101-113, ``spin`` :116-144;
- ``step_size`` :149-158, ``hit_boundary`` :162-178, ``hop`` :180-184, ``drop`` :188-193;
- ``cross_or_not`` :199-246, ``roulette`` :250-259, ``step_once`` :264-277;
- geometry (layer depths, critical cosines) config.hpp:71-101, computed in double;
- the stream: Philox4x32-10 words for counter (photon id, block) and key seed, taken in order
  (philox.hpp:62-83), each mapped to (x + 0.5) 2^-32 (philox.hpp:127-131). The words come from
  the oracle's Philox, which the Random123 known-answer vectors pin.

Small photon counts only (pure-Python loops): a few hundred photons of the three-layer stack take
about a second.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

import oracle

PI = 3.14159265358979323846
WEIGHT = 1e-4
CHANCE = 0.1
COSZERO = 1.0 - 1e-12
COS90D = 1e-6


class Stream:
    """Uniform draws of one photon: words of the Philox stream (seed, photon id) in order."""

    def __init__(self, seed: int, photon_id: int):
        self.seed, self.pid = seed, photon_id
        self.words = oracle.stream_u32(seed, photon_id, 1024)
        self.pos = 0

    def uniform(self) -> float:
        if self.pos == len(self.words):
            self.words = oracle.stream_u32(self.seed, self.pid, 2 * len(self.words))
        x = int(self.words[self.pos])
        self.pos += 1
        return (x + 0.5) * 2.0 ** -32


@dataclass
class Layer:
    n: float
    mua: float
    mus: float
    g: float
    z0: float
    z1: float
    cos_top: float
    cos_bottom: float


def critical_cos(ni: float, nt: float) -> float:
    return math.sqrt(1.0 - (nt * nt) / (ni * ni)) if ni > nt else 0.0


def geometry(spec: dict) -> list[Layer]:
    out, z = [], 0.0
    layers = spec["layers"]
    for i, s in enumerate(layers):
        n_top = spec["n_above"] if i == 0 else layers[i - 1]["n"]
        n_bot = spec["n_below"] if i + 1 == len(layers) else layers[i + 1]["n"]
        z0 = z
        z += s["d"]
        out.append(Layer(s["n"], s["mua"], s["mus"], s["g"], z0, z, critical_cos(s["n"], n_top),
                         critical_cos(s["n"], n_bot)))
    return out


def fresnel(n1: float, n2: float, ca1: float) -> tuple[float, float]:
    """(reflectance, cos of the transmission angle)."""
    if n1 == n2:
        return 0.0, ca1
    if ca1 > COSZERO:
        r = (n2 - n1) / (n2 + n1)
        return r * r, ca1
    if ca1 < COS90D:
        return 1.0, 0.0
    sa1 = math.sqrt(1.0 - ca1 * ca1)
    sa2 = n1 * sa1 / n2
    if sa2 >= 1.0:
        return 1.0, 0.0
    ca2 = math.sqrt(1.0 - sa2 * sa2)
    cap = ca1 * ca2 - sa1 * sa2
    cam = ca1 * ca2 + sa1 * sa2
    sap = sa1 * ca2 + ca1 * sa2
    sam = sa1 * ca2 - ca1 * sa2
    return 0.5 * sam * sam * (cam * cam + cap * cap) / (sap * sap * cam * cam), ca2


@dataclass
class PhotonRecord:
    """Everything one photon scored."""
    absorptions: list = field(default_factory=list)  # (x, y, z, layer, dwa), after the hop
    pre_hop: list = field(default_factory=list)      # (x, y, z) before that hop
    rd: float = 0.0
    tt: float = 0.0
    exit_xy: tuple | None = None                     # position where it left the stack
    boundary_steps: int = 0


def walk(spec: dict, seed: int, photon_id: int) -> PhotonRecord:
    geo = geometry(spec)
    n0, n1 = spec["n_above"], geo[0].n
    rs = (n0 - n1) / (n0 + n1)
    rng = Stream(seed, photon_id)
    rec = PhotonRecord()
    x = y = z = 0.0
    ux = uy = 0.0
    uz = 1.0
    w = 1.0 - rs * rs
    s = sleft = 0.0
    layer = 0
    nl = len(geo)
    dead = False
    while not dead:
        L = geo[layer]
        mut = L.mua + L.mus
        if sleft == 0.0:
            s = -math.log(rng.uniform()) / mut
        else:
            s = sleft / mut
            sleft = 0.0
        hit = False
        if uz != 0.0:
            dl_b = (L.z1 - z) / uz if uz > 0.0 else (L.z0 - z) / uz
            if s > dl_b:
                sleft = (s - dl_b) * (L.mua + L.mus)
                s = dl_b
                hit = True
        x0, y0, z0 = x, y, z
        x += s * ux
        y += s * uy
        z += s * uz
        if hit:
            rec.boundary_steps += 1
            ni = L.n
            if uz < 0.0:
                nt = spec["n_above"] if layer == 0 else geo[layer - 1].n
                r, uz1 = (1.0, 0.0) if -uz <= L.cos_top else fresnel(ni, nt, -uz)
                if rng.uniform() > r:
                    ux *= ni / nt
                    uy *= ni / nt
                    uz = -uz1
                    if layer == 0:
                        rec.rd += w
                        rec.exit_xy = (x, y)
                        dead = True
                    else:
                        layer -= 1
                else:
                    uz = -uz
            else:
                last = layer == nl - 1
                nt = spec["n_below"] if last else geo[layer + 1].n
                r, uz1 = (1.0, 0.0) if uz <= L.cos_bottom else fresnel(ni, nt, uz)
                if rng.uniform() > r:
                    ux *= ni / nt
                    uy *= ni / nt
                    uz = uz1
                    if last:
                        rec.tt += w
                        rec.exit_xy = (x, y)
                        dead = True
                    else:
                        layer += 1
                else:
                    uz = -uz
        else:
            dwa = w * L.mua / (L.mua + L.mus)
            w -= dwa
            rec.absorptions.append((x, y, z, layer, dwa))
            rec.pre_hop.append((x0, y0, z0))
            # spin
            g = L.g
            xi = rng.uniform()
            if g == 0.0:
                cost = 2.0 * xi - 1.0
            else:
                t = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi)
                cost = min(max((1.0 + g * g - t * t) / (2.0 * g), -1.0), 1.0)
            sint = 1.0 - cost * cost
            sint = math.sqrt(sint) if sint > 0.0 else 0.0
            psi = 2.0 * PI * rng.uniform()
            cosp, sinp = math.cos(psi), math.sin(psi)
            if abs(uz) > COSZERO:
                ux, uy = sint * cosp, sint * sinp
                uz = cost if uz >= 0.0 else -cost
            else:
                temp = math.sqrt(1.0 - uz * uz)
                ux, uy, uz = (sint * (ux * uz * cosp - uy * sinp) / temp + ux * cost,
                              sint * (uy * uz * cosp + ux * sinp) / temp + uy * cost,
                              -sint * cosp * temp + uz * cost)
            inv = 1.0 / math.sqrt(ux * ux + uy * uy + uz * uz)
            ux *= inv
            uy *= inv
            uz *= inv
        if not dead and w < WEIGHT:
            if w == 0.0:
                dead = True
            elif rng.uniform() < CHANCE:
                w /= CHANCE
            else:
                dead = True
    return rec


def bin_index(v: float, width: float, count: int) -> int:
    """Clamped bin: v <= 0 -> 0, beyond the grid -> the last bin (src/cpu/engine.cpp:59-66)."""
    if not v > 0:
        return 0
    i = v / width
    return count - 1 if i >= count else int(i)


def arz_of(records, spec: dict, *, positions: str = "after") -> np.ndarray:
    """A(r, z) of recorded photons: each absorption binned at r = sqrt(x^2 + y^2) and z of the
    chosen position (``after`` the hop, as the engines score; ``before`` it, to show a slip)."""
    nz, nr, dz, dr = spec["nz"], spec["nr"], spec["dz"], spec["dr"]
    out = np.zeros((nz, nr))
    for rec in records:
        pts = rec.absorptions if positions == "after" else \
            [(p[0], p[1], p[2], a[3], a[4]) for p, a in zip(rec.pre_hop, rec.absorptions)]
        for (x, y, z, _layer, dwa) in pts:
            out[bin_index(z, dz, nz), bin_index(math.sqrt(x * x + y * y), dr, nr)] += dwa
    return out
