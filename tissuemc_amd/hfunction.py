"""Diffuse reflectance of a semi-infinite, isotropically scattering medium from Chandrasekhar's
H-function (S. Chandrasekhar, *Radiative Transfer*, 1960, ch. V).

This is an independent physics reference for the Monte Carlo engine, used where it applies (one
isotropic layer deep enough to be semi-infinite), as the reference does in
python/tissuemc_analysis/oracle.py:84-105.

For single-scattering albedo w the H-function solves

    H(mu) = 1 / (1 - (w/2) mu Int_0^1 H(t) / (mu + t) dt),

computed here by fixed-point iteration on Gauss-Legendre nodes of (0, 1). In a matched medium a
beam entering at cosine mu0 leaves at cosine mu with density
K(mu, mu0) = w mu H(mu) H(mu0) / (2 (mu + mu0)). A refractive-index mismatch is a boundary
recursion: light reaching the surface from inside at cosine mu leaves with probability
1 - R(mu) (unpolarized Fresnel, total internal reflection included) and otherwise starts again as
a downward beam at the same cosine. With E(mu) the probability that a downward beam at mu
eventually escapes, E = K W (1 - R) + K W diag(R) E, a linear system on the nodes.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SemiInfinite:
    specular: float
    diffuse: float

    @property
    def total(self) -> float:
        return self.specular + self.diffuse


def nodes_01(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes and weights on (0, 1)."""
    x, wt = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, wt / 2.0


def h_nodes(albedo: float, n: int = 200, tol: float = 1e-15) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(mu, weights, H(mu)) on n nodes."""
    if not 0.0 <= albedo < 1.0:
        raise ValueError("the albedo must be in [0, 1): the medium has to absorb")
    mu, wt = nodes_01(n)
    inv_sum = 1.0 / (mu[:, None] + mu[None, :])
    h = np.ones(n)
    for _ in range(200_000):
        h_new = 1.0 / (1.0 - 0.5 * albedo * mu * (inv_sum @ (wt * h)))
        if np.max(np.abs(h_new - h)) < tol:
            return mu, wt, h_new
        h = h_new
    raise RuntimeError("H-function iteration did not converge")


def h_value(m: float, albedo: float, mu: np.ndarray, wt: np.ndarray, h: np.ndarray) -> float:
    """H at any cosine m, from its integral equation and the node values."""
    return float(1.0 / (1.0 - 0.5 * albedo * m * np.sum(wt * h / (m + mu))))


def fresnel_unpolarized(n_from: float, n_to: float, cos_i: np.ndarray) -> np.ndarray:
    """Unpolarized Fresnel reflectance from index n_from into n_to (1 for total reflection)."""
    cos_i = np.asarray(cos_i, dtype=np.float64)
    sin_t = n_from / n_to * np.sqrt(np.clip(1.0 - cos_i ** 2, 0.0, None))
    r = np.ones_like(cos_i)
    ok = sin_t < 1.0
    ct = np.sqrt(1.0 - sin_t[ok] ** 2)
    ci = cos_i[ok]
    rs = (n_from * ci - n_to * ct) / (n_from * ci + n_to * ct)
    rp = (n_from * ct - n_to * ci) / (n_from * ct + n_to * ci)
    r[ok] = 0.5 * (rs ** 2 + rp ** 2)
    return r


def semi_infinite_isotropic(mua: float, mus: float, n_medium: float, n_ambient: float = 1.0,
                            n_nodes: int = 200) -> SemiInfinite:
    """Specular and diffuse reflectance of a normally incident pencil beam."""
    if mua <= 0 or mus < 0:
        raise ValueError("need mua > 0 and mus >= 0")
    albedo = mus / (mua + mus)
    mu, wt, h = h_nodes(albedo, n_nodes)
    h1 = h_value(1.0, albedo, mu, wt, h)
    R = fresnel_unpolarized(n_medium, n_ambient, mu)
    K = albedo * np.outer(h, mu * h) / (2.0 * (mu[:, None] + mu[None, :]))  # K[i, j] = K(mu_j; mu_i)
    KW = K * wt[None, :]
    escape_from = np.linalg.solve(np.eye(len(mu)) - KW * R[None, :], KW @ (1.0 - R))
    # The refracted beam travels at cosine 1 inside the medium.
    k1 = albedo * mu * h * h1 / (2.0 * (mu + 1.0))
    escape = float(np.sum(wt * k1 * ((1.0 - R) + R * escape_from)))
    specular = ((n_medium - n_ambient) / (n_medium + n_ambient)) ** 2
    return SemiInfinite(specular=specular, diffuse=(1.0 - specular) * escape)


def escape_angle_fractions(mua: float, mus: float, na: int, n_quad: int = 32,
                           n_nodes: int = 200) -> np.ndarray:
    """Fraction of a normally incident pencil beam that leaves a matched (n = 1) semi-infinite
    isotropic medium at an angle from the normal in each of na bins of pi / (2 na): the integral
    of K(mu, 1) = w mu H(mu) H(1) / (2 (mu + 1)) over cos of the bin's angles. The bins sum to the
    diffuse reflectance."""
    if mua <= 0 or mus < 0:
        raise ValueError("need mua > 0 and mus >= 0")
    albedo = mus / (mua + mus)
    mu, wt, h = h_nodes(albedo, n_nodes)
    h1 = h_value(1.0, albedo, mu, wt, h)
    da = np.pi / (2.0 * na)
    x, wq = nodes_01(n_quad)
    out = np.zeros(na)
    for i in range(na):
        lo, hi = np.cos((i + 1) * da), np.cos(i * da)
        m = lo + (hi - lo) * x
        hm = np.array([h_value(float(v), albedo, mu, wt, h) for v in m])
        out[i] = (hi - lo) * float(np.sum(wq * albedo * m * hm * h1 / (2.0 * (m + 1.0))))
    return out

