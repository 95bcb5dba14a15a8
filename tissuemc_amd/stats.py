"""Batch statistics and the pass rule for comparisons with reference values.

Restates python/tissuemc_analysis/stats.py of the reference (batch_estimate :19-34, compare
:48-64, estimated_reference_se :67-72): a check passes if |ours - reference| <= 3 * combined
standard error + the rounding allowance of a published value. Adds the per-bin form used for
profiles.
"""

from __future__ import annotations

import math
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

Z_LIMIT = 3.0


@dataclass(frozen=True)
class Estimate:
    mean: float
    se: float  # nan with fewer than two batches


def batch_estimate(sums: Sequence[float], photons: Sequence[int]) -> Estimate:
    """Mean per-photon value and its standard error from per-batch weight sums."""
    if len(sums) != len(photons) or not len(sums):
        raise ValueError("need one photon count per batch and at least one batch")
    mean = math.fsum(sums) / sum(photons)
    n = len(sums)
    if n < 2:
        return Estimate(mean, math.nan)
    per_batch = [s / p for s, p in zip(sums, photons, strict=True)]
    center = math.fsum(per_batch) / n
    ss = math.fsum((x - center) ** 2 for x in per_batch)
    return Estimate(mean, math.sqrt(ss / (n - 1) / n))


@dataclass(frozen=True)
class Comparison:
    ours: float
    se_ours: float
    reference: float
    se_reference: float
    rounding: float
    z: float
    passed: bool


def compare(ours: float, se_ours: float, reference: float, se_reference: float = 0.0,
            rounding: float = 0.0) -> Comparison:
    se = math.hypot(se_ours, se_reference)
    diff = ours - reference
    z = diff / se if se > 0 else (0.0 if diff == 0 else math.copysign(math.inf, diff))
    passed = abs(diff) <= Z_LIMIT * se + rounding
    return Comparison(ours, se_ours, reference, se_reference, rounding, z, passed)


def estimated_reference_se(se_ours: float, photons_ours: int, photons_reference: int) -> float:
    """Standard error of a published value that came without one (same per-photon variance)."""
    return se_ours * math.sqrt(photons_ours / photons_reference)


def profile_estimate(per_batch: np.ndarray, photons: Sequence[int]) -> tuple[np.ndarray, np.ndarray]:
    """Per-bin mean per photon and standard error from a [batches, bins] array of weight sums."""
    pb = np.asarray(per_batch, dtype=np.float64)
    ph = np.asarray(photons, dtype=np.float64)
    mean = pb.sum(axis=0) / ph.sum()
    n = pb.shape[0]
    if n < 2:
        return mean, np.full(pb.shape[1], np.nan)
    per = pb / ph[:, None]
    return mean, per.std(axis=0, ddof=1) / math.sqrt(n)


def profile_z(mean_a: np.ndarray, se_a: np.ndarray, mean_b: np.ndarray,
              se_b: np.ndarray) -> np.ndarray:
    """Per-bin z-scores of a - b with the combined standard error (0 where both are exactly 0)."""
    se = np.hypot(se_a, se_b)
    diff = np.asarray(mean_a) - np.asarray(mean_b)
    z = np.zeros_like(diff)
    nz = se > 0
    z[nz] = diff[nz] / se[nz]
    z[(~nz) & (diff != 0)] = np.inf
    return z
