"""The H-function oracle (tissuemc_amd.hfunction), pinned.

- The reference's recorded outputs (results/validation-cpu.json): O1 (matched, albedo 0.9)
  R_d = 0.4149474791308157, V2 (n = 1.5) diffuse 0.2197039010071812 and total
  0.2597039010071812.
- Identities of the integral equation (the cases of the reference's
  tests/python/test_oracle.py:6-73, stated anew): convergence in the number of nodes, the
  absorbed-fraction identity, the matched boundary 1 - sqrt(1 - w) H(1), monotonicity, the
  published Giovanelli total 0.2600, Fresnel limits; and Chandrasekhar's zeroth moment
  Int_0^1 H = (2 / w)(1 - sqrt(1 - w)), which nothing in the solver uses.
"""

import math

import numpy as np
import pytest

from tissuemc_amd import hfunction as hf


def test_recorded_outputs_of_the_reference_oracle():
    o1 = hf.semi_infinite_isotropic(mua=10, mus=90, n_medium=1.0)
    assert o1.diffuse == pytest.approx(0.4149474791308157, rel=1e-14, abs=0)
    v2 = hf.semi_infinite_isotropic(mua=10, mus=90, n_medium=1.5)
    assert v2.diffuse == pytest.approx(0.2197039010071812, rel=1e-14, abs=0)
    assert v2.total == pytest.approx(0.2597039010071812, rel=1e-14, abs=0)


def test_h_converges_with_the_number_of_nodes():
    w = 0.9
    vals = []
    for n in (60, 120, 240):
        mu, wt, h = hf.h_nodes(w, n)
        vals.append(hf.h_value(1.0, w, mu, wt, h))
    assert abs(vals[1] - vals[2]) < 1e-10
    assert abs(vals[0] - vals[2]) < 1e-6


@pytest.mark.parametrize("w", [0.3, 0.9, 0.99])
def test_zeroth_moment(w):
    mu, wt, h = hf.h_nodes(w, 240)
    assert np.sum(wt * h) == pytest.approx(2.0 / w * (1.0 - math.sqrt(1.0 - w)), abs=1e-10)


@pytest.mark.parametrize("mu0", [1.0, 0.5, 0.1])
def test_absorbed_fraction_identity(mu0):
    """The escaping density integrates to 1 - sqrt(1 - w) H(mu0)."""
    w = 0.9
    mu, wt, h = hf.h_nodes(w, 240)
    h0 = hf.h_value(mu0, w, mu, wt, h)
    escaped = np.sum(wt * w * mu * h * h0 / (2.0 * (mu + mu0)))
    assert escaped == pytest.approx(1.0 - math.sqrt(1.0 - w) * h0, abs=1e-9)


def test_matched_boundary():
    w = 0.9
    r = hf.semi_infinite_isotropic(mua=10, mus=90, n_medium=1.0)
    mu, wt, h = hf.h_nodes(w, 200)
    assert r.specular == 0.0 and r.total == r.diffuse
    assert r.diffuse == pytest.approx(1.0 - math.sqrt(1.0 - w) * hf.h_value(1.0, w, mu, wt, h),
                                      abs=1e-9)


def test_mismatch_traps_light():
    matched = hf.semi_infinite_isotropic(10, 90, 1.0)
    mismatched = hf.semi_infinite_isotropic(10, 90, 1.5)
    assert mismatched.specular == pytest.approx(0.04)
    assert mismatched.diffuse < matched.diffuse


def test_reflectance_rises_with_albedo():
    rd = [hf.semi_infinite_isotropic(1 - w, w, 1.0).diffuse for w in (0.5, 0.9, 0.99)]
    assert rd[0] < rd[1] < rd[2] < 1.0


def test_published_giovanelli_total():
    """The MCML manual quotes 0.2600 for n = 1.5, albedo 0.9 (precision not stated)."""
    assert abs(hf.semi_infinite_isotropic(10, 90, 1.5).total - 0.2600) < 5e-4


def test_fresnel_limits():
    r = hf.fresnel_unpolarized(1.5, 1.0, np.array([1.0, 0.7, 0.3]))
    assert r[0] == pytest.approx(0.04)
    assert r[1] == 1.0 and r[2] == 1.0  # beyond the critical angle (cos = 0.745)
    # s/p average against the textbook angle form at 30 degrees from glass into air
    ti = math.radians(30.0)
    tt = math.asin(1.5 * math.sin(ti))
    rs = (math.sin(ti - tt) / math.sin(ti + tt)) ** 2
    rp = (math.tan(ti - tt) / math.tan(ti + tt)) ** 2
    got = hf.fresnel_unpolarized(1.5, 1.0, np.array([math.cos(ti)]))[0]
    assert got == pytest.approx(0.5 * (rs + rp), rel=1e-12)


def test_angle_fractions_sum_to_the_diffuse_reflectance():
    f = hf.escape_angle_fractions(10, 90, 12)
    assert f.sum() == pytest.approx(hf.semi_infinite_isotropic(10, 90, 1.0).diffuse, rel=1e-8)
    assert np.all(f > 0)


def test_rejects_non_absorbing_or_invalid_media():
    with pytest.raises(ValueError):
        hf.h_nodes(1.0)
    with pytest.raises(ValueError):
        hf.semi_infinite_isotropic(mua=0.0, mus=10, n_medium=1.0)
