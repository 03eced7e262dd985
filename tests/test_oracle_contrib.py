"""Pins of the oracle's KV_ML_INVZ2, convective adjustment and double
diffusion (oracle/epbl_oracle.c) -- CPU only.

Restates the reference's tests/test_ocean_kv_ml_invz2.F90,
tests/test_ocean_convection.F90 and tests/test_ocean_double_diffusion.F90
on the oracle (same columns, same bounds and known answers), plus the
closed forms those leave open.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

G = 9.80665


def _col(a):
    """A fresh (1, 1, n) Fortran-ordered copy of a profile."""
    a = np.array(a, dtype=np.float64).ravel()
    return np.array(a.reshape(1, 1, a.size), order="F", copy=True)


# ---------------------------------------------------------------------------
# KV_ML_INVZ2 (tests/test_ocean_kv_ml_invz2.F90)
# ---------------------------------------------------------------------------
def _invz2(orc, nz, h_each, kv0, hmix):
    c = orc.default_config()
    c.kv_ml_invz2, c.hmix_fixed = kv0, hmix
    return orc.vmix_invz2(c, _col(np.full(nz, h_each)), np.zeros((1, 1, nz + 1), order="F"))[0, 0]


def test_invz2_zero_amplitude_is_noop(orc):
    assert not np.any(_invz2(orc, 4, 5.0, 0.0, 20.0))
    assert not np.any(_invz2(orc, 4, 5.0, 0.01, 0.0))


def test_invz2_interface_below_hmix_unchanged(orc):
    assert not np.any(_invz2(orc, 2, 100.0, 0.01, 20.0))


def test_invz2_profile_decreases_with_depth(orc):
    kv = _invz2(orc, 50, 1.0, 0.01, 20.0)
    # Fortran interfaces nz and nz-9 are 0-based nz-1 and nz-10
    assert kv[49] > kv[40] > 0.0


def test_invz2_value_is_the_inverse_square_profile(orc):
    kv = _invz2(orc, 10, 5.0, 0.01, 20.0)
    assert abs(kv[7] - 0.01 * (20.0 / 15.0) ** 2) < 1e-14
    # every interface shallower than hmix, none at or below it, the surface
    # and the bed untouched; the floor dz_min = hmix / (2 nz) near the top
    want = np.zeros(11)
    for k in range(9, 0, -1):
        z = (10 - k) * 5.0
        if z >= 20.0:
            break
        want[k] = 0.01 * 400.0 / max(z, 1.0) ** 2
    assert np.allclose(kv, want, rtol=1e-15, atol=0)
    kv = _invz2(orc, 40, 0.01, 0.01, 20.0)  # thin layers: z < dz_min = 0.25 m
    assert kv[39] == pytest.approx(0.01 * 400.0 / 0.25**2, rel=1e-15)


# ---------------------------------------------------------------------------
# Convective adjustment (tests/test_ocean_convection.F90)
# ---------------------------------------------------------------------------
def _conv_run(orc, rho, h, bld, on=True, kd=1.0, pr=1.0, thresh=0.0):
    c = orc.default_config()
    c.conv_enable, c.conv_kd, c.conv_prandtl, c.conv_n2_thresh = int(on), kd, pr, thresh
    nz = len(rho)
    hh, rr = _col(h), _col(rho)
    zero = np.zeros_like(hh)
    kv, kt = orc.pp81(c, zero, zero, hh, rr)
    if on:
        kv, kt = orc.vmix_convection(c, hh, rr, np.full((1, 1), float(bld), order="F"), kv, kt)
    ks = kt.copy(order="F")
    orc.assemble(c, kv, kt, ks)
    return kv[0, 0], kt[0, 0]


def test_conv_disabled_and_stable_bit_identical(orc):
    rho_unst = 1030.0 + np.arange(4.0)
    a = _conv_run(orc, rho_unst, np.full(4, 10.0), 0.0, on=False)
    b = _conv_run(orc, rho_unst, np.full(4, 10.0), 0.0, on=False)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    rho_st = 1030.0 - np.arange(4.0)
    a = _conv_run(orc, rho_st, np.full(4, 10.0), 0.0, on=True)
    b = _conv_run(orc, rho_st, np.full(4, 10.0), 0.0, on=False)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_conv_unstable_sets_kd_conv(orc):
    kv, kt = _conv_run(orc, 1030.0 + np.arange(4.0), np.full(4, 10.0), 0.0, kd=1.0, pr=2.0)
    assert np.all(np.abs(kt[1:4] - 1.0) < 1e-12) and np.all(np.abs(kv[1:4] - 2.0) < 1e-12)
    assert kt[0] == kv[0] == kt[4] == kv[4] == 0.0


def test_conv_below_bl_only(orc):
    kv, kt = _conv_run(orc, 1028.0 + np.arange(10.0), np.full(10, 20.0), 100.0)
    # Fortran interfaces 7..10 (0-based 6..9) lie inside the 100 m layer
    assert np.all(np.abs(kt[6:10] - (1e-5 + 1e-2)) < 1e-10)
    assert np.all(np.abs(kt[1:6] - 1.0) < 1e-10)


def test_conv_n2_thresh_live(orc):
    drho = 5.0e-6 * 1035.0 * 10.0 / G
    rho = 1030.0 - drho * np.arange(4.0)
    for thresh, want in ((0.0, 1e-5), (1e-6, 1e-5), (1e-5, 1.0)):
        _, kt = _conv_run(orc, rho, np.full(4, 10.0), 0.0, thresh=thresh)
        assert abs(kt[1] - want) < 1e-9


def test_conv_homogenises_inversion(orc):
    nz, dt = 20, 1200.0
    T0 = 1028.0 + 2.0 * np.arange(nz) / (nz - 1)
    res = {}
    for on in (True, False):
        c = orc.default_config()
        c.conv_enable = int(on)
        h, T, S = _col(np.full(nz, 10.0)), _col(T0), _col(np.full(nz, 35.0))
        zero = np.zeros_like(h)
        kv, kt = orc.pp81(c, zero, zero, h, T)
        if on:
            kv, kt = orc.vmix_convection(c, h, T, np.zeros((1, 1), order="F"), kv, kt)
        ks = kt.copy(order="F")
        orc.assemble(c, kv, kt, ks)
        for _ in range(7):
            orc.vdiff_tracers(dt, kt, kt, h, T, S)
        res[on] = T[0, 0]
    assert np.ptp(res[True]) < 0.6 and np.ptp(res[False]) > 1.5
    for on in (True, False):
        assert abs(res[on].sum() * 10.0 - T0.sum() * 10.0) < 1e-9 * T0.sum() * 10.0


def test_conv_vanished_interface_skipped(orc):
    h = np.array([10.0, 1e-6, 1e-6, 10.0])
    rho = np.array([1028.0, 1028.0, 1040.0, 1040.0])
    kv, kt = _conv_run(orc, rho, h, 0.0)
    assert math.isfinite(kt[2]) and abs(kt[2] - (1e-5 + 1e-2)) < 1e-9


def test_conv_depth_mask_and_n2_closed_form(orc):
    # z is the depth of the interface's upper layer bottom; N^2 from the
    # centred face thickness; floors, never lowers
    c = orc.default_config()
    c.conv_enable, c.conv_kd, c.conv_prandtl, c.conv_n2_thresh = 1, 0.5, 3.0, 1e-6
    h = np.array([8.0, 6.0, 4.0, 2.0])
    rho = np.array([1027.0, 1027.0 - 6e-4, 1027.0 + 1e-3, 1026.0])
    kv0 = np.full(5, 0.7)
    kt0 = np.full(5, 0.2)
    kv, kt = orc.vmix_convection(c, _col(h), _col(rho), np.full((1, 1), 2.5, order="F"), _col(kv0), _col(kt0))
    kv, kt = kv[0, 0], kt[0, 0]
    for k in range(3, 0, -1):
        z = h[k:].sum()
        n2 = -G * (rho[k] - rho[k - 1]) / (1035.0 * 0.5 * (h[k - 1] + h[k]))
        fire = z >= 2.5 and n2 < 1e-6
        assert kt[k] == (max(0.2, 0.5) if fire else 0.2)
        assert kv[k] == (max(0.7, 1.5) if fire else 0.7)
    assert kt[3] == 0.2  # z = 2 m < 2.5 m: inside the boundary layer


# ---------------------------------------------------------------------------
# Double diffusion (tests/test_ocean_double_diffusion.F90)
# ---------------------------------------------------------------------------
K_S_FINGER, K_T_FINGER = 5.2448726125e-5, 3.6714108288e-5
K_T_CONV, K_S_CONV = 5.2441171668e-5, 2.3336321392e-5


def _ddiff(orc, T, S, kt0=0.0, **kw):
    c = orc.default_config()
    c.ddiff_enable = 1
    for k, v in kw.items():
        setattr(c, k, v)
    nz = len(T)
    kt = np.asfortranarray(np.broadcast_to(np.asarray(kt0, dtype=np.float64), (nz + 1,)).reshape(1, 1, nz + 1))
    kt2, ks = orc.vmix_split_ddiff(c, _col(np.full(nz, 10.0)), _col(T), _col(S), kt)
    return kt2[0, 0], ks[0, 0]


def _close(a, b, rtol=1e-7):
    return abs(a - b) <= rtol * abs(b) + 1e-18


def test_ddiff_fingering_exact(orc):
    kt, ks = _ddiff(orc, [5.0, 7.0, 9.0], [34.0, 34.3441295547, 34.6882591094])
    for k in (1, 2):
        assert _close(ks[k], K_S_FINGER) and _close(kt[k], K_T_FINGER) and ks[k] > kt[k]
    assert kt[0] == ks[0] == kt[3] == ks[3] == 0.0


def test_ddiff_convection_exact(orc):
    kt, ks = _ddiff(orc, [9.0, 7.0, 5.0], [35.0, 34.3609022556, 33.7218045112])
    for k in (1, 2):
        assert _close(kt[k], K_T_CONV) and _close(ks[k], K_S_CONV) and kt[k] > ks[k]


def test_ddiff_above_cutoff_and_doubly_stable(orc):
    dS = 1.7e-4 * 2.0 / (7.6e-4 * 3.0)
    kt, ks = _ddiff(orc, [5.0, 7.0, 9.0], [34.0, 34.0 + dS, 34.0 + 2 * dS])
    assert np.max(np.abs(ks)) < 1e-18 and np.max(np.abs(kt)) < 1e-18
    kt, ks = _ddiff(orc, [5.0, 7.0, 9.0], [35.0, 34.5, 34.0])
    assert np.array_equal(kt, ks) and np.max(np.abs(kt)) < 1e-18


def test_ddiff_disabled_is_plain_split(orc):
    c = orc.default_config()
    kt = _col(1e-3 * np.arange(1, 5))
    ks = kt.copy(order="F")
    assert np.array_equal(ks, kt)  # the disabled path is orc_split (ks := kt)


def test_ddiff_closed_forms(orc):
    # Large et al. (1994) fingering: K_S = kappa_s (1 - ((R-1)/(Rmax-1))^e1)^e2,
    # K_T = 0.7 K_S; Marmorino-Caldwell (1976) convection K_T = nu 0.909
    # exp(4.6 exp(-0.54 (1/R - 1))), K_S = K_T (0.15 R | 1.85 R - 0.85);
    # Kelley (1990) K_T = nu 8.7 R^1.1.  Background kt adds to both.
    aT, bS = 1.7e-4, 7.6e-4
    for R in (1.2, 1.8, 2.5):
        dS = aT * 2.0 / (bS * R)
        for e1, e2 in ((1.0, 3.0), (1.5, 2.0)):
            kt, ks = _ddiff(orc, [5.0, 7.0], [34.0, 34.0 + dS], kt0=1e-5, ddiff_exp1=e1, ddiff_exp2=e2)
            Rr = (aT * 2.0) / (bS * dS)
            want_s = 1e-4 * (1 - ((Rr - 1) / 1.55) ** e1) ** e2
            assert ks[1] == pytest.approx(1e-5 + want_s, rel=1e-13)
            assert kt[1] == pytest.approx(1e-5 + 0.7 * want_s, rel=1e-13)
    for R in (0.3, 0.7):  # adT / bdS with both negative (warm, salty above cold, fresh... inverted)
        dT = -2.0
        dS = aT * dT / (bS * R)
        for k90 in (0, 1):
            kt, ks = _ddiff(orc, [5.0, 5.0 + dT], [34.0, 34.0 + dS], ddiff_use_k90=k90)
            Rr = (aT * dT) / (bS * dS)
            dd = 1.5e-6 * 8.7 * Rr**1.1 if k90 else 1.5e-6 * 0.909 * math.exp(4.6 * math.exp(-0.54 * (1 / Rr - 1)))
            assert kt[1] == pytest.approx(dd, rel=1e-13)
            assert ks[1] == pytest.approx((0.15 * Rr if Rr < 0.5 else 1.85 * Rr - 0.85) * dd, rel=1e-13)
    # a vanished layer on either side switches the interface off
    kt, ks = _ddiff(orc, [5.0, 7.0], [34.0, 34.2], kt0=1e-5)
    c = orc.default_config()
    c.ddiff_enable = 1
    kt2, ks2 = orc.vmix_split_ddiff(c, _col([10.0, 1e-4]), _col([5.0, 7.0]), _col([34.0, 34.2]),
                                    _col(np.full(3, 1e-5)))
    assert ks[1] != 1e-5 and ks2[0, 0, 1] == 1e-5 and kt2[0, 0, 1] == 1e-5
