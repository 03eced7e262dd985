"""Pins of the EPBL oracle (oracle/epbl_oracle.c) -- CPU only.

Restates the reference's tests/test_ocean_epbl.F90 on the oracle (same
columns, forcings and bounds; the step harness is run_epbl_steps :142-187:
surface flux -> EPBL -> tracer diffusion with kd_int) and pins the
position-independent math bit for bit against the reference's golden rows
(tests/test_safe_math_golden.inc, extracted by tests/golden/
extract_reference_vectors.py).

Differences from the reference's setup, each forced by this path's scope:
the budget-closure column runs without penetrating shortwave (the column
engine has no shortwave band); every other case is the reference's.
"""

from __future__ import annotations

import json
import math
import struct

import numpy as np
import pytest

from conftest import ROOT

G = 9.80665
RHO0 = 1035.0
ALPHA_LIN, BETA_LIN = 0.2, 0.8  # tests/test_ocean_epbl.F90:38-40
CP = 3992.0


def _f(h: str) -> float:
    return struct.unpack(">d", bytes.fromhex(h))[0]


def _hx(x: float) -> str:
    return struct.pack(">d", x).hex().upper()


GOLD = json.loads((ROOT / "tests" / "golden" / "safe_math_golden.json").read_text())


def _same(y: float, want: str) -> bool:
    w = _f(want)
    return (math.isnan(y) and math.isnan(w)) or _hx(y) == want


@pytest.mark.parametrize("table,fn", [("G_EXP", "sm_exp"), ("G_LOG", "sm_log"), ("G_ERFC", "sm_erfc")])
def test_safe_math_unary_golden_bits(orc, table, fn):
    rows = GOLD[table]["rows"]
    assert len(rows) == 48
    bad = [r for r in rows if not _same(getattr(orc, fn)(_f(r[0])), r[1])]
    assert not bad, bad[:4]


@pytest.mark.parametrize("table", ["G_POW", "G_POWBIG"])
def test_safe_math_pow_golden_bits(orc, table):
    rows = GOLD[table]["rows"]
    assert len(rows) == 48
    bad = [r for r in rows if not _same(orc.sm_pow(_f(r[0]), _f(r[1])), r[2])]
    assert not bad, bad[:4]


def test_safe_math_identities(orc):
    # exact cases the polynomial must reproduce and the IEEE special values
    assert orc.sm_exp(0.0) == 1.0 and orc.sm_log(1.0) == 0.0
    assert orc.sm_pow(2.0, 10.0) == 1024.0 and orc.sm_pow(-2.0, 3.0) == -8.0
    assert math.isnan(orc.sm_pow(-2.0, 0.5)) and orc.sm_pow(0.0, -1.0) == math.inf
    assert orc.sm_erfc(0.0) == 1.0 and orc.sm_erfc(30.0) == 0.0 and orc.sm_erfc(-30.0) == 2.0
    xs = np.linspace(-30, 30, 241)
    assert max(abs(orc.sm_exp(x) / math.exp(x) - 1) for x in xs) < 4e-16
    for x in (1e-300, 3e-5, 0.7, 1.0 + 2**-40, 10.0, 1e300):
        assert abs(orc.sm_log(x) - math.log(x)) <= 2e-16 * max(1.0, abs(math.log(x)))
        lh, ll = orc.sm_log_dd(x)
        assert lh + ll == pytest.approx(math.log(x), rel=1e-16, abs=1e-300)
    for x in np.linspace(-5, 5, 81):
        assert orc.sm_erfc(x) == pytest.approx(math.erfc(x), rel=3e-15, abs=1e-300)


def _sv_wright(orc, T, S, p):
    rho = orc.eos_wright(np.full((1, 1, 1), 1.0), np.full((1, 1, 1), S), np.full((1, 1, 1), T), RHO0, p)
    return 1.0 / rho[0, 0, 0]


def test_specvol_derivs_match_finite_differences(orc):
    # tests/test_ocean_epbl.F90:201-242 (Wright, three (T, S, p) points)
    for T, S, p in ((2.0, 33.0, 0.0), (10.0, 35.0, 2.0e6), (25.0, 37.5, 4.0e7)):
        dt, ds = orc.specvol_derivs(T, S, p)
        e = 1.0e-4
        fd_t = (_sv_wright(orc, T + e, S, p) - _sv_wright(orc, T - e, S, p)) / (2 * e)
        fd_s = (_sv_wright(orc, T, S + e, p) - _sv_wright(orc, T, S - e, p)) / (2 * e)
        assert abs(dt - fd_t) < 1e-7 * abs(fd_t)
        assert abs(ds - fd_s) < 1e-7 * abs(fd_s)
        assert dt > 0 and ds < 0


def _epbl_cfg(orc, **kw):
    c = orc.default_config()
    c.epbl_enable = 1
    c.epbl_mstar_scheme = 1  # constant, 1.2 (setup_epbl :77-92)
    c.epbl_mstar = 1.2
    c.epbl_nstar = 0.2
    c.epbl_tke_diags = 1
    c.alpha_T, c.beta_S, c.rho0 = ALPHA_LIN, BETA_LIN, RHO0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_find_mstar_closed_forms(orc):
    # tests/test_ocean_epbl.F90:244-279
    USTAR, ABSF, BLD = 0.01, 1.0e-4, 50.0
    c = orc.default_config()
    c.epbl_mstar_scheme = 2
    exp_n = min(1.25, 0.085 * math.log(USTAR / (ABSF * BLD)))
    assert abs(orc.epbl_find_mstar(c, 0.0, USTAR, BLD, ABSF) - exp_n) < 1e-14
    b0 = 1.0e-6
    exp_s = 0.3 * math.sqrt(b0 / (USTAR**2 * ABSF))
    assert abs(orc.epbl_find_mstar(c, b0, USTAR, BLD, ABSF) - max(exp_s, exp_n)) < 1e-14
    c.epbl_mstar_cap = 0.5
    assert orc.epbl_find_mstar(c, b0, USTAR, BLD, ABSF) == 0.5
    c.epbl_mstar_cap = -1.0
    # Ekman limit inactive when u* <= |f| bld
    assert orc.epbl_find_mstar(c, 0.0, 1e-3, BLD, ABSF) == 0.0
    c.epbl_mstar_scheme = 1
    assert orc.epbl_find_mstar(c, b0, USTAR, BLD, ABSF) == 1.2
    # RH18 fits (Reichl & Hallberg 2018 eq. fits, rdb_ocean_epbl.F90:665-671)
    c.epbl_mstar_scheme = 3
    for b0 in (0.0, 1e-7, 1e-6):
        t = 8.0 * math.exp(-5.0 * BLD * ABSF / USTAR)
        want = 0.275 * t / (1 + t) + 0.2 * (b0**2 * BLD / (USTAR**5 * ABSF)) ** 0.4
        assert orc.epbl_find_mstar(c, b0, USTAR, BLD, ABSF) == pytest.approx(want, rel=1e-14)
    # convective reduction: (1-adj) at u* -> 0 limit, blend otherwise
    c.epbl_mstar_scheme = 1
    c.epbl_mstar_conv_adj = 0.5
    t1, t2 = -BLD * -1e-7, 2 * 1.2 * USTAR**3
    assert orc.epbl_find_mstar(c, -1e-7, USTAR, BLD, ABSF) == pytest.approx(1.2 * (0.5 * t1 + t2) / (t1 + t2), rel=1e-15)


def test_mixlen_shape(orc):
    assert orc.epbl_mixlen_shape(0.0, 50.0, 0.1, 2.0) == 1.0
    assert orc.epbl_mixlen_shape(50.0, 50.0, 0.1, 2.0) == 0.1
    assert orc.epbl_mixlen_shape(25.0, 50.0, 0.1, 2.0) == pytest.approx(0.1 + 0.9 * 0.25, rel=1e-15)
    assert orc.epbl_mixlen_shape(25.0, 50.0, 0.1, 3.0) == pytest.approx(0.1 + 0.9 * 0.125, rel=1e-15)
    assert orc.epbl_mixlen_shape(25.0, 50.0, 0.1, 2.0, shaped=False) == 1.0


def test_lf17_spot_values(orc):
    # tests/test_ocean_epbl.F90:604-628
    for us, bld, la_ref, u10_ref in zip((0.005, 0.01, 0.01, 0.02), (20.0, 20.0, 50.0, 50.0),
                                        (0.4129, 0.3627, 0.4372, 0.3947), (5.08, 8.51, 8.51, 14.01)):
        u10, ustokes, kphil = orc.epbl_lf17_wave_state(us, RHO0)
        assert abs(u10 - u10_ref) < 0.02 * u10_ref
        la = orc.epbl_lf17_la(us, 0.04 * bld, ustokes, kphil)
        assert abs(la - la_ref) < 0.01 * la_ref
    # the small-argument branch of the surface-layer Stokes average agrees with
    # the erfc form where they meet
    u10, ustokes, kphil = orc.epbl_lf17_wave_state(0.01, RHO0)
    zs = 1e-3 ** 2 / (2 * kphil)
    a = orc.epbl_lf17_la(0.01, zs * 0.999, ustokes, kphil)
    b = orc.epbl_lf17_la(0.01, zs * 1.001, ustokes, kphil)
    assert a == pytest.approx(b, rel=1e-4)


def test_lt_enhance(orc):
    # tests/test_ocean_epbl.F90:630-660
    c = orc.default_config()
    c.epbl_lt_lac1 = c.epbl_lt_lac4 = c.epbl_lt_lac5 = 0.0
    hi = orc.epbl_lt_enhance(c, 0.6, 0.0, 0.01, 50.0, 0.0, 1.0)
    lo = orc.epbl_lt_enhance(c, 0.3, 0.0, 0.01, 50.0, 0.0, 1.0)
    tiny = orc.epbl_lt_enhance(c, 0.01, 0.0, 0.01, 50.0, 0.0, 1.0)
    assert hi > 1.0 and lo > hi and abs(tiny - 5.0) < 1e-12
    assert abs(lo - (1 + 0.447 * 0.3 ** -1.33)) < 1e-12
    c.epbl_lt_scheme = 2
    assert orc.epbl_lt_enhance(c, 0.3, 0.0, 0.01, 50.0, 0.0, 1.0) == pytest.approx(1 + 0.447 * 0.3 ** -1.33, rel=1e-14)
    # stability modification: La_mod = La (1 + max(-1/2, lac1 MLD/Ek) + lac4 Ek/Ob + lac2 MLD/Ob)
    c = orc.default_config()
    la, b0, us, bld, f = 0.3, 2e-8, 0.01, 50.0, 1e-4
    ek, eo, mo = bld * f / us, abs(b0 * 0.41) / (f * us * us), abs(bld * b0 * 0.41) / us**3
    la_mod = la * ((1 + max(-0.5, -0.87 * ek)) + 0.95 * eo + 0.0 * mo)
    assert orc.epbl_lt_enhance(c, la, b0, us, bld, f, 1.0) == pytest.approx(min(5, 1 + 0.447 * la_mod ** -1.33),
                                                                              rel=1e-14)


# ---------------------------------------------------------------------------
# Column cases: run_epbl_steps (tests/test_ocean_epbl.F90:142-187)
# ---------------------------------------------------------------------------
class Column:
    def __init__(self, orc, cfg, nz, dz, t_surf, dT_dz, f=0.0):
        self.orc, self.cfg, self.nz = orc, cfg, nz
        self.h = np.full((1, 1, nz), float(dz), order="F")
        zc = (np.arange(nz)[::-1] + 0.5) * dz  # centre depth of layer k (0 = bed)
        self.T = np.asfortranarray((t_surf - dT_dz * zc).reshape(1, 1, nz))
        self.S = np.full((1, 1, nz), 35.0, order="F")
        self.f = f
        self.mld = 0.0
        self.wet = np.ones((1, 1), order="F")

    def step(self, dt, tau, Q):
        o = self.orc
        o.surface_flux(self.cfg, dt, self.T, self.S, self.h, self.wet, np.full((1, 1), float(Q), order="F"),
                       np.zeros((1, 1), order="F"))
        r = o.epbl_column(self.cfg, dt, self.h.ravel(), self.T.ravel(), self.S.ravel(), tau, Q, 0.0, 1.0, self.f,
                          self.mld, eos_linear=True)
        self.mld = r["mld"]
        kd = np.asfortranarray(r["kd_int"].reshape(1, 1, self.nz + 1))
        o.vdiff_tracers(dt, kd, kd, self.h, self.T, self.S)
        return r


def _run(col, dt, n, tau, Q, every):
    mlds, ts, r = [], [], None
    for s in range(1, n + 1):
        r = col.step(dt, tau, Q)
        if s % every == 0:
            mlds.append(r["mld"])
            ts.append(s * dt)
    return np.array(mlds), np.array(ts), r


def test_ctke_seed(orc):
    # :281-341
    cfg = _epbl_cfg(orc)
    col = Column(orc, cfg, 8, 5.0, 15.0, 0.05)
    r = col.step(1800.0, 0.0, -200.0)
    q = -200.0 / (RHO0 * CP)
    b0 = G * ALPHA_LIN * q / RHO0
    ctke = -0.5 * RHO0 * 5.0 * b0 * 1800.0
    assert ctke > 0
    assert abs(r["tke"]["conv"] - 0.2 * ctke / 1800.0) < 1e-10 * 0.2 * ctke / 1800.0
    assert r["b0"] == pytest.approx(b0, rel=1e-13)
    assert abs(r["kd_int"][0]) < 1e-30 and abs(r["kd_int"][8]) < 1e-30


def test_wind_deepening_kraus_turner(orc):
    # :343-407
    cfg = _epbl_cfg(orc)
    N2, us = 1.0e-4, 0.01
    col = Column(orc, cfg, 100, 3.0, 15.0, N2 * RHO0 / (G * ALPHA_LIN))
    mld, t, _ = _run(col, 1800.0, 240, RHO0 * us**2, 0.0, 12)
    assert np.all(np.diff(mld) >= -1e-9)
    x, y = np.log(t[9:]), np.log(mld[9:])
    slope = np.sum((x - x.mean()) * (y - y.mean())) / np.sum((x - x.mean()) ** 2)
    assert 0.20 < slope < 0.47
    d = (12 * 1.2 * us**3 * t[-1] / N2) ** (1 / 3)
    assert 0.9 < mld[-1] / d < 1.8


def test_budget_ledger_closes(orc):
    # :409-485 (no penetrating shortwave)
    cfg = _epbl_cfg(orc)
    col = Column(orc, cfg, 40, 5.0, 15.0, 0.0527, f=1e-4)
    _, _, r = _run(col, 1800.0, 12, 0.1, -300.0, 1)
    t = r["tke"]
    res = t["wind"] + t["conv"] + t["forcing"] - t["mixing"] - t["mech_decay"] - t["conv_decay"]
    assert abs(res) < 1e-9 * (t["wind"] + t["conv"] + abs(t["forcing"]) + 1e-30)
    assert t["wind"] > 0 and t["conv"] > 0
    # a heating surface spends wind TKE on the forcing term instead
    col = Column(orc, cfg, 40, 5.0, 15.0, 0.0527, f=1e-4)
    _, _, r = _run(col, 1800.0, 12, 0.1, 300.0, 1)
    t = r["tke"]
    res = t["wind"] + t["conv"] + t["forcing"] - t["mixing"] - t["mech_decay"] - t["conv_decay"]
    assert t["forcing"] < 0 and t["conv"] == 0.0
    assert abs(res) < 1e-9 * (t["wind"] + abs(t["forcing"]))


def test_stable_forcing_shoals(orc):
    # :487-534
    cfg = _epbl_cfg(orc)
    dT = 1.0e-4 * RHO0 / (G * ALPHA_LIN)
    a, _, _ = _run(Column(orc, cfg, 60, 3.0, 15.0, dT), 1800.0, 96, RHO0 * 0.01**2, 0.0, 24)
    b, _, _ = _run(Column(orc, cfg, 60, 3.0, 15.0, dT), 1800.0, 96, RHO0 * 0.01**2, 500.0, 24)
    assert b[3] < 0.95 * a[3]


def test_convective_deepening(orc):
    # :536-598
    cfg = _epbl_cfg(orc)
    N2 = 1.0e-4
    col = Column(orc, cfg, 60, 2.5, 15.0, N2 * RHO0 / (G * ALPHA_LIN))
    mld, t, r = _run(col, 1800.0, 96, 0.0, -200.0, 12)
    assert np.all(np.diff(mld) >= -1e-9)
    b0 = abs(G * ALPHA_LIN * (-200.0 / (RHO0 * CP)) / RHO0)
    ratio = mld[-1] / (math.sqrt(2 * b0 * t[-1]) / math.sqrt(N2))
    assert 0.5 < ratio < 1.4
    assert r["kd_int"][59] > 1e-6


def test_langmuir_deepens(orc):
    # :662-715
    dT = 1.0e-4 * RHO0 / (G * ALPHA_LIN)
    off = _epbl_cfg(orc)
    on = _epbl_cfg(orc, epbl_use_lt=1, epbl_lt_lac1=0.0, epbl_lt_lac4=0.0, epbl_lt_lac5=0.0)
    a, _, ra = _run(Column(orc, off, 80, 3.0, 15.0, dT), 1800.0, 48, RHO0 * 0.01**2, 0.0, 12)
    b, _, rb = _run(Column(orc, on, 80, 3.0, 15.0, dT), 1800.0, 48, RHO0 * 0.01**2, 0.0, 12)
    assert b[3] > 1.1 * a[3]
    assert 0.2 < rb["la"] < 0.6 and ra["la"] == 0.0


@pytest.mark.parametrize("nlive,ndead", [(8, 1), (8, 2), (1, 1)])
def test_collapsed_top_layers_equal_column_without_them(orc, nlive, ndead):
    # :738-824
    cfg = _epbl_cfg(orc)
    ref = Column(orc, cfg, nlive, 5.0, 15.0, 0.05)
    r0 = orc.epbl_column(cfg, 1800.0, ref.h.ravel(), ref.T.ravel(), ref.S.ravel(), 0.05, -200.0, 0.0, 1.0, 0.0,
                         eos_linear=True)
    h = np.concatenate([ref.h.ravel(), np.full(ndead, 1e-5)])
    T = np.concatenate([ref.T.ravel(), np.full(ndead, 40.0)])
    S = np.concatenate([ref.S.ravel(), np.full(ndead, 35.0)])
    r1 = orc.epbl_column(cfg, 1800.0, h, T, S, 0.05, -200.0, 0.0, 1.0, 0.0, eos_linear=True)
    c0, c1 = r0["tke"]["conv"], r1["tke"]["conv"]
    assert c0 > 0 and abs(c1 - c0) < 1e-10 * c0
    assert abs(r1["tke"]["wind"] - r0["tke"]["wind"]) < 1e-10 * r0["tke"]["wind"]
    if nlive >= 2:
        scale = np.max(np.abs(r0["kd_int"][1:nlive]))
        assert scale > 0
        assert np.max(np.abs(r1["kd_int"][1:nlive] - r0["kd_int"][1:nlive])) < 1e-9 * scale
    assert abs(r1["mld"] - r0["mld"]) < 1e-6
    assert np.max(np.abs(r1["kd_int"][nlive:])) < 1e-30


def test_dry_column_and_mld_iteration_knobs(orc):
    cfg = _epbl_cfg(orc)
    col = Column(orc, cfg, 20, 5.0, 15.0, 0.05)
    r = orc.epbl_column(cfg, 1800.0, col.h.ravel(), col.T.ravel(), col.S.ravel(), 0.1, -100.0, 0.0, 0.0, 1e-4,
                        mld_prev=7.0, eos_linear=True)
    assert r["mld"] == 0.0 and not np.any(r["kd_int"]) and r["la"] == 0.0
    # without the MLD iteration one pass is taken; with it the converged MLD is
    # a fixed point of the sweep to within mld_tol
    args = (1800.0, col.h.ravel(), col.T.ravel(), col.S.ravel(), 0.1, -100.0, 0.0, 1.0, 1e-4)
    rb = orc.epbl_column(_epbl_cfg(orc, epbl_mld_bisection=1, epbl_mld_tol=1e-3), *args, eos_linear=True)
    rf = orc.epbl_column(_epbl_cfg(orc, epbl_mld_tol=1e-3), *args, eos_linear=True)
    assert rb["mld"] == pytest.approx(rf["mld"], abs=5e-3)
    r1 = orc.epbl_column(_epbl_cfg(orc, epbl_mld_iteration=0), *args, eos_linear=True)
    assert r1["mld"] > 0 and r1["mld"] != rf["mld"]
    # the previous MLD as first guess converges to the same root
    rp = orc.epbl_column(_epbl_cfg(orc, epbl_mld_use_prev_guess=1, epbl_mld_tol=1e-3), *args,
                         mld_prev=rf["mld"] * 1.3, eos_linear=True)
    assert rp["mld"] == pytest.approx(rf["mld"], abs=5e-3)


# ---------------------------------------------------------------------------
# Pins that the reference's range tests leave open
# ---------------------------------------------------------------------------
def test_lt_enhance_stability_branches(orc):
    # lac4/lac2 under a stabilising flux, lac5/lac3 under a destabilising one
    c = orc.default_config()
    c.epbl_lt_lac2, c.epbl_lt_lac3, c.epbl_lt_lac4, c.epbl_lt_lac5 = 0.1, 0.2, 0.3, 0.7
    la, us, bld, f = 0.3, 0.01, 50.0, 1e-4
    for b0, lo, ek in ((2e-8, 0.1, 0.3), (-2e-8, 0.2, 0.7)):
        eo, mo = abs(b0 * 0.41) / (f * us * us), abs(bld * b0 * 0.41) / us**3
        la_mod = la * ((1 + max(-0.5, -0.87 * bld * f / us)) + ek * eo + lo * mo)
        assert orc.epbl_lt_enhance(c, la, b0, us, bld, f, 1.0) == pytest.approx(min(5, 1 + 0.447 * la_mod ** -1.33),
                                                                                  rel=1e-14)


def test_collapsed_top_layers_are_bit_identical(orc):
    # the live part of the column is computed exactly as without the dead
    # layers: the pressure stack starts at the first live layer
    cfg = _epbl_cfg(orc)
    ref = Column(orc, cfg, 12, 5.0, 15.0, 0.05)
    for lin in (True, False):
        r0 = orc.epbl_column(cfg, 1800.0, ref.h.ravel(), ref.T.ravel(), ref.S.ravel(), 0.05, -200.0, 0.0, 1.0, 1e-4,
                             eos_linear=lin)
        h = np.concatenate([ref.h.ravel(), [1.4e-4, 1e-5]])
        T = np.concatenate([ref.T.ravel(), [40.0, 40.0]])
        S = np.concatenate([ref.S.ravel(), [35.0, 35.0]])
        r1 = orc.epbl_column(cfg, 1800.0, h, T, S, 0.05, -200.0, 0.0, 1.0, 1e-4, eos_linear=lin)
        assert np.array_equal(r1["kd_int"][:13], r0["kd_int"]) and r1["mld"] == r0["mld"]


@pytest.mark.parametrize("iterate", [0, 1])
def test_homogeneous_column_kd_profile(orc, iterate):
    # An unstratified column costs no energy to mix (PE change 0), so every
    # interface takes the unlimited diffusivity Kd = v* kappa l with the
    # mechanical TKE decayed as exp(-z |f| tke_decay / u*) over the layers
    # above, v* = (TKE / (dt rho0))^(1/3), l = htot hbs v* / (|f| htot hbs + v*)
    # and hbs = (H - z) / H: no MLD iteration, or one whose shape function
    # stays above (H - z) / H (transition-layer floor 0.9999).
    cfg = _epbl_cfg(orc, epbl_mld_iteration=iterate, epbl_translay_scale=0.9999)
    nz, dz, dt, f, tau = 30, 4.0, 1800.0, 1.2e-4, 0.08
    h = np.full(nz, dz)
    T = np.full(nz, 12.0)
    S = np.full(nz, 35.0)
    r = orc.epbl_column(cfg, dt, h, T, S, tau, 0.0, 0.0, 1.0, f, eos_linear=True)
    us = math.sqrt(tau / RHO0)
    H = nz * dz
    m0 = 1.2 * dt * RHO0 * us**3
    for i in range(1, nz):
        z = i * dz
        tke = m0 * math.exp(-z * 2.5 * f / us)
        vs = (tke / (dt * RHO0)) ** (1 / 3)
        hbs = (H - z) / H
        l = z * hbs * vs / (f * z * hbs + vs)
        assert r["kd_int"][nz - i] == pytest.approx(vs * 0.41 * l, rel=1e-9)  # round-off PE of the eliminated column
    # the whole column is reached: MLD is the full depth
    assert r["mld"] == pytest.approx(H, abs=1.0 if iterate else 1e-9)


def _brute_energy(orc, h, T, S, kd, dt, rho0=RHO0):
    """Energy the mixing of kd_int (bed-first, nz+1) costs, interface by
    interface from the surface down, from direct solves of the implicit
    diffusion of the column (no elimination shortcuts): the change of the
    pressure-weighted PE sum(dPE/dT dT + dPE/dS dS), plus, where mixing
    shrinks the column, the work against the pressure at that interface."""
    nz = h.size
    live = np.nonzero(h > 1.5e-4)[0]
    ks = int(live[-1]) if live.size else nz - 1
    hh = h + 1e-20
    t0 = np.where(h > 0, T, 0.0)
    s0 = np.where(h > 0, S, 0.0)
    dpe_t, dpe_s, dch_t, dch_s = (np.zeros(nz) for _ in range(4))
    pres, h_sum = 0.0, 0.0
    for k in range(nz - 1, -1, -1):
        dmass = rho0 * h[k]
        dpres = G * dmass
        pm = pres + 0.5 * dpres
        a, b = orc.specvol_derivs(t0[k], s0[k], pm)
        dpe_t[k], dpe_s[k], dch_t[k], dch_s[k] = dmass * pm * a, dmass * pm * b, dmass * a, dmass * b
        if k <= ks:
            pres += dpres
            h_sum += h[k]
    h_sum += 1e-20

    def solve(kd_used):
        A = np.diag(hh.copy())
        for ki, kv in kd_used.items():
            ka, kb = ki, ki - 1
            kdd = kv * dt / max(0.5 * (hh[ka] + hh[kb]), 1e-15 * h_sum)
            A[ka, ka] += kdd
            A[kb, kb] += kdd
            A[ka, kb] -= kdd
            A[kb, ka] -= kdd
        Tn = np.linalg.solve(A, hh * t0)
        Sn = np.linalg.solve(A, hh * s0)
        return (np.sum(dpe_t * (Tn - t0) + dpe_s * (Sn - s0)), np.sum(dch_t * (Tn - t0) + dch_s * (Sn - s0)))

    used, total, scale = {}, 0.0, 0.0
    pe_prev, ch_prev = 0.0, 0.0
    z = h[ks]
    for ki in range(ks, 0, -1):
        used[ki] = kd[ki]
        pe, ch = solve(used)
        e = pe - pe_prev
        if ch - ch_prev < 0:
            e -= G * rho0 * z * (ch - ch_prev)
        total += e
        scale += abs(e)
        pe_prev, ch_prev = pe, ch
        z += h[ki - 1]
    return total, scale


@pytest.mark.parametrize("case", ["wind_stable", "wind_partial", "interior_inversion", "cooling"])
def test_energy_solve_matches_direct_implicit_diffusion(orc, case):
    # Wright EOS (nonlinear, pressure dependent: cabbeling shrinks the
    # column, so the column-height work term is live).  The TKE the sweep
    # books as spent on mixing plus the PE it books as released by
    # convective interfaces must equal the energy of the final diffusivities
    # measured by direct solves.
    cfg = _epbl_cfg(orc)
    nz, dz, dt = 40, 5.0, 1800.0
    zc = (np.arange(nz)[::-1] + 0.5) * dz
    T = 22.0 - 0.12 * zc
    S = 34.5 + 0.004 * zc
    tau, Q, f = 0.1, 0.0, 1e-4
    if case == "wind_partial":
        tau = 0.02
    if case == "interior_inversion":
        T = T.copy()
        T[nz - 8] = T[nz - 7] + 1.5  # warm water under cold: statically unstable
        tau = 0.05
    if case == "cooling":
        Q = -400.0
    h = np.full(nz, dz)
    r = orc.epbl_column(cfg, dt, h, T, S, tau, Q, 0.0, 1.0, f)
    total, scale = _brute_energy(orc, h, T, S, r["kd_int"], dt)
    t = r["tke"]
    nstar = 0.2
    ctke_pos = 0.0
    if Q < 0:
        a, b = orc.specvol_derivs(T[-1], S[-1], 0.5 * G * RHO0 * dz)
        ctke_pos = -(0.5 * G * RHO0**2 * dz * ((1 / (RHO0 * CP)) * Q * dt * a))
    released = (nstar * ctke_pos - t["conv"] * dt) / nstar
    booked = t["mixing"] * dt + released
    assert scale > 0
    assert booked == pytest.approx(total, rel=1e-9, abs=1e-9 * scale)
    if case == "interior_inversion":
        assert released < 0  # PE released by mixing the unstable interface
    if case == "cooling":
        assert released == 0.0 and t["conv"] > 0  # the seed itself, no unstable interface below
