"""Pinning of the oracle's conservative vertical remap (oracle/remap_oracle.c)
before it judges the HIP kernels.

* The reference's own tests restated: tests/test_remap.F90 (identity,
  conservation, constancy, monotonicity, thin and single layers, PPM less
  diffusive than PCM), tests/test_remap_boundary_extrap.F90 (linear exactness
  with the boundary closure, the default's first-order boundary error),
  tests/test_remap_nonuniform.F90 (stretched-column exactness, the
  uniform-source reductions, vanished layers, the precondition predicate),
  tests/test_remap_pqm.F90 (high order, monotone, curvature limiter,
  small-N fallback), tests/test_ocean_ppm_h4_remap.F90 (the nz = 6 golden
  vector, order of accuracy), tests/test_ocean_remap_vanished.F90 (the
  concentration-form filler rules) and tests/test_ocean_remap.F90 (faces,
  precondition scan).
* What the mathematics fixes beyond them: each reconstruction reproduces
  the polynomials its edge estimates are exact for (PPM interior: a monotone
  quadratic on equal cells; PPM_H4 interior: a monotone quadratic on any
  cells; PQM interior: a monotone cubic on any cells), so the overlap
  integral of a remap onto any other grid is the exact cell mean there --
  a dropped term, a wrong weight or a transposed stencil index fails it.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

METHODS = (O.REMAP_PCM, O.REMAP_PLM, O.REMAP_PPM, O.REMAP_PPM_H4, O.REMAP_PQM)
RECON = METHODS[1:]


def _nonuniform(n, amp, seed, total=1.0):
    r = np.random.default_rng(seed).random(n)
    dz = 1.0 + amp * r
    return dz / dz.sum() * total


def _pair(n, amp, s1, s2, total=1.0):
    a = _nonuniform(n, amp, s1, total)
    b = _nonuniform(n, amp, s2, total)
    b[-1] += a.sum() - b.sum()
    return a, b


def _means(f_int, dz):
    """Cell means of a profile with antiderivative f_int over the stack dz
    (z = 0 at the bed)."""
    z = np.concatenate([[0.0], np.cumsum(dz)])
    return (f_int(z[1:]) - f_int(z[:-1])) / dz


# ---- tests/test_remap.F90 ---------------------------------------------------

@pytest.mark.parametrize("method", METHODS)
def test_identity_remap(method):
    dz = np.array([1.0, 2.0, 1.5, 3.0, 2.0, 1.0, 1.2, 0.8, 1.4])
    q = np.array([10.0, 20.0, 30.0, 25.0, 15.0, 18.0, 22.0, 19.0, 21.0])
    for be in (0, 1):
        for nu in (0, 1):
            assert np.max(np.abs(O.remap_column(method, dz, dz, q, be, nu) - q)) < 1e-11


@pytest.mark.parametrize("method", METHODS)
def test_conservation_and_constancy(method):
    for seed in range(6):
        n = 5 + 7 * seed
        a, b = _pair(n, 0.8, seed, 100 + seed, total=500.0)
        q = 30.0 + np.sin(np.arange(n) * 0.9) * 4.0
        for be in (0, 1):
            for nu in (0, 1):
                qn = O.remap_column(method, a, b, q, be, nu)
                assert abs((qn * b).sum() - (q * a).sum()) <= 1e-13 * abs((q * a).sum())
                qc = O.remap_column(method, a, b, np.full(n, 25.0), be, nu)
                assert np.max(np.abs(qc - 25.0)) < 1e-12


@pytest.mark.parametrize("method", RECON)
def test_step_profile_bounded(method):
    dz_old = np.ones(4)
    dz_new = np.array([0.5, 1.5, 1.5, 0.5])
    q = np.array([0.0, 0.0, 35.0, 35.0])
    qn = O.remap_column(method, dz_old, dz_new, q)
    assert abs((qn * dz_new).sum() - (q * dz_old).sum()) < 1e-10
    assert qn.min() >= -1e-10 and qn.max() <= 35.0 + 1e-10
    # monotone profile (test_monotonicity_plm/ppm)
    q6 = np.array([0.0, 5.0, 10.0, 20.0, 35.0, 40.0])
    qn = O.remap_column(method, np.ones(6), np.array([0.5, 1.5, 0.5, 1.5, 0.5, 1.5]), q6)
    assert qn.min() >= -1e-10 and qn.max() <= 40.0 + 1e-10


def test_thin_and_single_layer():
    dz_old = np.array([1.0, 1.0e-14, 2.0, 1.0])
    dz_new = np.array([1.0, 1.0, 1.0, 1.0e-14 + 1.0])
    q = np.array([10.0, 0.0, 20.0, 30.0])
    for m in RECON:
        qn = O.remap_column(m, dz_old, dz_new, q)
        assert abs((qn * dz_new).sum() - (q * dz_old).sum()) < 1e-8
    for m in METHODS:
        assert abs(O.remap_column(m, [5.0], [5.0], [42.0])[0] - 42.0) < 1e-12
    # every reconstructing method is the identity at nz = 1, whatever the target
    for m in RECON:
        assert O.remap_column(m, [5.0], [3.0], [42.0])[0] == 42.0


def test_ppm_less_diffusive_than_pcm():
    dz_old = np.ones(8)
    dz_new = np.array([0.5, 1.5] * 4)
    q = np.sin(np.arange(1, 9) * 0.4)
    err = {}
    for m in (O.REMAP_PCM, O.REMAP_PPM):
        back = O.remap_column(m, dz_new, dz_old, O.remap_column(m, dz_old, dz_new, q))
        err[m] = ((back - q) ** 2).sum()
    assert err[O.REMAP_PPM] <= err[O.REMAP_PCM] + 1e-14


# ---- tests/test_remap_boundary_extrap.F90 -----------------------------------

Q0, DQDZ = 3.5, -0.025
DZ_NEW8 = np.array([5.0, 5.0, 4.0, 4.0, 6.0, 6.0, 3.0, 3.0])
DZ_NU8 = np.array([3.0, 4.0, 5.0, 6.0, 6.0, 5.0, 4.0, 3.0])


def _linear(dz, q0=Q0, g=DQDZ):
    z = np.concatenate([[0.0], np.cumsum(dz)])
    return q0 + g * (z[:-1] + 0.5 * dz)


@pytest.mark.parametrize("method,uniform_src", [(O.REMAP_PLM, True), (O.REMAP_PPM, True), (O.REMAP_PPM_H4, True),
                                                (O.REMAP_PPM_H4, False), (O.REMAP_PQM, True),
                                                (O.REMAP_PQM, False)])
def test_linear_exact_with_boundary_extrapolation(method, uniform_src):
    dz_old = np.full(8, 4.5) if uniform_src else DZ_NU8
    qn = O.remap_column(method, dz_old, DZ_NEW8, _linear(dz_old), bnd_extrap=True)
    assert np.max(np.abs(qn - _linear(DZ_NEW8))) < 1e-12


def test_default_closure_is_first_order_at_the_boundary_only():
    dz_old = np.full(8, 4.5)
    q = O.remap_column(O.REMAP_PPM, dz_old, DZ_NEW8, _linear(dz_old))
    ex = _linear(DZ_NEW8)
    assert max(abs(q[0] - ex[0]), abs(q[-1] - ex[-1])) > 1e-4
    assert np.max(np.abs(q[2:-2] - ex[2:-2])) < 1e-12


def test_extrapolation_knob_pcm_inert_uniform_kept_bounded():
    q = _linear(DZ_NU8)
    assert np.array_equal(O.remap_column(O.REMAP_PCM, DZ_NU8, DZ_NEW8, q, 0),
                          O.remap_column(O.REMAP_PCM, DZ_NU8, DZ_NEW8, q, 1))
    qu = O.remap_column(O.REMAP_PPM, DZ_NU8, DZ_NEW8, np.full(8, 34.75), bnd_extrap=True)
    assert np.max(np.abs(qu - 34.75)) < 1e-13
    q_step = np.array([0.0, 0, 0, 30, 30, 30, 30, 30])
    qn = O.remap_column(O.REMAP_PPM, np.full(8, 4.0), np.array([2.0, 6, 4, 4, 4, 4, 6, 2]), q_step, bnd_extrap=True)
    assert qn.min() >= -30.0 - 1e-10 and qn.max() <= 60.0 + 1e-10


# ---- tests/test_remap_nonuniform.F90 ----------------------------------------

SRC = np.array([11.0, 10, 9, 8, 7, 6, 6, 5, 5, 4, 4, 3])
TGT = np.array([2.0, 3, 4, 4, 6, 7, 8, 8, 9, 9, 9, 9])


@pytest.mark.parametrize("method", [O.REMAP_PLM, O.REMAP_PPM])
def test_stretched_linear_exact_with_both_knobs(method):
    q = O.remap_column(method, SRC, TGT, _linear(SRC, 4.25, -0.037), bnd_extrap=True, nonunif=True)
    assert np.max(np.abs(q - _linear(TGT, 4.25, -0.037))) < 1e-13
    # knob off: first order on the stretched source
    q_off = O.remap_column(method, SRC, TGT, _linear(SRC, 4.25, -0.037), bnd_extrap=True)
    assert np.max(np.abs(q_off - _linear(TGT, 4.25, -0.037))) > 1e-6


@pytest.mark.parametrize("method", [O.REMAP_PPM_H4, O.REMAP_PQM])
def test_h4_pqm_already_exact_on_stretched_source(method):
    q = O.remap_column(method, SRC, TGT, _linear(SRC, 4.25, -0.037), bnd_extrap=True)
    assert np.max(np.abs(q - _linear(TGT, 4.25, -0.037))) < 1e-13


def test_uniform_source_reductions():
    dz_u = np.full(12, 6.5)
    dz_t = TGT * (dz_u.sum() / TGT.sum())
    k = np.arange(1, 13)
    q = 12.0 - 0.7 * k + 0.4 * np.cos(1.7 * k)
    a = O.remap_column(O.REMAP_PPM, dz_u, dz_t, q)
    b = O.remap_column(O.REMAP_PPM, dz_u, dz_t, q, nonunif=True)
    assert np.max(np.abs(a - b)) < 1e-12  # CW84 (1.6) reduces to (7/12, -1/12)
    a = O.remap_column(O.REMAP_PLM, dz_u, dz_t, q)
    b = O.remap_column(O.REMAP_PLM, dz_u, dz_t, q, nonunif=True)
    assert np.max(np.abs(a - b)) > 1e-6  # PLM swaps minmod for the CW84 limiter
    step = np.where(k >= 7, 30.0, 0.0)
    b = O.remap_column(O.REMAP_PLM, dz_u, dz_t, step, nonunif=True)
    assert b.min() >= -1e-12 and b.max() <= 30.0 + 1e-12


def test_vanished_layers_stay_finite_and_conservative():
    dz_o = np.full(12, 50.0)
    dz_o[4], dz_o[5] = 0.0, 1.0e-12
    dz_n = np.full(12, dz_o.sum() / 12)
    q = 10.0 + np.arange(1, 13)
    for m in (O.REMAP_PLM, O.REMAP_PPM):
        qn = O.remap_column(m, dz_o, dz_n, q, bnd_extrap=True, nonunif=True)
        assert np.all(np.isfinite(qn))
        assert abs((qn * dz_n).sum() - (q * dz_o).sum()) < 1e-11 * (q * dz_o).sum()


def test_preconditions_predicate():
    assert O.remap_preconditions_ok(SRC, TGT, 1e-10)
    bad = SRC.copy()
    bad[2], bad[3] = -2.0, SRC[3] + 2.0
    assert not O.remap_preconditions_ok(bad, TGT, 1e-10)
    short = TGT.copy()
    short[-1] -= 1.0
    assert not O.remap_preconditions_ok(SRC, short, 1e-10)
    assert O.remap_preconditions_ok(np.zeros(12), np.zeros(12), 1e-10)


# ---- tests/test_remap_pqm.F90, tests/test_ocean_ppm_h4_remap.F90 ------------

def test_pqm_high_order_and_small_n():
    ratios = []
    for n in (16, 32, 64):
        s, t = _pair(n, 0.4, 11, 97)
        F = lambda z: np.exp(1.5 * z) / 1.5  # noqa: E731
        us, ur = _means(F, s), _means(F, t)
        lo, hi = n // 4, 3 * n // 4
        eq = np.max(np.abs(O.remap_column(O.REMAP_PQM, s, t, us)[lo:hi] - ur[lo:hi]))
        ep = np.max(np.abs(O.remap_column(O.REMAP_PPM, s, t, us)[lo:hi] - ur[lo:hi]))
        assert eq < ep
        ratios.append(eq / ep)
    assert ratios[-1] < ratios[0] and ratios[-1] < 0.1
    for n in (2, 3, 4):
        a, b = _pair(n, 0.5, 42 + n, 424 + n)
        q = np.sin(np.arange(1, n + 1))
        assert np.array_equal(O.remap_column(O.REMAP_PQM, a, b, q), O.remap_column(O.REMAP_PPM, a, b, q))


def test_pqm_monotone_and_curvature_limiter():
    s, t = _pair(30, 0.3, 77, 707)
    zm = np.cumsum(s) - 0.5 * s
    q = np.where(zm < 0.5, 1.0, 5.0)
    qn = O.remap_column(O.REMAP_PQM, s, t, q)
    assert qn.min() >= 1.0 - 1e-9 and qn.max() <= 5.0 + 1e-9
    s, t = _pair(12, 0.5, 131, 919)
    zm = np.cumsum(s) - 0.5 * s
    q = 1.0 + 4.0 / (1.0 + np.exp(-14.0 * (zm - 0.5)))
    qn = O.remap_column(O.REMAP_PQM, s, t, q)
    assert abs((qn * t).sum() - (q * s).sum()) < 1e-13 * (q * s).sum()
    assert qn.min() >= q.min() - 1e-12 and qn.max() <= q.max() + 1e-12
    assert np.all(np.diff(qn) >= -1e-12)  # no dip in the monotone ramp


def test_ppm_h4_golden_nz6():
    """The reference's golden vector (tests/test_ocean_ppm_h4_remap.F90:
    237-256)."""
    dz_old = np.array([0.15, 0.20, 0.18, 0.22, 0.12, 0.13])
    u_old = np.array([1.0, 2.5, 2.0, 3.5, 2.8, 1.5])
    dz_new = np.array([0.13, 0.12, 0.22, 0.18, 0.20, 0.15])
    golden = np.array([1.000000000000, 2.250000000000, 2.227272727273, 3.000000000000, 3.192607815674,
                       1.616522912434])
    assert np.max(np.abs(O.remap_column(O.REMAP_PPM_H4, dz_old, dz_new, u_old) - golden)) < 1e-9


def test_ppm_h4_converges_faster_than_ppm():
    errs = {O.REMAP_PPM_H4: [], O.REMAP_PPM: []}
    ns = (12, 24, 48, 96)
    for n in ns:
        s, t = _pair(n, 0.5, 11, 97)
        F = lambda z: -np.cos(2 * np.pi * z) / (2 * np.pi)  # noqa: E731
        us, ur = _means(F, s), _means(F, t)
        for m in errs:
            errs[m].append(np.sqrt((t * (O.remap_column(m, s, t, us) - ur) ** 2).sum()))
    slope = {m: np.log(e[-1] / e[0]) / np.log(ns[-1] / ns[0]) for m, e in errs.items()}
    assert slope[O.REMAP_PPM_H4] < slope[O.REMAP_PPM] - 0.1


# ---- polynomial exactness of the reconstructions ----------------------------

def _targets_inside(s, t, lo_cell, hi_cell):
    """Indices of target cells lying within source cells lo_cell..hi_cell."""
    zs = np.concatenate([[0.0], np.cumsum(s)])
    zt = np.concatenate([[0.0], np.cumsum(t)])
    return [k for k in range(len(t)) if zt[k] >= zs[lo_cell] and zt[k + 1] <= zs[hi_cell + 1]]


def test_ppm_interior_exact_for_monotone_quadratic_on_equal_cells():
    n = 20
    s = np.full(n, 2.0)
    t = _nonuniform(n, 1.0, 3, total=40.0)
    F = lambda z: 3.0 * z + 0.02 * z ** 3 / 3.0  # noqa: E731  f = 3 + 0.02 z^2
    q = O.remap_column(O.REMAP_PPM, s, t, _means(F, s))
    ks = _targets_inside(s, t, 2, n - 3)
    assert len(ks) > 8
    assert np.max(np.abs(q[ks] - _means(F, t)[ks])) < 1e-11


def test_ppm_h4_interior_exact_for_monotone_quadratic_on_any_cells():
    n = 20
    s, t = _pair(n, 1.5, 8, 9, total=40.0)
    F = lambda z: 3.0 * z + 0.02 * z ** 3 / 3.0  # noqa: E731
    q = O.remap_column(O.REMAP_PPM_H4, s, t, _means(F, s))
    ks = _targets_inside(s, t, 1, n - 2)
    assert len(ks) > 8
    assert np.max(np.abs(q[ks] - _means(F, t)[ks])) < 1e-11
    # and PPM is not (its edge weights assume equal cells)
    qp = O.remap_column(O.REMAP_PPM, s, t, _means(F, s))
    assert np.max(np.abs(qp[ks] - _means(F, t)[ks])) > 1e-8


def test_pqm_interior_exact_for_monotone_cubic_on_any_cells():
    n = 24
    s, t = _pair(n, 1.0, 21, 22, total=30.0)
    F = lambda z: 2.0 * z + 0.01 * z ** 3 / 3.0 + 0.0004 * z ** 4 / 4.0  # noqa: E731  f' > 0
    q = O.remap_column(O.REMAP_PQM, s, t, _means(F, s))
    ks = _targets_inside(s, t, 1, n - 2)
    assert len(ks) > 8
    assert np.max(np.abs(q[ks] - _means(F, t)[ks])) < 1e-10


def test_pqm_end_fit_reproduces_a_cubic():
    dz = np.array([0.7, 1.3, 0.9, 1.6])
    F = lambda z: 1.0 * z - 0.5 * z ** 2 / 2 + 0.3 * z ** 3 / 3 + 0.05 * z ** 4 / 4  # noqa: E731
    c = O.pqm_end_value_h4(dz, _means(F, dz))
    assert abs(c[0] - 1.0) < 1e-12 and abs(c[1] - (-0.5)) < 1e-12  # value and slope at the boundary


# ---- tests/test_ocean_remap_vanished.F90 (concentration form) ---------------

def _content(h, c):
    return float((h * c).sum())


def test_donor_column():
    h = np.array([1e-4, 1e-4, 20.0, 1e-4, 30.0, 1e-4])
    c = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    assert np.array_equal(O.vl_column_conc(h, c), [3.0, 3.0, 3.0, 5.0, 5.0, 5.0])
    assert np.array_equal(O.vl_column_conc(np.full(3, 1e-4), c[:3]), [0.0, 0.0, 0.0])


def test_waking_bed_layer_takes_the_neighbour_concentration():
    h_old = np.array([1.0e-4, 20.0, 20.0, 1.0e-4])
    h_new = np.array([1.758272e-3, 20.0, 20.0, 1.0e-4])
    h_new[1] += h_old.sum() - h_new.sum()
    c = np.full(4, 34.5)
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, h_old, h_new, c)
    assert np.all(np.abs(cn - 34.5) <= 1e-13 * 34.5)
    assert abs(_content(h_new, cn) - _content(h_old, c)) <= 1e-12 * _content(h_old, c)


def test_mixed_fillers_conserved_in_range():
    h_old = np.array([1.0e-4, 12.0, 1.0e-4, 30.0, 25.0, 5.0e-5, 18.0, 1.0e-4])
    h_new = np.array([3.0, 9.0, 1.0e-4, 28.0, 27.0, 1.0e-4, 18.9, 1.0e-4])
    h_new[6] = h_old.sum() - (h_new.sum() - h_new[6])
    c = 34.0 + 0.1 * np.arange(1, 9)
    # fillers hold their neighbour's concentration (fillers_hold_neighbour_c)
    c = O.vl_column_conc(h_old, c)
    cn, cl = O.remap_tracer_column(O.REMAP_PPM, h_old, h_new, c)
    before = _content(h_old, c)
    assert abs(_content(h_new, cn) - before) <= 1e-12 * before
    assert cn.min() >= c.min() and cn.max() <= c.max()
    assert abs(cl.sum()) <= 1e-12 * before


def test_random_filler_columns():
    rng = np.random.default_rng(20260921)
    for trial in range(60):
        nz = 12
        h_old = 1.0 + 40.0 * rng.random(nz)
        h_new = 1.0 + 40.0 * rng.random(nz)
        c = 30.0 + 8.0 * rng.random(nz)
        place = trial % 3
        k0, k1 = [(0, trial % 4), (nz - 1 - trial % 4, nz - 1), (3, 3 + trial % 3)][place]
        for k in range(k0, k1 + 1):
            h_old[k] = 1.0e-4
            h_new[k] = 1.0e-4 if trial % 2 == 0 else 1.0e-3 + rng.random()
        live = h_new > 1.5e-4
        h_new[live] *= 1.0 + (h_old.sum() - h_new.sum()) / h_new[live].sum()
        h_new[-1 if live[-1] else np.flatnonzero(live)[-1]] += h_old.sum() - h_new.sum()
        for m in METHODS:
            for be in (0, 1):
                cn, cl = O.remap_tracer_column(m, h_old, h_new, c, be)
                before = _content(h_old, c)
                # the content is conserved once the bounding's recorded amount is removed
                assert abs(_content(h_new, cn) - cl.sum() - before) <= 1e-12 * before
                assert np.all(np.isfinite(cn))


def test_no_filler_column_equals_the_plain_pipeline_but_recorded_clamp():
    rng = np.random.default_rng(991)
    for _ in range(40):
        h_old = 0.5 + 30.0 * rng.random(10)
        h_new = 0.5 + 30.0 * rng.random(10)
        h_new[-1] += h_old.sum() - h_new.sum()
        if h_new[-1] <= 0.2:
            continue
        c = 25.0 + 12.0 * rng.random(10)
        ref = O.remap_column(O.REMAP_PPM, h_old, h_new, c)
        cn, cl = O.remap_tracer_column(O.REMAP_PPM, h_old, h_new, c)
        same = cl == 0.0
        assert np.array_equal(cn[same], ref[same])
        d = np.abs(h_new * (cn - ref) - cl)
        assert np.all(d[~same] <= 4 * np.finfo(float).eps * np.abs(h_new * cn)[~same])


def test_dead_column_and_uniform_constancy_and_mismatch():
    h = np.full(5, 1.0e-4)
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, h, h, np.full(5, 7.0))
    assert np.all(np.abs(cn - 7.0) <= 1e-13 * 7.0)
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, h, h, np.arange(1.0, 6.0))
    assert cn.min() >= 1.0 - 1e-13 and cn.max() <= 5.0 + 1e-13
    h_old = np.array([1.0e-4, 12.0, 1.0e-4, 30.0, 25.0, 5.0e-5, 18.0, 1.0e-4])
    h_new = np.array([3.0, 9.0, 1.0e-4, 28.0, 27.0, 1.0e-4, 18.9 - 5.0e-5, 1.0e-4])
    h_new[1] += h_old.sum() - h_new.sum()
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, h_old, h_new, np.full(8, 35.0))
    assert np.max(np.abs(cn / 35.0 - 1.0)) <= 1e-13
    h_old = np.array([40.0, 48.0, 48.0, 36.0, 1.0e-4, 1.0e-4])
    h_new = h_old.copy()
    h_new[0] -= 1.72e-7
    dh = h_new.sum() - h_old.sum()
    c = np.array([34.6, 34.5, 34.4, 34.2, 34.2, 34.2])
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, h_old, h_new, c)
    before = _content(h_old, c)
    assert abs((_content(h_new, cn) - before) - 34.2 * dh) <= 64 * np.finfo(float).eps * before


def test_field_budget_and_column_helper():
    nx, ny, nz = 5, 4, 6
    rng = np.random.default_rng(4)
    h_old = np.asfortranarray(1.0 + 20.0 * rng.random((nx, ny, nz)))
    h_new = np.asfortranarray(1.0 + 20.0 * rng.random((nx, ny, nz)))
    h_new[:, :, -1] += h_old.sum(axis=2) - h_new.sum(axis=2)
    c = np.asfortranarray(30.0 + rng.random((nx, ny, nz)))
    bud = np.zeros((nx, ny, nz), order="F")
    cn, cl = O.remap_tracer_field(O.REMAP_PPM, h_old, h_new, c, budget=bud)
    for i in range(nx):
        for j in range(ny):
            ref, rcl = O.remap_tracer_column(O.REMAP_PPM, h_old[i, j], h_new[i, j], c[i, j])
            assert np.array_equal(cn[i, j], ref) and np.array_equal(cl[i, j], rcl)
    assert np.array_equal(bud, h_new * cn - h_old * c)
    # the reference's analytic column helper case (test_ocean_remap.F90:385-420)
    cn, _ = O.remap_tracer_column(O.REMAP_PLM, np.full(4, 10.0), np.array([5.0, 10, 10, 15]), [1.0, 2, 3, 4])
    assert abs((np.array([5.0, 10, 10, 15]) * cn).sum() - 100.0) < 1e-10 and cn[-1] < 4.0


# ---- tests/test_ocean_remap.F90: faces, precondition scan -------------------

def test_faces_uniform_and_momentum():
    nx, ny, nz = 4, 3, 4
    h_old = np.asfortranarray(np.broadcast_to([10.0, 10.0, 10.0, 10.0], (nx, ny, nz)).copy())
    h_new = np.asfortranarray(np.broadcast_to([5.0, 10.0, 10.0, 15.0], (nx, ny, nz)).copy())
    for axis, shp in ((0, (nx + 1, ny, nz)), (1, (nx, ny + 1, nz))):
        u = O.remap_face_velocity(axis, h_old, h_new, np.full(shp, 0.25, order="F"), O.REMAP_PPM)
        assert np.max(np.abs(u - 0.25)) < 1e-10
        u0 = np.asfortranarray(np.broadcast_to(0.05 * np.arange(1, nz + 1), shp).copy())
        u = O.remap_face_velocity(axis, h_old, h_new, u0, O.REMAP_PPM)
        assert np.max(np.abs((u * 5.0 * np.array([1, 2, 2, 3])).sum(axis=2) - (u0 * 10.0).sum(axis=2))) < 1e-12
    # KE rescale: momentum kept, anomaly KE restored where the factor is in range
    rng = np.random.default_rng(7)
    h_old = np.asfortranarray(1.0 + 9.0 * rng.random((nx, ny, nz)))
    h_new = np.asfortranarray(1.0 + 9.0 * rng.random((nx, ny, nz)))
    h_new[:, :, -1] += h_old.sum(axis=2) - h_new.sum(axis=2)
    u0 = np.asfortranarray(rng.normal(0, 0.1, (nx + 1, ny, nz)))
    u1 = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PPM, conserve_ke=True)
    hf_o = 0.5 * (h_old[:-1] + h_old[1:])
    hf_n = 0.5 * (h_new[:-1] + h_new[1:])
    mo, mn = (hf_o * u0[1:-1]).sum(axis=2), (hf_n * u1[1:-1]).sum(axis=2)
    assert np.max(np.abs(mo - mn)) < 1e-12
    ubo, ubn = mo / hf_o.sum(axis=2), mn / hf_n.sum(axis=2)
    ke_o = (0.5 * hf_o * (u0[1:-1] - ubo[..., None]) ** 2).sum(axis=2)
    ke_n = (0.5 * hf_n * (u1[1:-1] - ubn[..., None]) ** 2).sum(axis=2)
    plain = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PPM)
    ke_p = (0.5 * hf_n * (plain[1:-1] - ubn[..., None]) ** 2).sum(axis=2)
    ok = np.sqrt(ke_o / ke_p) < 1.25
    assert ok.any() and np.max(np.abs(ke_n[ok] / ke_o[ok] - 1.0)) < 1e-12


def test_zlevel_faces_drop_closed_layers():
    nx, ny, nz = 3, 2, 4
    h_old = np.asfortranarray(np.broadcast_to([10.0, 10.0, 10.0, 10.0], (nx, ny, nz)).copy())
    h_old[1, :, 0] = 1.0e-4  # a closed bed layer in the middle column
    h_old[1, :, 1] += 10.0 - 1.0e-4
    h_new = h_old.copy(order="F")
    h_new[:, :, 3] += 2.0
    h_new[:, :, 2] -= 2.0
    u0 = np.asfortranarray(np.full((nx + 1, ny, nz), 0.3))
    u = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PPM, zlevel_faces=True)
    assert np.all(u[1:3, :, 0] == 0.0)  # faces beside the closed layer: no water, no momentum
    assert np.max(np.abs(u[:, :, 1:] - 0.3)) < 1e-12


def test_precondition_scan():
    nx, ny, nz = 4, 3, 5
    h = np.asfortranarray(np.full((nx, ny, nz), 10.0))
    hn = h.copy(order="F")
    assert O.remap_scan_preconditions(h, hn, 1e-9) == (0, 0.0, 0.0)
    hn[1, 1, 2] = -3.0
    hn[1, 1, 3] = 23.0
    hn[2, 0, 4] += 1e-6
    n_bad, worst_rel, worst_neg = O.remap_scan_preconditions(h, hn, 1e-9)
    assert n_bad == 2 and worst_neg == -3.0 and abs(worst_rel - 1e-6 / 50.000001) < 1e-15
    hz = np.zeros((nx, ny, nz), order="F")
    assert O.remap_scan_preconditions(hz, hz, 1e-9) == (0, 0.0, 0.0)


@pytest.mark.parametrize("method", RECON)
def test_limited_reconstruction_is_monotone_inside_each_cell(method):
    """Colella-Woodward (and W&A) limiting leaves every cell's profile
    monotone and inside its neighbours' means: a monotone source, remapped
    onto a grid that splits every source cell in ten, gives ten monotone
    sub-cell means per cell, within the means around the cell (the
    overshoot branch moves the offending edge so that the parabola's
    extremum sits on the cell edge)."""
    rng = np.random.default_rng(12)
    for trial in range(20):
        n = 6  # live source cells, then 54 empty ones (same layer count as the target)
        s = np.zeros(10 * n)
        s[:n] = 1.0 + 3.0 * rng.random(n)
        q = np.cumsum(np.exp(4.0 * rng.random(10 * n)))  # steep, uneven increments
        t = np.repeat(s[:n] / 10.0, 10)
        t[-1] += s.sum() - t.sum()
        qn = O.remap_column(method, s, t, q)
        tol = 1e-9 * q[:n + 1].max()
        for j in range(n):
            seg = qn[10 * j:10 * j + 10]
            d = np.diff(seg)
            # monotone, in either direction: the edge clip bounds an edge by
            # the three means around the cell, not by the two beside the edge
            assert np.all(d >= -tol) or np.all(d <= tol), (trial, j)
            lo, hi = q[max(j - 1, 0)], q[j + 1]
            assert seg.min() >= lo - tol and seg.max() <= hi + tol, (trial, j)


@pytest.mark.parametrize("method", [O.REMAP_PPM, O.REMAP_PPM_H4, O.REMAP_PQM])
def test_limited_reconstruction_stays_inside_the_local_range(method):
    """On rough (oscillating) data every cell's limited profile stays inside
    the range of the means of the cell and its two neighbours: sub-cell
    means of a ten-fold refinement never leave it."""
    rng = np.random.default_rng(31)
    for trial in range(30):
        n = 6
        s = np.zeros(10 * n)
        s[:n] = 1.0 + 3.0 * rng.random(n)
        q = rng.normal(0.0, 50.0, 10 * n)
        t = np.repeat(s[:n] / 10.0, 10)
        t[-1] += s.sum() - t.sum()
        qn = O.remap_column(method, s, t, q)
        tol = 1e-9 * np.abs(q[:n + 1]).max()
        for j in range(n):
            seg = qn[10 * j:10 * j + 10]
            nb = q[max(j - 1, 0):j + 2]
            assert seg.min() >= nb.min() - tol and seg.max() <= nb.max() + tol, (trial, j)


def test_filler_keeps_its_own_concentration_and_empty_target_takes_the_cell_below():
    # identity remap with a stale bed filler: every layer keeps its value,
    # the filler too (its range is its donor's shifted by its own excess)
    h = np.array([1.0e-4, 12.0, 9.0, 20.0, 1.0e-4])
    c = np.array([99.0, 35.0, 34.0, 33.0, -5.0])
    for m in METHODS:
        cn, _ = O.remap_tracer_column(m, h, h, c)
        assert np.max(np.abs(cn - c)) < 1e-9, m
    # an empty target takes the mean of the source cell containing its
    # bottom interface (remap_bound_column's fallback)
    cn, cl = O.remap_tracer_column(O.REMAP_PPM, np.array([10.0, 10, 10]), np.array([10.0, 0, 20]), [1.0, 2, 4])
    assert cn[1] == 2.0 and cl[1] == 0.0
    cn, _ = O.remap_tracer_column(O.REMAP_PPM, np.array([10.0, 10, 10]), np.array([20.0, 10, 0]), [1.0, 2, 4])
    assert cn[2] == 4.0  # at the very top: the topmost source cell


def test_field_pipeline_keeps_linear_exactness_under_boundary_extrapolation():
    """The bounding widens a boundary cell's range by its mirror point, so the
    extrapolated line is not clamped (the stream test's
    bnd_extrap_linear_unclamped_conserved)."""
    for m in RECON:
        for src in (np.full(8, 4.5), DZ_NU8):
            if m in (O.REMAP_PLM, O.REMAP_PPM) and src is DZ_NU8:
                continue
            for g in (DQDZ, -DQDZ):  # both trends: each mirror widens one side
                # thin outermost targets sit in the outer halves of the boundary cells
                for tgt in (DZ_NEW8, np.array([1.0, 5, 5, 4, 6, 6, 7, 2])):
                    cn, cl = O.remap_tracer_column(m, src, tgt, _linear(src, Q0, g), bnd_extrap=True)
                    assert np.max(np.abs(cn - _linear(tgt, Q0, g))) < 1e-12, m
                    assert np.max(np.abs(cl)) < 1e-11


def test_zlevel_faces_take_the_overlap_thickness():
    nx, ny, nz = 3, 2, 5
    rng = np.random.default_rng(9)
    h_old = np.asfortranarray(5.0 + 10.0 * rng.random((nx, ny, nz)))
    h_old[0, :, 0] = 1.0e-4   # closed in the old grid, open in the new one
    h_new = np.asfortranarray(h_old * (1.0 + 0.1 * rng.random((nx, ny, nz))))
    h_new[0, :, 0] = 3.0
    h_new[:, :, -1] += h_old.sum(axis=2) - h_new.sum(axis=2)
    u0 = np.asfortranarray(rng.normal(0, 0.2, (nx + 1, ny, nz)))
    u = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PPM, zlevel_faces=True)

    def faces(h):
        f = np.empty((nx + 1, ny, nz))
        f[0], f[-1] = h[0], h[-1]
        f[1:-1] = np.minimum(h[:-1], h[1:])
        return f
    fo, fn = faces(h_old), faces(h_new)
    closed = (fo <= 1.5e-4) | (fn <= 1.5e-4)
    fo[closed] = 0.0
    fn[closed] = 0.0
    assert np.all(u[closed] == 0.0)  # a layer closed on either side gets no momentum
    ok = np.abs(fo.sum(axis=2) - fn.sum(axis=2)) <= 1e-12 * fo.sum(axis=2)
    assert ok.sum() >= 2
    mo, mn = (fo * u0).sum(axis=2), (fn * u).sum(axis=2)
    assert np.max(np.abs(mo - mn)[ok]) < 1e-12


def test_ke_rescale_factor_is_capped():
    nx, ny, nz = 2, 2, 12
    h_old = np.asfortranarray(np.full((nx, ny, nz), 10.0))
    h_new = np.asfortranarray(np.broadcast_to(np.array([5.0, 15.0] * 6), (nx, ny, nz)).copy())
    u0 = np.asfortranarray(np.broadcast_to(0.1 * (-1.0) ** np.arange(nz), (nx + 1, ny, nz)).copy())
    plain = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PCM)
    kec = O.remap_face_velocity(0, h_old, h_new, u0, O.REMAP_PCM, conserve_ke=True)
    hf_o, hf_n = np.full(nz, 10.0), np.array([5.0, 15.0] * 6)

    def ke(u, hf):
        ub = (hf * u).sum() / hf.sum()
        return (0.5 * hf * (u - ub) ** 2).sum(), ub
    k_o, _ = ke(u0[1, 0], hf_o)
    k_p, ub = ke(plain[1, 0], hf_n)
    k_n, ub_n = ke(kec[1, 0], hf_n)
    assert np.sqrt(k_o / k_p) > 1.25  # PCM damps the zig-zag hard
    assert abs(k_n / k_p - 1.25 ** 2) < 1e-12 and abs(ub_n - ub) < 1e-15
