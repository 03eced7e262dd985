"""Pinning of the oracle's heat/salt_budget_vdiff slots
(apply_factored_tracer's budget += h*c_new - h*c_old,
rdb_ocean_vdiff.F90:1397-1405; rdb_multilayer_state.F90:403-410):

* the reference's own test restated (tests/test_ocean_vdiff.F90:494-566):
  with closed boundaries the contributor totals telescope to round-off and
  equal the change of the heat and salt contents;
* in the composed column step the budget is the content change of the
  tracer solve alone: after the surface deposit (which the budget must not
  count) and before the non-local tendency (likewise), checked against the
  composition restated from the staged oracle routines with a snapshot
  around the solve; per column the budget telescopes to round-off.
"""

from __future__ import annotations

import numpy as np

from oracle import oracle as O
from roundabout_amd import synth

DT = 3600.0


def test_reference_budget_contributor_telescopes():
    nx, ny, nz = 6, 4, 5
    h = np.full((nx, ny, nz), 10.0, order="F")
    k = np.arange(1, nz + 1)
    T = np.asfortranarray(np.broadcast_to(5.0 + 2.0 * k, (nx, ny, nz)).copy())
    S = np.asfortranarray(np.broadcast_to(30.0 + 0.5 * k, (nx, ny, nz)).copy())
    K = np.full((nx, ny, nz + 1), 1.0e-2, order="F")
    K[:, :, 0] = K[:, :, nz] = 0.0
    hb = np.zeros((nx, ny, nz), order="F")
    sb = np.zeros((nx, ny, nz), order="F")
    t0, s0 = (h * T).sum(), (h * S).sum()
    O.vdiff_tracers_budget(60.0, K, K, h, T, S, hb, sb)
    t1, s1 = (h * T).sum(), (h * S).sum()
    assert abs(hb.sum()) / abs(t0) < 1e-12 and abs(sb.sum()) / abs(s0) < 1e-12
    assert abs((t1 - t0) - hb.sum()) < 1e-12 * abs(t0)
    assert abs((s1 - s0) - sb.sum()) < 1e-12 * abs(s0)
    assert np.abs(hb).max() > 0  # the solve did move heat between layers


def _staged_step(cfg, o, smag, shelf, hb, sb):
    """One column step composed from the staged oracle routines in the order
    of apply_deferred_thermo, with the budget formed around the solve."""
    rho = O.eos_wright(o.h, o.S, o.T, cfg.rho0, cfg.p_ref)
    kv, kt = O.pp81(cfg, o.u, o.v, o.h, rho)
    gt, gs, _, _ = O.kpp_overlay(cfg, o.u, o.v, o.h, rho, smag, o.Q_heat, o.Q_salt, o.bl_depth, kv, kt,
                                 stress_shelf=shelf)
    ks = kt.copy(order="F")
    O.assemble(cfg, kv, kt, ks)
    O.surface_flux(cfg, DT, o.T, o.S, o.h, o.wet, o.Q_heat, o.Q_salt, o.k_top)
    T_old, S_old = o.T.copy(), o.S.copy()
    O.vdiff_tracers(DT, kt, ks, o.h, o.T, o.S, o.passive, cfg.use_harmonic)
    hb += o.h * o.T - o.h * T_old
    sb += o.h * o.S - o.h * S_old
    O.nonlocal_tendencies(DT, gt, gs, o.h, o.T, o.S)


def test_column_step_budget_is_the_solve_alone():
    ocean = synth.small(8)
    cfg = O.default_config()
    st = O.OracleState(ocean, cfg, budgets=True)
    st.step(DT, 3)
    o = ocean.copy()
    hb = np.zeros_like(o.h, order="F")
    sb = np.zeros_like(o.h, order="F")
    smag = O.stress_mag(o.tau_x, o.tau_y)
    shelf = np.zeros((o.nx, o.ny), order="F")
    for _ in range(3):
        _staged_step(cfg, o, smag, shelf, hb, sb)
    assert np.array_equal(st.ocean.T, o.T)  # the staged composition is the column step
    assert np.array_equal(st.heat_budget, hb) and np.array_equal(st.salt_budget, sb)
    # per column the solve only moves content between layers
    scale = (o.h * np.abs(o.T)).sum(axis=2)
    assert np.all(np.abs(st.heat_budget.sum(axis=2)) <= 1e-13 * scale)
    scale = (o.h * np.abs(o.S)).sum(axis=2)
    assert np.all(np.abs(st.salt_budget.sum(axis=2)) <= 1e-13 * scale)
    # the surface deposit and the non-local transport are not in it: the
    # step's total content change differs from the budget by the deposit
    assert np.abs(st.heat_budget).max() > 0
