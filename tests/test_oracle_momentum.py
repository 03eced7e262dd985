"""Pinning of the momentum vertical-viscosity oracle (oracle/mom_oracle.c,
the restatement of diffuse_velocity_columns_impl,
rdb_ocean_vdiff.F90:1416-1926) before it is trusted as the HIP kernel's
checker.

* the reference's own tests, restated: constancy and the K_v = 0 no-op
  (tests/test_ocean_vdiff.F90:368-444), implicit stability (:446-492), the
  no-drag remnant of exactly 1 (tests/test_ocean_visc_rem.F90), stress
  placement, sign and land masking and the thin-layer drag bound
  (tests/test_ocean_vdiff_implicit_stress_drag.F90), the upwind-blend gate
  (tests/test_ocean_vdiff_bbl.F90:357-430);
* closed forms: a single layer under drag, top drag and wind, and the
  constant-piston glue;
* mathematics: every face column solved as a dense linear system built
  independently in numpy (walls, thin and vanished layers, z-level cuts,
  k_top/k_bot rows, stress/drag/top-drag folds, harmonic face thickness and
  both hvel_mom6 branches); the remnant equals the response of the momentum
  solve to a uniform acceleration (the operator is linear); and the
  flux form conserves sum(hf * u) when no boundary term acts.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

HV = 1.5e-4


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _grid(nx=6, ny=4, nz=5, seed=0, zlevel=False):
    rng = np.random.default_rng(seed)
    h = rng.uniform(2.0, 40.0, (nx, ny, nz))
    if zlevel:
        # bed fillers of uneven depth, and a few thin live layers
        nfill = rng.integers(0, nz - 1, (nx, ny))
        for i in range(nx):
            for j in range(ny):
                h[i, j, :nfill[i, j]] = 1e-4
        h[rng.random(h.shape) < 0.05] = 5e-3
    u = rng.normal(0, 0.2, (nx + 1, ny, nz))
    v = rng.normal(0, 0.2, (nx, ny + 1, nz))
    kv = rng.uniform(1e-5, 5e-2, (nx, ny, nz + 1))
    kv[:, :, 0] = 0.0
    kv[:, :, nz] = 0.0
    wet = np.ones((nx, ny))
    return [np.asfortranarray(a) for a in (h, u, v, kv, wet)]


# ---------------------------------------------------------- reference tests
def test_constancy_and_zero_kv():
    nx, ny, nz = 6, 4, 5
    h = np.full((nx, ny, nz), 10.0, order="F")
    u = np.full((nx + 1, ny, nz), 0.4, order="F")
    v = np.full((nx, ny + 1, nz), -0.3, order="F")
    wet = np.ones((nx, ny), order="F")
    u2, v2, _, _ = O.vdiff_apply_momentum(h, u, v, wet, 60.0, K_v_momentum=1e-2)
    assert np.abs(u2 - 0.4).max() < 1e-12 and np.abs(v2 + 0.3).max() < 1e-12
    k = np.arange(1, nz + 1)
    u = np.asfortranarray(np.broadcast_to(0.1 * k, (nx + 1, ny, nz)).copy())
    v = np.asfortranarray(np.broadcast_to(-0.05 * k, (nx, ny + 1, nz)).copy())
    u2, v2, _, _ = O.vdiff_apply_momentum(h, u, v, wet, 60.0, K_v_momentum=0.0)
    assert np.array_equal(u2, u) and np.array_equal(v2, v)


def test_implicit_stability():
    nx, ny, nz = 4, 4, 6
    h = np.ones((nx, ny, nz), order="F")
    u = np.ones((nx + 1, ny, nz), order="F")
    u[:, :, nz // 2:] = -1.0
    v = np.zeros((nx, ny + 1, nz), order="F")
    u2, _, _, _ = O.vdiff_apply_momentum(h, u, v, np.ones((nx, ny)), 1.0, K_v_momentum=1000.0)
    assert np.all(np.isfinite(u2)) and np.abs(u2).max() <= 1.0 + 1e-6
    assert np.abs(u2).max() < 0.5 * 2.0


def test_no_drag_remnant_is_one():
    nx, ny, nz = 4, 4, 5
    h = np.asfortranarray(np.broadcast_to(17.0 * np.arange(1, nz + 1), (nx, ny, nz)).copy())
    z = lambda s: np.zeros(s, order="F")
    _, _, ru, rv = O.vdiff_apply_momentum(h, z((nx + 1, ny, nz)), z((nx, ny + 1, nz)), np.ones((nx, ny)), 900.0,
                                          K_v_momentum=3.5, remnant=True)
    assert max(np.abs(ru - 1).max(), np.abs(rv - 1).max()) < 1e-12
    # the remnant is clamped to 1 as each level is finalised: with random
    # layer thicknesses the unclamped back substitution exceeds 1 by an ulp
    rng = np.random.default_rng(1)
    nx, ny, nz = 40, 30, 12
    h = np.asfortranarray(rng.uniform(0.5, 80.0, (nx, ny, nz)))
    _, _, ru, rv = O.vdiff_apply_momentum(h, z((nx + 1, ny, nz)), z((nx, ny + 1, nz)), np.ones((nx, ny)), 900.0,
                                          K_v_momentum=3.5, remnant=True)
    assert ru.max() <= 1.0 and rv.max() <= 1.0 and ru.min() > 1.0 - 1e-12


def test_stress_placement_sign_and_land_mask():
    nx, ny, nz = 6, 4, 4
    h = np.full((nx, ny, nz), 25.0, order="F")
    u = np.zeros((nx + 1, ny, nz), order="F")
    v = np.zeros((nx, ny + 1, nz), order="F")
    wet = np.ones((nx, ny), order="F")
    wet[2, 1] = 0.0
    tau_u = np.full((nx + 1, ny), 0.1, order="F")
    tau_v = np.zeros((nx, ny + 1), order="F")
    dt, rho0 = 600.0, 1035.0
    u2, _, _, _ = O.vdiff_apply_momentum(h, u, v, wet, dt, K_v_momentum=0.0, implicit_stress=True, tau_u=tau_u,
                                         tau_v=tau_v, rho0=rho0)
    want = dt * 0.1 / rho0 / 25.0
    top = u2[:, :, nz - 1]
    # faces touching the land cell (u faces i = 2 and 3 of row 1) take no stress
    mask = np.ones((nx + 1, ny), dtype=bool)
    mask[2, 1] = mask[3, 1] = False
    assert np.allclose(top[mask], want, rtol=1e-14, atol=0)
    assert np.all(top[~mask] == 0.0)
    assert np.all(u2[:, :, :nz - 1] == 0.0)  # K_v = 0: the bed is untouched
    u3, _, _, _ = O.vdiff_apply_momentum(h, u, v, wet, dt, K_v_momentum=1e-2, implicit_stress=True, tau_u=tau_u,
                                         tau_v=tau_v, rho0=rho0)
    col = u3[0, 0, :]
    assert np.all(np.diff(col) > 0) and col[0] > 0  # diffused downward, surface gains most
    with pytest.raises(ValueError):
        O.vdiff_apply_momentum(h, u, v, wet, dt, implicit_stress=True, tau_u=tau_u, tau_v=tau_v)


def test_thin_layer_drag_bounded_and_single_layer_closed_forms():
    nx, ny, nz = 3, 3, 1
    h = np.full((nx, ny, nz), 0.5, order="F")
    u = np.full((nx + 1, ny, nz), 0.8, order="F")
    v = np.full((nx, ny + 1, nz), -0.6, order="F")
    dt = 3600.0
    lam_u = np.full((nx + 1, ny), 6.0 / dt, order="F")  # dt * lambda = 6: explicit Euler flips
    lam_v = np.full((nx, ny + 1), 6.0 / dt, order="F")
    uu = u
    for _ in range(5):
        un, vn, ru, rv = O.vdiff_apply_momentum(h, uu, v, np.ones((nx, ny)), dt, implicit_drag=True,
                                                lambda_bot_u=lam_u, lambda_bot_v=lam_v, remnant=True)
        assert np.all(un > 0) and np.all(un < uu)
        assert np.allclose(un, uu / 7.0, rtol=1e-15)
        uu = un
    assert np.allclose(ru, 1.0 / 7.0, rtol=1e-15)
    # top drag and wind on one layer: u' = (u + dt tau (1 - cover) / (rho0 h)) / (1 + dt lambda_top)
    cov_u = np.full((nx + 1, ny), 0.25, order="F")
    tau_u = np.full((nx + 1, ny), 0.2, order="F")
    un, _, _, _ = O.vdiff_apply_momentum(h, u, v, np.ones((nx, ny)), dt, implicit_stress=True, tau_u=tau_u,
                                         tau_v=np.zeros((nx, ny + 1)), rho0=1035.0, implicit_top_drag=True,
                                         lambda_top_u=lam_u, lambda_top_v=lam_v, cover_u=cov_u,
                                         cover_v=np.zeros((nx, ny + 1)))
    want = (0.8 + dt * 0.2 * 0.75 / 1035.0 / 0.5) / 7.0
    assert np.allclose(un, want, rtol=1e-14)
    # constant-piston glue on one layer: b = 1 + dt kv_bbl / (h (min(h/2, hbbl) + eps))
    _, _, ru, _ = O.vdiff_apply_momentum(h, u, v, np.ones((nx, ny)), dt, hvel_mom6=True, bbl_glue=True,
                                         remnant=True, hbbl_visc=10.0, bbl_piston=3e-4)
    b = 1.0 + dt * 3e-4 * 10.0 / (0.5 * (min(0.25, 10.0) + 1e-30))
    assert np.allclose(ru, 1.0 / b, rtol=1e-14)


def test_upwind_gate():
    nx, ny, nz = 6, 4, 3

    def run(uniform, upwind):
        h = np.zeros((nx, ny, nz), order="F")
        h[:, :, 1], h[:, :, 2] = 12.0, 20.0
        for i in range(nx):
            h[i, :, 0] = 8.0 if uniform else max(0.5, 16.0 - 2.0 * (i + 1))
        u = np.asfortranarray(np.broadcast_to(np.array([0.4, 0.1, -0.2]), (nx + 1, ny, nz)).copy())
        v = np.zeros((nx, ny + 1, nz), order="F")
        return O.vdiff_apply_momentum(h, u, v, np.ones((nx, ny)), 300.0, K_v_momentum=0.3, hvel_mom6=True,
                                      hvel_harmonic=True, hvel_upwind=upwind)[0]

    assert np.array_equal(run(True, True), run(True, False))
    assert np.abs(run(False, True) - run(False, False)).max() > 1e-12


# ------------------------------------------------------------ dense solves
def _hvel_column(hl, hr, uc, kb, mode, d_min=None, i_hbbl=0.1):
    """Face thickness per layer of one face column, from the layer
    thicknesses of its two cells: arithmetic mean, z-level min, or the two
    hvel_mom6 blends (harmonic with the thick-to-thin upwind blend; the
    z_clear arithmetic/harmonic blend)."""
    nz = len(hl)
    out = np.zeros(nz)
    zacc = 0.0
    zl, zr = -hl.sum(), -hr.sum()
    dmin = min(hl.sum(), hr.sum())
    for k in range(nz):
        harm = 2.0 * hl[k] * hr[k] / (hl[k] + hr[k] + 1e-30)
        arith = 0.5 * (hl[k] + hr[k])
        if mode == "mom6_harm":
            bot = 1.0 / (1.0 + 0.09 * zacc ** 6)
            out[k] = (1 - bot) * harm + bot * arith if uc[k] * (hr[k] - hl[k]) < 0 else harm
            if k >= kb:
                zacc += harm * i_hbbl
        elif mode == "mom6_zclear":
            zl += hl[k]
            zr += hr[k]
            if k >= kb:
                zacc += harm
            zi = max(zacc, max(zl, zr) + dmin) * i_hbbl
            bot = 1.0 / (1.0 + 0.09 * zi ** 6)
            out[k] = (1 - bot) * arith + bot * harm if uc[k] * (hr[k] - hl[k]) > 0 else arith
        elif mode == "zlevel":
            out[k] = min(hl[k], hr[k])
        else:
            out[k] = arith
    return out


def _dense_face_solve(u, hvel, nu_i, dt, kb, ktop, zlevel, harmonic, mom6, drag=0.0, topdrag=0.0, wind=0.0):
    """Solve one face column's backward-Euler system as a dense matrix:
    rows kb_live..ktop couple through dt * nu / (hf_k * dz) at each
    interface (none across a z-level cut), every other row is the identity;
    the folds add drag to the bed row's diagonal, top drag to the top row's
    diagonal and wind to the top row's right-hand side."""
    nz = len(u)
    hf = np.maximum(hvel, HV)
    kbl = kb
    if zlevel:
        live = [k for k in range(kb, nz) if hvel[k] > HV]
        kbl = live[0] if live else kb
    A = np.eye(nz)
    r = u.copy()

    def coup(k):  # interface between layers k and k + 1
        if mom6:
            dz = 0.5 * (hf[k] + hf[k + 1])
        elif harmonic:
            dz = 2 * hf[k] * hf[k + 1] / (hf[k] + hf[k + 1])
        else:
            dz = 0.5 * (hf[k] + hf[k + 1])
        cut = zlevel and (hvel[k] <= HV or hvel[k + 1] <= HV)
        return 0.0 if cut else dt * nu_i[k + 1] / dz

    for k in range(nz):
        if kbl <= k <= ktop:
            if k == kbl and k < nz - 1:
                A[k, k + 1] = -coup(k) / hf[k]
                A[k, k] += coup(k) / hf[k]
            if k > kbl:
                A[k, k - 1] = -coup(k - 1) / hf[k]
                A[k, k] += coup(k - 1) / hf[k]
            if kbl < k < ktop:
                A[k, k + 1] = -coup(k) / hf[k]
                A[k, k] += coup(k) / hf[k]
            if k == kbl:
                A[k, k] += dt * drag
        if k == ktop:  # the top row's folds apply even to an identity row
            A[k, k] += dt * topdrag
            r[k] += wind / max(hf[k], 1e-3)
    return np.linalg.solve(A, r)


@pytest.mark.parametrize("mode", ["default", "harmonic", "zlevel", "mom6_harm", "mom6_zclear"])
def test_oracle_matches_dense_solve(mode):
    nx, ny, nz = 5, 4, 7
    h, u, v, kv, wet = _grid(nx, ny, nz, seed=3, zlevel=(mode == "zlevel"))
    wet[1, 2] = 0.0
    rng = np.random.default_rng(9)
    dt, rho0 = 1800.0, 1035.0
    tau_u = np.asfortranarray(rng.normal(0, 0.1, (nx + 1, ny)))
    tau_v = np.asfortranarray(rng.normal(0, 0.1, (nx, ny + 1)))
    lb_u = np.asfortranarray(rng.uniform(0, 1e-3, (nx + 1, ny)))
    lb_v = np.asfortranarray(rng.uniform(0, 1e-3, (nx, ny + 1)))
    lt_u = np.asfortranarray(rng.uniform(0, 1e-3, (nx + 1, ny)))
    lt_v = np.asfortranarray(rng.uniform(0, 1e-3, (nx, ny + 1)))
    cov_u = np.asfortranarray((rng.random((nx + 1, ny)) < 0.3).astype(float))
    cov_v = np.asfortranarray((rng.random((nx, ny + 1)) < 0.3).astype(float))
    kt_u = np.asfortranarray(rng.integers(nz - 2, nz + 1, (nx + 1, ny)))
    kb_u = np.asfortranarray(rng.integers(1, 3, (nx + 1, ny)))
    kt_v = np.asfortranarray(rng.integers(nz - 2, nz + 1, (nx, ny + 1)))
    kb_v = np.asfortranarray(rng.integers(1, 3, (nx, ny + 1)))
    mom6 = mode.startswith("mom6")
    un, vn, _, _ = O.vdiff_apply_momentum(
        h, u, v, wet, dt, kv_source=kv, use_harmonic=(mode == "harmonic"), implicit_stress=True, tau_u=tau_u,
        tau_v=tau_v, rho0=rho0, implicit_drag=True, lambda_bot_u=lb_u, lambda_bot_v=lb_v, implicit_top_drag=True,
        lambda_top_u=lt_u, lambda_top_v=lt_v, cover_u=cov_u, cover_v=cov_v, zlevel_faces=(mode == "zlevel"),
        hvel_mom6=mom6, hvel_harmonic=(mode == "mom6_harm"), k_top_u=kt_u, k_bot_u=kb_u, k_top_v=kt_v,
        k_bot_v=kb_v)
    for xf, (nu, nv), arr, out, tau, lb, lt, cov, kt, kbb in (
            (True, (nx + 1, ny), u, un, tau_u, lb_u, lt_u, cov_u, kt_u, kb_u),
            (False, (nx, ny + 1), v, vn, tau_v, lb_v, lt_v, cov_v, kt_v, kb_v)):
        for j in range(nv):
            for i in range(nu):
                if xf:
                    cl, cr = (max(i - 1, 0), j), (min(i, nx - 1), j)
                else:
                    cl, cr = (i, max(j - 1, 0)), (i, min(j, ny - 1))
                hl, hr = h[cl], h[cr]
                kb, ktop = int(kbb[i, j]) - 1, int(kt[i, j]) - 1
                hvel = _hvel_column(hl, hr, arr[i, j], kb, mode if mode != "harmonic" else "default")
                nu_i = 0.5 * (kv[cl] + kv[cr])
                wind = dt * tau[i, j] * (1.0 - cov[i, j]) * min(wet[cl], wet[cr]) / rho0
                want = _dense_face_solve(arr[i, j], hvel, nu_i, dt, kb, ktop, mode == "zlevel", mode == "harmonic",
                                         mom6, drag=lb[i, j], topdrag=lt[i, j], wind=wind)
                got = out[i, j]
                assert np.allclose(got, want, rtol=1e-11, atol=1e-14), (mode, xf, i, j, got - want)


def test_remnant_is_the_response_to_uniform_acceleration():
    nx, ny, nz = 4, 3, 6
    h, u, v, kv, wet = _grid(nx, ny, nz, seed=5)
    dt, acc = 900.0, 1e-3
    rng = np.random.default_rng(2)
    lb_u = np.asfortranarray(rng.uniform(0, 1e-3, (nx + 1, ny)))
    lb_v = np.asfortranarray(rng.uniform(0, 1e-3, (nx, ny + 1)))
    kw = dict(kv_source=kv, implicit_drag=True, lambda_bot_u=lb_u, lambda_bot_v=lb_v)
    u0, v0, ru, rv = O.vdiff_apply_momentum(h, u, v, wet, dt, remnant=True, **kw)
    u1, v1, _, _ = O.vdiff_apply_momentum(h, u + dt * acc, v + dt * acc, wet, dt, **kw)
    assert np.allclose((u1 - u0) / (dt * acc), ru, rtol=1e-9, atol=1e-12)
    assert np.allclose((v1 - v0) / (dt * acc), rv, rtol=1e-9, atol=1e-12)
    assert np.all(ru > 0) and np.all(ru <= 1.0)
    # remnant-only mode leaves u alone and gives the same remnant
    u2, v2, ru2, rv2 = O.vdiff_apply_momentum(h, u, v, wet, dt, remnant=True, remnant_only=True, **kw)
    assert np.array_equal(u2, u) and np.array_equal(ru2, ru) and np.array_equal(rv2, rv)


def test_flux_form_conserves_face_momentum():
    nx, ny, nz = 5, 4, 8
    h, u, v, kv, wet = _grid(nx, ny, nz, seed=7)
    un, _, _, _ = O.vdiff_apply_momentum(h, u, v, wet, 3600.0, kv_source=kv)
    for i in range(nx + 1):
        for j in range(ny):
            hf = 0.5 * (h[max(i - 1, 0), j] + h[min(i, nx - 1), j])
            assert abs((hf * un[i, j]).sum() - (hf * u[i, j]).sum()) <= 1e-13 * np.abs(hf * u[i, j]).sum()
