"""Batched ensembles on the GPU: several members of different grids and nz
stepped by one persistent launch per step.

The plain ensemble must reproduce, bit for bit, each member stepped alone
(which the single-grid parity tests tie to the oracle); the guarded learned
closure over an uneven-nz ensemble must conserve heat and salt and is compared
with the all-plain ensemble.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from roundabout_amd import synth
from roundabout_amd.rdb import (CLOSURE_SURROGATE, GUARD_SURFACE, ColumnEnsemble, ColumnPhysics, RdbError,
                                default_config)

pytestmark = pytest.mark.gpu
DT = 3600.0


def members(seed=40):
    return [
        synth.small(seed),                                                        # 64 x 64 x 20 z-star
        synth.tile(seed + 1, nx=56, ny=40),                                       # z-level, nz 75
        synth.make_ocean(48, 36, 50, seed + 2, land_frac=0.3, topo="zlevel", dmin=50.0, dmax=5500.0),
        synth.thin(seed + 3, nx=40, ny=32),                                       # thin layers, nz 75
        synth.make_ocean(24, 20, 50, seed + 4, land_frac=0.3, topo="zlevel", dmin=50.0, dmax=5500.0),
    ]


FIELDS = ("temperature", "salinity", "tracer_1", "tracer_2", "kv", "kt", "ks", "bl_depth", "kbl", "b0")


@pytest.mark.parametrize("nsteps", [1, 24])
def test_ensemble_matches_members_alone_bitwise(gpu_lib, nsteps):
    print(f"[case] ensemble of 5 (nz 20/75/50/75/50), {nsteps} steps, bitwise vs members alone", flush=True)
    oceans = members()
    cfg = default_config()
    alone = []
    for o in oceans:
        with ColumnPhysics.from_ocean(o, config=cfg) as cp:
            cp.step(DT, nsteps)
            alone.append({f: cp.get_field(f) for f in FIELDS})
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        assert ens.n_members == 5
        ens.step(DT, nsteps)
        for m in range(len(oceans)):
            for f in FIELDS:
                got = ens.member_get_field(m, f)
                assert np.array_equal(got, alone[m][f]), (m, f, float(np.nanmax(np.abs(got - alone[m][f]))))
        st = ens.stats()
        assert st["wet_columns"] == sum(o.n_wet for o in oceans)


def test_ensemble_autotune_and_graph_replay(gpu_lib):
    print("[case] ensemble autotune + repeated graph replays, bitwise vs untuned", flush=True)
    oceans = members(50)
    cfg = default_config()
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.step(DT, 6)
        ref = [ens.member_get_field(m, "temperature") for m in range(len(oceans))]
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.autotune(DT)  # restores the state it times on
        for _ in range(3):
            ens.step(DT, 2)  # one cached graph, replayed
        for m in range(len(oceans)):
            assert np.array_equal(ens.member_get_field(m, "temperature"), ref[m])


def test_ensemble_refuses_single_grid_entry_points(gpu_lib):
    print("[case] ensemble error paths", flush=True)
    oceans = members(60)[:2]
    with ColumnEnsemble.from_oceans(oceans) as ens:
        with pytest.raises(RdbError):
            ens.set_field("temperature", oceans[0].T)
        with pytest.raises(RdbError):
            ens.eos()
        with pytest.raises(RdbError):
            ens.member_get_field(2, "temperature")
        with pytest.raises(RdbError):
            ColumnPhysics.from_ocean(oceans[0])  # one live handle


def test_ensemble_guarded_surrogate(gpu_lib, tmp_path):
    print("[case] ensemble surrogate: train on redrawn members, 24 guarded steps vs all-plain", flush=True)
    oceans = members(70)
    cfg = default_config()
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.step(DT, 24)
        assert ens.stats()["graph_replay"] == 1  # the plain ensemble step replays its captured graph
        T_ref = [ens.member_get_field(m, "temperature") for m in range(len(oceans))]
    net = tmp_path / "ens.bin"
    with ColumnEnsemble.from_oceans([synth.redraw(o, 700 + i) for i, o in enumerate(oceans)], config=cfg) as ens:
        n = ens.mlp_collect(DT, 12)
        tl, vl = ens.mlp_train(epochs=60, batch=4096, guard_mode=GUARD_SURFACE)
        ens.mlp_save(net)
    assert n == 12 * sum(o.n_wet for o in oceans)
    assert vl[-1] < 0.5 * vl[0]
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.mlp_load(net)
        ens.set_guard(GUARD_SURFACE)
        ens.set_closure(CLOSURE_SURROGATE)
        h0, s0 = ens.heat_salt_totals()
        ens.step(DT, 24)
        h1, s1 = ens.heat_salt_totals()
        fb = ens.stats()["fallback_fraction"]
        assert ens.stats()["graph_replay"] == 1  # so does the guarded-surrogate step
        T = [ens.member_get_field(m, "temperature") for m in range(len(oceans))]
        kbl = [ens.member_get_field(m, "kbl") for m in range(len(oceans))]
    dep = 24 * DT / (cfg.rho0 * cfg.cp) * sum(float((o.Q_heat * o.wet).sum()) for o in oceans)
    assert abs((h1 - h0) - dep) <= 1e-12 * abs(h0)
    assert abs(s1 - s0) <= 1e-13 * abs(s0)
    rms = []
    for o, a, b in zip(oceans, T, T_ref):
        live = (o.h > 1.5e-4) & (o.wet > 0.5)[:, :, None]
        d = (a - b)[live]
        rms.append(float(np.sqrt(np.mean(d * d))))
    used = sum(int((k[o.wet > 0.5] == -1).sum()) for o, k in zip(oceans, kbl))
    print("ensemble-surrogate " + json.dumps({"T_rms_K_per_member": rms, "fallback_fraction": fb,
                                              "surrogate_columns_last_step": used, "val_loss": [vl[0], vl[-1]]}),
          flush=True)
    assert all(np.isfinite(r) for r in rms) and max(rms) < 0.2
    assert 0.0 <= fb < 1.0 and used > 0


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_ensemble_edge_members_bitwise(gpu_lib, precision):
    """No passive tracers, a single-layer member and a z-star member next to
    z-level ones; FP64 and split precision, each against the members alone."""
    from roundabout_amd.rdb import PREC_FP64, PREC_SPLIT
    print(f"[case] ensemble edge members ({precision}): nz 1/20/75, no passives, 6 steps", flush=True)
    mode = PREC_SPLIT if precision == "split" else PREC_FP64
    oceans = [synth.make_ocean(30, 20, 1, 80, land_frac=0.2, topo="zstar", n_passive=0),
              synth.make_ocean(40, 30, 20, 81, land_frac=0.1, topo="zstar", n_passive=0),
              synth.make_ocean(36, 28, 75, 82, land_frac=0.3, topo="zlevel", dmin=50.0, dmax=5500.0, n_passive=0)]
    cfg = default_config()
    fields = ("temperature", "salinity", "kv", "kt", "ks", "bl_depth", "kbl")
    alone = []
    for o in oceans:
        with ColumnPhysics.from_ocean(o, config=cfg) as cp:
            cp.set_precision(mode)
            cp.step(DT, 6)
            alone.append({f: cp.get_field(f) for f in fields})
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.set_precision(mode)
        ens.step(DT, 6)
        for m in range(len(oceans)):
            for f in fields:
                assert np.array_equal(ens.member_get_field(m, f), alone[m][f]), (m, f)


def test_learned_closure_error_paths(gpu_lib, tmp_path):
    from roundabout_amd.rdb import CLOSURE_SURROGATE
    print("[case] learned-closure error paths", flush=True)
    o = synth.small(90)
    with ColumnPhysics.from_ocean(o) as cp:
        with pytest.raises(RdbError):
            cp.mlp_train(epochs=1)  # nothing collected
        with pytest.raises(RdbError):
            cp.set_closure(CLOSURE_SURROGATE)  # no network
        with pytest.raises(RdbError):
            cp.mlp_load(tmp_path / "missing.bin")
        cp.mlp_collect(DT, 1)
        with pytest.raises(RdbError):
            cp.mlp_train(epochs=1, batch=1 << 21)  # batch above the limit
        with pytest.raises(RdbError):
            cp.mlp_train(epochs=1, batch=1 << 16)  # more than the collected columns: no full batch
        tl, vl = cp.mlp_train(epochs=1, batch=512)
        assert len(tl) == 1 and np.isfinite(tl[0]) and np.isfinite(vl[0])
        cp.mlp_save(tmp_path / "small.bin")
    # a network trained for nz = 20 does not load into a 75-level grid
    with ColumnPhysics.from_ocean(synth.tile(91, nx=16, ny=12)) as cp:
        with pytest.raises(RdbError):
            cp.mlp_load(tmp_path / "small.bin")


def test_ensemble_members_match_oracle_bitwise(orc, gpu_lib):
    """Each member of an uneven ensemble against the oracle stepping that
    member alone: a 256 x 256 x 50 z-level member (the smallest of the
    16-member configuration), a 160 x 128 x 75 z-level member and the
    64 x 64 x 20 z-star case, 24 one-hour steps in one persistent launch per
    step."""
    print("[case] ensemble 256x256x50 + 160x128x75 + 64x64x20, 24 steps, bitwise vs oracle", flush=True)
    oceans = [synth.tile(61, nx=256, ny=256, nz=50), synth.tile(62, nx=160, ny=128, nz=75), synth.small(63)]
    cfg = default_config()
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.autotune(DT)
        ens.step(DT, 24)
        got = [{f: ens.member_get_field(m, f) for f in FIELDS} for m in range(len(oceans))]
    for m, o in enumerate(oceans):
        ref = orc.OracleState(o, cfg)
        ref.step(DT, 24)
        want = {"temperature": ref.ocean.T, "salinity": ref.ocean.S, "tracer_1": ref.ocean.passive[0],
                "tracer_2": ref.ocean.passive[1], "kv": ref.kv, "kt": ref.kt, "ks": ref.ks,
                "bl_depth": ref.ocean.bl_depth, "kbl": ref.kbl, "b0": ref.b0}
        for f in FIELDS:
            assert np.array_equal(got[m][f], want[f]), (m, f, int(np.sum(got[m][f] != want[f])))
