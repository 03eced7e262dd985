"""GPU checks that pin what the bitwise column-step tests cannot reach on
their own: the device cube root on the reference's golden bit table, the
Adam update of the training step against an FP64 oracle, the output-major
forward layout the surrogate uses, the guard of the learned closure against
a numpy oracle (per column), a surrogate step whose guard trips everywhere
against the plain oracle (bitwise), and retraining after inference."""

from __future__ import annotations

import ctypes
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import guard_oracle as GO
from oracle import mlp_oracle as MO
from roundabout_amd import synth
from roundabout_amd.rdb import (CLOSURE_SURROGATE, GUARD_RANGE, GUARD_STRICT, GUARD_SURFACE, ColumnPhysics,
                                default_config)

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
DT = 3600.0
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _h2d(h):
    return struct.unpack(">d", bytes.fromhex(h))[0]


def test_device_cbrt_golden_bits(gpu_lib):
    print("[case] device cube root on the 48 golden rows + exact cubes", flush=True)
    rows = json.loads((GOLDEN / "cbrt_golden.json").read_text())["rows"]
    x = np.array([_h2d(a) for a, _ in rows] + [27.0, -8.0, 0.125, 0.0, -0.0, 5e-324, np.inf, -np.inf])
    y = np.zeros_like(x)
    assert gpu_lib.rdb_colphys_cbrt(x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), len(x)) == 0
    got = [struct.pack(">d", v).hex().upper() for v in y[:48]]
    bad = [(a, r, g) for (a, r), g in zip(rows, got) if g != r]
    assert not bad, bad
    assert list(y[48:51]) == [3.0, -2.0, 0.5]
    assert struct.pack(">d", y[52]) == struct.pack(">d", -0.0) and y[51] == 0.0 and y[53] == 0.0
    assert y[54] == np.inf and y[55] == -np.inf


def _grad(lib, widths, p, x, t, w):
    wa = np.asarray(widths, dtype=np.int32)
    loss = ctypes.c_float()
    g = np.zeros(MO.n_params(widths), dtype=np.float32)
    assert lib.rdb_mlp_loss_grad(wa.ctypes.data_as(_ip), len(widths), p.ctypes.data_as(_fp), x.shape[0],
                                 x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), w.ctypes.data_as(_fp),
                                 ctypes.byref(loss), g.ctypes.data_as(_fp)) == 0
    return g


def _adam(lib, widths, p, x, t, w, n, lr, b1, b2, eps):
    wa = np.asarray(widths, dtype=np.int32)
    out = np.zeros_like(p)
    assert lib.rdb_mlp_adam_steps(wa.ctypes.data_as(_ip), len(widths), p.ctypes.data_as(_fp), x.shape[0],
                                  x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), w.ctypes.data_as(_fp), n, lr, b1, b2,
                                  eps, out.ctypes.data_as(_fp)) == 0
    return out


@pytest.mark.parametrize("batch", [128, 200])
def test_adam_update_matches_oracle(gpu_lib, batch):
    """Two training steps on one batch against the FP64 Adam of
    oracle/mlp_oracle.py applied to the kernels' own gradients (their parity
    with the oracle's gradient is test_mlp.py's): this pins the moments,
    beta1/beta2, the bias corrections at t = 1 and 2, lr and eps, and -- as
    the second gradient is taken from the weights the update refreshed --
    the BF16 copies of W and W^T the update writes."""
    print(f"[case] Adam: 2 training steps, widths 45-64-64-29, batch {batch}", flush=True)
    widths = [45, 64, 64, 29]
    rng = np.random.default_rng(11)
    p0 = MO.init_params(widths, seed=7)
    x = rng.normal(0, 1, (batch, widths[0])).astype(np.float32)
    t = rng.normal(0, 1, (batch, widths[-1])).astype(np.float32)
    w = (rng.random((batch, widths[-1])) > 0.25).astype(np.float32)
    lr, b1, b2, eps = 3e-3, 0.85, 0.99, 1e-7  # not the defaults: a swapped or ignored knob shows
    g1 = _grad(gpu_lib, widths, p0, x, t, w).astype(np.float64)
    p1 = _adam(gpu_lib, widths, p0, x, t, w, 1, lr, b1, b2, eps)
    e1, m, v = MO.adam_step(p0, g1, 0.0, 0.0, 1, lr, b1, b2, eps)

    def tol(ref, lr):
        # a few FP32 ulps of the parameter, plus 1e-6 of a step (FP32 moments)
        return 4 * np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-6 * lr

    assert np.all(np.abs(p1 - e1) <= tol(e1, lr)), np.max(np.abs(p1 - e1))
    # the update moved every parameter with a non-negligible gradient by ~lr
    big = np.abs(g1) > 1e-4
    assert np.allclose(np.abs(p1 - p0)[big], lr, rtol=1e-2)
    g2 = _grad(gpu_lib, widths, p1, x, t, w).astype(np.float64)
    p2 = _adam(gpu_lib, widths, p0, x, t, w, 2, lr, b1, b2, eps)
    e2, _, _ = MO.adam_step(p1.astype(np.float64), g2, m, v, 2, lr, b1, b2, eps)
    assert np.all(np.abs(p2 - e2) <= tol(e2, lr)), np.max(np.abs(p2 - e2))
    # and the same two steps with the default Adam knobs
    p2d = _adam(gpu_lib, widths, p0, x, t, w, 2, 1e-3, 0.9, 0.999, 1e-8)
    g1d = g1
    q1, md, vd = MO.adam_step(p0, g1d, 0.0, 0.0, 1)
    p1d = _adam(gpu_lib, widths, p0, x, t, w, 1, 1e-3, 0.9, 0.999, 1e-8)
    assert np.all(np.abs(p1d - q1) <= tol(q1, 1e-3)), np.max(np.abs(p1d - q1))
    g2d = _grad(gpu_lib, widths, p1d, x, t, w).astype(np.float64)
    q2, _, _ = MO.adam_step(p1d.astype(np.float64), g2d, md, vd, 2)
    assert np.all(np.abs(p2d - q2) <= tol(q2, 1e-3)), np.max(np.abs(p2d - q2))


def test_mlp_forward_output_major_matches_oracle(gpu_lib):
    print("[case] MLP forward, output-major layout, 381-256x4-149, batch 300 (3 row blocks)", flush=True)
    widths = [381, 256, 256, 256, 256, 149]
    rng = np.random.default_rng(13)
    p = MO.init_params(widths, seed=17)
    x = rng.normal(0, 1, (300, 381)).astype(np.float32)
    wa = np.asarray(widths, dtype=np.int32)
    yt = np.zeros((149, 300), dtype=np.float32)
    y = np.zeros((300, 149), dtype=np.float32)
    assert gpu_lib.rdb_mlp_eval_t(wa.ctypes.data_as(_ip), len(widths), p.ctypes.data_as(_fp), 300,
                                  x.ctypes.data_as(_fp), yt.ctypes.data_as(_fp)) == 0
    assert gpu_lib.rdb_mlp_eval(wa.ctypes.data_as(_ip), len(widths), p.ctypes.data_as(_fp), 300,
                                x.ctypes.data_as(_fp), y.ctypes.data_as(_fp)) == 0
    ye, _, _ = MO.forward(widths, p, x, bf16=True)
    scale = np.abs(ye).max()
    assert np.max(np.abs(yt.T.astype(np.float64) - ye)) < 2e-3 * scale
    # the two layouts hold the same numbers
    assert np.array_equal(yt.T, y)


# ------------------------------------------------------------------ the guard
def _read_net(path):
    """The network file: header, parameters, then the statistics
    fmean | fstd | fmin | fmax | tmean | tstd | tmin | tmax (version 5)."""
    b = Path(path).read_bytes()
    assert b[:8] == b"RDBSURR5"
    nz, din, dout, nw = struct.unpack("<4i", b[8:24])
    widths = np.frombuffer(b, dtype="<i4", count=nw, offset=24)
    off = 24 + 4 * nw
    npub = int(sum(int(o) * int(i) + int(o) for i, o in zip(widths[:-1], widths[1:])))
    params = np.frombuffer(b, dtype="<f4", count=npub, offset=off)
    off += 4 * npub
    stats = np.frombuffer(b, dtype="<f4", count=4 * din + 4 * dout, offset=off).copy()
    assert off + 4 * stats.size == len(b)
    return {"head": b[:off], "nz": nz, "din": din, "dout": dout, "params": params, "stats": stats}


def _write_net(path, net, fmin, fmax, tmin=None, tmax=None):
    din, dout = net["din"], net["dout"]
    st = net["stats"].copy()
    st[2 * din:3 * din] = fmin
    st[3 * din:4 * din] = fmax
    if tmin is not None:
        st[4 * din + 2 * dout:4 * din + 3 * dout] = tmin
    if tmax is not None:
        st[4 * din + 3 * dout:] = tmax
    Path(path).write_bytes(net["head"] + st.astype("<f4").tobytes())


@pytest.fixture(scope="module")
def small_net(tmp_path_factory, gpu_lib):
    """A briefly trained network for the 64 x 64 x 20 case (the guard tests
    need a network, not an accurate one)."""
    ocean = synth.small(31)
    path = tmp_path_factory.mktemp("net") / "small.bin"
    with ColumnPhysics.from_ocean(synth.redraw(ocean, 32), config=default_config()) as cp:
        cp.mlp_collect(DT, 2)
        cp.mlp_train(epochs=2, batch=1024, guard_mode=GUARD_SURFACE)
        cp.mlp_save(path)
    return ocean, _read_net(path)


def test_all_fallback_surrogate_step_is_the_plain_step(orc, gpu_lib, small_net, tmp_path):
    """Every feature range empty: the guard trips on every column, so each
    column takes the plain PP81 + KPP code inside the surrogate launch and
    the result must be the oracle's bitwise."""
    print("[case] surrogate, guard tripped everywhere, 64x64x20 seed 31, 3 steps vs oracle", flush=True)
    ocean, net = small_net
    path = tmp_path / "never.bin"
    _write_net(path, net, np.float32(3.0e38), np.float32(-3.0e38))
    cfg = default_config()
    ref = orc.OracleState(ocean, cfg)
    ref.step(DT, 3)
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.mlp_load(path)
        cp.set_guard(GUARD_RANGE)
        cp.set_closure(CLOSURE_SURROGATE)
        cp.step(DT, 3)
        assert cp.stats()["fallback_fraction"] == 1.0
        for name, want in (("temperature", ref.ocean.T), ("salinity", ref.ocean.S), ("kv", ref.kv),
                           ("kt", ref.kt), ("ks", ref.ks), ("bl_depth", ref.ocean.bl_depth), ("b0", ref.b0)):
            got = cp.get_field(name)
            assert np.array_equal(got, want), (name, np.max(np.abs(got - want)))
        for t in range(ocean.n_passive):
            assert np.array_equal(cp.get_field(f"tracer_{t + 1}"), ref.ocean.passive[t])
        assert np.array_equal(cp.get_field("kbl"), ref.kbl)


def _surrogate_columns(ocean, path, guard):
    """(mask of columns that used the network after one step, fallback fraction)."""
    with ColumnPhysics.from_ocean(ocean, config=default_config()) as cp:
        cp.mlp_load(path)
        cp.set_guard(guard)
        cp.set_closure(CLOSURE_SURROGATE)
        cp.step(DT, 1)
        kbl = cp.get_field("kbl")
        fb = cp.stats()["fallback_fraction"]
    return kbl == -1, fb


@pytest.mark.parametrize("guard", [GUARD_STRICT, GUARD_SURFACE, GUARD_RANGE])
def test_guard_flags_match_oracle(gpu_lib, small_net, tmp_path, guard):
    """Every feature range open: only the static-stability part of the guard
    can trip; the columns that took the network (kbl = -1) must be exactly
    the wet columns the numpy guard oracle passes."""
    print(f"[case] guard mode {guard} vs numpy oracle, 64x64x20 seed 31, ranges open", flush=True)
    ocean, net = small_net
    path = tmp_path / "open.bin"
    _write_net(path, net, np.float32(-3.0e38), np.float32(3.0e38))
    used, fb = _surrogate_columns(ocean, path, guard)
    wet = ocean.wet > 0.5
    flag = GO.stability_flags(guard, ocean.h, ocean.T, ocean.S)
    want = wet & ~flag
    assert np.array_equal(used, want), int(np.sum(used != want))
    assert fb == pytest.approx(float((wet & flag).sum()) / wet.sum(), abs=1e-15)
    if guard == GUARD_SURFACE:
        assert 0 < (wet & flag).sum() < wet.sum()  # the case exercises both outcomes
    if guard == GUARD_RANGE:
        assert fb == 0.0


def test_guard_range_on_one_input_matches_oracle(gpu_lib, small_net, tmp_path):
    """Only the Q_heat feature has a finite training range: the columns whose
    surface heat flux falls outside it go to the plain path."""
    print("[case] guard: range of the Q_heat feature only, 64x64x20 seed 31", flush=True)
    ocean, net = small_net
    din, nz = net["din"], net["nz"]
    q = ocean.Q_heat[ocean.wet > 0.5]
    lo, hi = np.float32(np.quantile(q, 0.2)), np.float32(np.quantile(q, 0.9))
    fmin = np.full(din, -3.0e38, dtype=np.float32)
    fmax = np.full(din, 3.0e38, dtype=np.float32)
    fmin[5 * nz + 1], fmax[5 * nz + 1] = lo, hi
    path = tmp_path / "qrange.bin"
    _write_net(path, net, fmin, fmax)
    used, fb = _surrogate_columns(ocean, path, GUARD_RANGE)
    wet = ocean.wet > 0.5
    want = wet & ~GO.range_flags(ocean.Q_heat, lo, hi)
    assert np.array_equal(used, want), int(np.sum(used != want))
    assert fb == pytest.approx(0.3, abs=0.02)


def test_training_records_the_targets_range(small_net):
    """The file carries each target's range over the training set (weighted
    rows only): finite, ordered, and around the mean."""
    print("[case] network file: targets' range, 64x64x20 seed 31", flush=True)
    _, net = small_net
    din, dout = net["din"], net["dout"]
    st = net["stats"]
    tmean, tmin, tmax = st[4 * din:4 * din + dout], st[4 * din + 2 * dout:4 * din + 3 * dout], st[4 * din + 3 * dout:]
    assert np.all(np.isfinite(tmin)) and np.all(np.isfinite(tmax)) and np.all(tmin <= tmax)
    on = tmax < 1e38  # outputs no training row weights are left unclamped (the widest range)
    assert on.sum() > dout // 2 and np.all(tmin[~on] < -1e38)
    assert np.all(tmin[on] <= tmean[on] + 1e-3) and np.all(tmean[on] <= tmax[on] + 1e-3)


def test_predictions_are_clamped_to_the_targets_range(gpu_lib, small_net, tmp_path):
    """A targets' range of a single value pins every prediction: log10 kt =
    -3 at every interface and a boundary-layer depth of 0 (no KPP profile to
    merge), so each surrogate column's kt is 10^-3 between live layers."""
    print("[case] surrogate with the targets' range pinned, 64x64x20 seed 31, 1 step", flush=True)
    ocean, net = small_net
    nz, dout = net["nz"], net["dout"]
    tmin = np.full(dout, -3.0, dtype=np.float32)  # log10 kt and log10 kv
    tmin[-1] = 0.0                                # boundary-layer depth
    path = tmp_path / "pinned.bin"
    _write_net(path, net, np.float32(-3.0e38), np.float32(3.0e38), tmin, tmin)
    cfg = default_config()
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.mlp_load(path)
        cp.set_guard(GUARD_RANGE)
        cp.set_closure(CLOSURE_SURROGATE)
        cp.step(DT, 1)
        assert cp.stats()["fallback_fraction"] == 0.0
        kt, kv, bl = cp.get_field("kt"), cp.get_field("kv"), cp.get_field("bl_depth")
    ten_m3 = float(np.exp2(np.float32(-3.0) * np.float32(3.32192809)))  # 10^-3 in single precision
    want = min(max(ten_m3, cfg.kt_bg), cfg.kd_max)
    want_v = min(max(ten_m3, cfg.kv_bg), cfg.kv_max)
    wet = ocean.wet > 0.5
    live = ocean.h > 1.5e-4
    both = live[:, :, 1:] & live[:, :, :-1] & wet[:, :, None]  # interfaces 1 .. nz-1 between live layers
    assert both.sum() > 1000
    # (the hardware exp2 and numpy's may differ in the last bits of single precision)
    assert kt.shape[2] == nz + 1  # interfaces 0 .. nz
    np.testing.assert_allclose(kt[:, :, 1:nz][both], want, rtol=4e-7, atol=0)
    np.testing.assert_allclose(kv[:, :, 1:nz][both], want_v, rtol=4e-7, atol=0)
    assert np.all(bl[wet] == 0.0)


def test_version4_network_file_loads_without_clamp(gpu_lib, small_net, tmp_path):
    """A file written before the targets' range existed (version 4) still
    loads; its predictions are not clamped, so it steps like the version-5
    file whose range is wide open."""
    print("[case] version-4 network file, 64x64x20 seed 31, 2 steps", flush=True)
    ocean, net = small_net
    din, dout = net["din"], net["dout"]
    st = net["stats"].copy()
    st[2 * din:4 * din] = np.concatenate([np.full(din, -3.0e38), np.full(din, 3.0e38)]).astype(np.float32)
    p4, p5 = tmp_path / "v4.bin", tmp_path / "v5.bin"
    p4.write_bytes(b"RDBSURR4" + net["head"][8:] + st[:4 * din + 2 * dout].astype("<f4").tobytes())
    _write_net(p5, net, np.float32(-3.0e38), np.float32(3.0e38), np.float32(-3.0e38), np.float32(3.0e38))
    out = []
    for path in (p4, p5):
        with ColumnPhysics.from_ocean(ocean, config=default_config()) as cp:
            cp.mlp_load(path)
            cp.set_guard(GUARD_SURFACE)
            cp.set_closure(CLOSURE_SURROGATE)
            cp.step(DT, 2)
            out.append((cp.get_field("temperature"), cp.get_field("kt")))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_retraining_after_surrogate_steps(gpu_lib, tmp_path):
    """collect -> train -> surrogate step -> retrain with a batch smaller than
    the wet-column count -> surrogate step: the inference workspace is set up
    again for every wet row (the retraining resized the network's buffers)."""
    print("[case] retrain between surrogate steps, tile 160x128x75 seed 8 (> 8192 wet columns)", flush=True)
    ocean = synth.tile(8, nx=160, ny=128)
    assert ocean.n_wet > 8192
    cfg = default_config()
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.mlp_collect(DT, 1)
        cp.mlp_train(epochs=1, batch=2048, guard_mode=GUARD_SURFACE)
        cp.set_closure(CLOSURE_SURROGATE)
        h0, s0 = cp.heat_salt_totals()
        cp.step(DT, 1)
        fb1 = cp.stats()["fallback_fraction"]
        cp.mlp_train(epochs=1, batch=256, guard_mode=GUARD_SURFACE)
        cp.step(DT, 2)
        fb2 = cp.stats()["fallback_fraction"]
        h1, s1 = cp.heat_salt_totals()
        T = cp.get_field("temperature")
        kt = cp.get_field("kt")
    assert 0.0 <= fb1 <= 1.0 and 0.0 <= fb2 <= 1.0
    assert np.all(np.isfinite(T)) and np.all(np.isfinite(kt))
    dep = 3 * DT / (cfg.rho0 * cfg.cp) * float((ocean.Q_heat * ocean.wet).sum())
    assert abs((h1 - h0) - dep) <= 1e-12 * abs(h0)
    assert abs(s1 - s0) <= 1e-13 * abs(s0)
