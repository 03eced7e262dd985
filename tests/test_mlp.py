"""The surrogate closure's MLP: the FP64 oracle's analytic gradient against
central finite differences (CPU), and the BF16 MFMA kernels (forward and
loss gradient through the C ABI) against the oracle (GPU)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as MO

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def test_oracle_gradient_matches_finite_differences():
    widths = [7, 9, 9, 5]
    rng = np.random.default_rng(0)
    p = MO.init_params(widths, seed=1).astype(np.float64) + rng.normal(0, 0.05, MO.n_params(widths))
    x = rng.normal(0, 1, (6, 7))
    t = rng.normal(0, 1, (6, 5))
    w = (rng.random((6, 5)) > 0.3).astype(np.float64)
    loss, g = MO.loss_grad(widths, p, x, t, w)
    eps = 1e-6
    for idx in rng.choice(len(p), 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[idx] += eps
        pm[idx] -= eps
        fd = (MO.loss_grad(widths, pp, x, t, w)[0] - MO.loss_grad(widths, pm, x, t, w)[0]) / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])


def test_gelu_derivative_and_bf16_rounding():
    x = np.linspace(-6, 6, 1001)
    fd = (MO.gelu(x + 1e-6) - MO.gelu(x - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - MO.gelu_grad(x))) < 1e-8
    # bf16 keeps 8 significant bits, ties to even
    assert MO.bf16_round(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert MO.bf16_round(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6


def _call_eval(lib, widths, params, x):
    wa = np.asarray(widths, dtype=np.int32)
    y = np.zeros((x.shape[0], widths[-1]), dtype=np.float32)
    st = lib.rdb_mlp_eval(wa.ctypes.data_as(_ip), len(widths), params.ctypes.data_as(_fp), x.shape[0],
                          np.ascontiguousarray(x, dtype=np.float32).ctypes.data_as(_fp), y.ctypes.data_as(_fp))
    assert st == 0
    return y


def _call_grad(lib, widths, params, x, t, w):
    wa = np.asarray(widths, dtype=np.int32)
    loss = ctypes.c_float()
    g = np.zeros(MO.n_params(widths), dtype=np.float32)
    st = lib.rdb_mlp_loss_grad(wa.ctypes.data_as(_ip), len(widths), params.ctypes.data_as(_fp), x.shape[0],
                               np.ascontiguousarray(x, dtype=np.float32).ctypes.data_as(_fp),
                               np.ascontiguousarray(t, dtype=np.float32).ctypes.data_as(_fp),
                               np.ascontiguousarray(w, dtype=np.float32).ctypes.data_as(_fp), ctypes.byref(loss),
                               g.ctypes.data_as(_fp))
    assert st == 0
    return loss.value, g


@pytest.mark.gpu
@pytest.mark.parametrize("widths,batch", [
    ([381, 256, 256, 256, 256, 149], 200),  # the closure network
    ([100, 64, 192, 149], 300),             # layers narrower than the waves' tiles; 2 k-steps ahead
    ([320, 64, 256, 40], 130),              # first layer past column 256 with a narrow output
])
def test_mlp_forward_matches_oracle(gpu_lib, widths, batch):
    print(f"[case] MLP forward {widths}, batch {batch}, vs FP64 oracle", flush=True)
    rng = np.random.default_rng(2)
    p = MO.init_params(widths, seed=3)
    x = rng.normal(0, 1, (batch, widths[0])).astype(np.float32)
    y = _call_eval(gpu_lib, widths, p, x).astype(np.float64)
    ye, _, _ = MO.forward(widths, p, x, bf16=True)
    y64, _, _ = MO.forward(widths, p, x, bf16=False)
    scale = np.abs(ye).max()
    assert np.max(np.abs(y - ye)) < 2e-3 * scale  # same roundings: FP32 accumulation order only
    assert np.max(np.abs(y - y64)) < 5e-2 * scale  # BF16 storage vs exact FP64


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [256, 200])
def test_mlp_gradient_matches_oracle(gpu_lib, batch):
    # 200 rows: the last 64-row tile is partly padding (masked rows)
    print(f"[case] MLP loss gradient 381-256x4-149, batch {batch}, vs FP64 oracle", flush=True)
    widths = [381, 256, 256, 256, 256, 149]
    rng = np.random.default_rng(4)
    p = MO.init_params(widths, seed=5)
    x = rng.normal(0, 1, (batch, 381)).astype(np.float32)
    t = rng.normal(0, 1, (batch, 149)).astype(np.float32)
    w = (rng.random((batch, 149)) > 0.2).astype(np.float32)
    loss, g = _call_grad(gpu_lib, widths, p, x, t, w)
    le, ge = MO.loss_grad(widths, p, x, t, w, bf16=True)
    l64, g64 = MO.loss_grad(widths, p, x, t, w, bf16=False)
    assert abs(loss - le) < 1e-3 * abs(le) and abs(loss - l64) < 2e-2 * abs(l64)
    g = g.astype(np.float64)
    rel_e = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    rel_64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    cos = float(g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)))
    print(f"mlp-grad rel_err_vs_bf16_emulation={rel_e:.3e} rel_err_vs_fp64={rel_64:.3e} cos={cos:.6f}", flush=True)
    assert rel_e < 1e-2 and rel_64 < 5e-2 and cos > 0.999
