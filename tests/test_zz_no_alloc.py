"""The allocation-free MLP entry point (rdb_mlp_eval_noalloc) against the
allocating one (rdb_mlp_eval): the same kernels, so the outputs must be the
same bits.  Two calls in a row with the second on a smaller batch (the
scratch array keeps the first call's contents), on the fused single-launch
path and on the per-layer path, and the size limit refused before any
work."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _eval(lib, fn, widths, p, x):
    wa = np.asarray(widths, dtype=np.int32)
    y = np.full((x.shape[0], widths[-1]), np.nan, dtype=np.float32)
    st = getattr(lib, fn)(wa.ctypes.data_as(_ip), len(widths), p.ctypes.data_as(_fp), x.shape[0],
                          x.ctypes.data_as(_fp), y.ctypes.data_as(_fp))
    return st, y


@pytest.mark.parametrize("widths,batches", [
    ([381, 256, 256, 256, 256, 149], (5000, 300)),   # the closure network: one fused launch
    ([500, 96, 64, 29], (700, 130)),                 # input wider than the fused kernel takes: per-layer GEMMs
])
def test_noalloc_matches_allocating_entry_point(gpu_lib, widths, batches):
    rng = np.random.default_rng(5)
    for i, batch in enumerate(batches):
        print(f"[case] mlp_eval_noalloc widths {widths} batch {batch} (call {i + 1})", flush=True)
        p = MO.init_params(widths, seed=40 + i)
        x = rng.normal(0, 1, (batch, widths[0])).astype(np.float32)
        st_a, y_a = _eval(gpu_lib, "rdb_mlp_eval", widths, p, x)
        st_n, y_n = _eval(gpu_lib, "rdb_mlp_eval_noalloc", widths, p, x)
        assert st_a == 0 and st_n == 0
        assert np.all(np.isfinite(y_n))
        assert np.array_equal(y_a, y_n), float(np.max(np.abs(y_a - y_n)))
    # and against the FP64 oracle of the network with BF16 rounding: the
    # FP32 accumulation order and the hardware exp/reciprocal of the GELU can
    # round a hidden activation to the neighbouring BF16 value (2^-8
    # relative), so a few outputs of a large batch move by more than the
    # typical error; the RMS stays small
    ye, _, _ = MO.forward(widths, p, x, bf16=True)
    err = y_n.astype(np.float64) - ye
    assert np.sqrt(np.mean(err ** 2)) < 1e-3 * np.sqrt(np.mean(ye ** 2))
    assert np.max(np.abs(err)) < 5e-3 * np.abs(ye).max()


def test_noalloc_refuses_a_batch_beyond_the_scratch(gpu_lib):
    print("[case] mlp_eval_noalloc: batch above the scratch limit is refused", flush=True)
    widths = [381, 256, 256, 256, 256, 149]
    p = MO.init_params(widths, seed=3)
    x = np.zeros((40000, 381), dtype=np.float32)
    st, y = _eval(gpu_lib, "rdb_mlp_eval_noalloc", widths, p, x)
    assert st == 14  # RDB_COLPHYS_ERR_BAD_SHAPE
    assert np.all(np.isnan(y))  # nothing was written
    buf = ctypes.create_string_buffer(256)
    gpu_lib.rdb_colphys_last_error(0, buf, 256)
    assert b"scratch" in buf.value
