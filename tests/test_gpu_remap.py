"""Parity of the HIP conservative vertical remap with the oracle's
restatement of the reference (src/ALE/rdb_remap_column.F90,
src/ALE/rdb_ocean_remap.F90): every output compared bitwise (``==``).

* rdb_remap_columns against remap_column for every method, both knobs and
  layer counts 1..75, on columns with zero, vanished, absorbed (a thickness
  below the running sum's resolution) and identical layers;
* rdb_remap_tracer_field against ocean_remap_tracer_field (donor column,
  filler excess, bounding, budget, clamp hook), in place and out of place,
  on fields with filler runs at the bed, inside and at the top and dead
  columns, and on a z-level crop of the 1440 x 1080 x 75 tile with a z*
  target;
* rdb_remap_face_velocity against remap_x/y_face_velocity with the mean or
  z-level face thickness and the KE rescale;
* rdb_remap_scan_preconditions against the oracle's scan.
"""

from __future__ import annotations

import numpy as np
import pytest

from roundabout_amd import synth
from roundabout_amd.rdb import remap as R

pytestmark = pytest.mark.gpu
# device arrays come from the HIP runtime the library runs on, not from torch:
# torch's bundled HIP runtime finds no GPU in a process where the library has
# already initialised the system one (as the earlier GPU tests do)
METHODS = (1, 2, 3, 4, 5)


def _columns(ncol, nz, seed):
    """(h_old, h_new, q) column batches, (nz, ncol) (so column-fastest in
    memory), with the awkward cases mixed in."""
    rng = np.random.default_rng(seed)
    ho = 0.5 + 30.0 * rng.random((nz, ncol))
    hn = 0.5 + 30.0 * rng.random((nz, ncol))
    q = 20.0 + 10.0 * rng.normal(size=(nz, ncol))
    kind = np.arange(ncol) % 8
    for c in range(ncol):
        k = kind[c]
        if k == 1:
            hn[:, c] = ho[:, c]                       # identity grid
        elif k == 2 and nz > 2:
            z = rng.integers(0, nz, size=max(1, nz // 4))
            ho[z, c] = 0.0                            # empty source layers
        elif k == 3 and nz > 2:
            z = rng.integers(0, nz, size=max(1, nz // 4))
            hn[z, c] = 0.0                            # empty targets
        elif k == 4:
            ho[: max(1, nz // 3), c] = 1.0e-4         # vanished bed run
            hn[: max(1, nz // 3), c] = 1.0e-4
        elif k == 5 and nz > 1:
            hn[rng.integers(0, nz), c] = 1.0e-20      # absorbed by the running sum
        elif k == 6:
            q[:, c] = np.where(np.arange(nz) < nz // 2, 0.0, 35.0)  # a step
        elif k == 7:
            q[:, c] = 34.5                            # uniform
        # same column total (scaled, then the last bits into the thickest layer)
        s_o, s_n = ho[:, c].sum(), hn[:, c].sum()
        if s_n > 0 and k != 1:
            hn[:, c] *= s_o / s_n
            j = int(np.argmax(hn[:, c]))
            hn[j, c] += s_o - hn[:, c].sum()
    return np.ascontiguousarray(ho), np.ascontiguousarray(hn), np.ascontiguousarray(q)


@pytest.mark.parametrize("nz", [1, 2, 3, 4, 5, 7, 20, 75])
def test_remap_columns_bitwise(orc, nz):
    ncol = 257
    ho, hn, q = _columns(ncol, nz, 100 + nz)
    # the device arrays are (ncol, nz) Fortran-ordered: the same bytes
    d_ho, d_hn, d_q = (R.DeviceArray.from_host(np.asfortranarray(a.T)) for a in (ho, hn, q))
    for method in METHODS:
        for be in (0, 1):
            for nu in (0, 1):
                print(f"[case] remap_columns nz={nz} method={method} be={be} nu={nu}", flush=True)
                got = R.remap_column(d_ho, d_hn, d_q, method, bnd_extrap=be, nonunif=nu).to_host().T
                want = np.stack([orc.remap_column(method, ho[:, c], hn[:, c], q[:, c], be, nu)
                                 for c in range(ncol)], axis=1)
                bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
                assert not bad.any(), (method, be, nu, int(bad.sum()), np.argwhere(bad)[:3].tolist())
    # in place
    got = R.remap_column(d_ho, d_hn, d_q, 3, q_new=d_q).to_host().T
    want = np.stack([orc.remap_column(3, ho[:, c], hn[:, c], q[:, c]) for c in range(ncol)], axis=1)
    assert np.array_equal(got, want)


def _field(nx, ny, nz, seed):
    """A field with filler runs (bed, interior, top), dead columns, stale
    filler concentrations and a z*-like or random target of the same column
    totals, Fortran-ordered (nx, ny, nz)."""
    rng = np.random.default_rng(seed)
    ho = 1.0 + 25.0 * rng.random((nx, ny, nz))
    hn = ho * (1.0 + 0.3 * (rng.random((nx, ny, nz)) - 0.5))
    c = 30.0 + 5.0 * rng.random((nx, ny, nz))
    for i in range(nx):
        for j in range(ny):
            kind = (i + 3 * j) % 6
            if kind == 1:
                n = 1 + (i + j) % max(1, nz // 2)
                ho[i, j, :n] = 1.0e-4
                hn[i, j, :n] = 1.0e-4 if (i + j) % 2 else 1.0e-3 + rng.random(n)
            elif kind == 2 and nz > 4:
                ho[i, j, 2:4] = 5.0e-5
                hn[i, j, 2] = 1.0e-4
            elif kind == 3:
                ho[i, j, -1] = 1.0e-4
                hn[i, j, -1] = 1.0e-4
            elif kind == 4:
                ho[i, j, :] = 1.0e-4                      # dead column
                hn[i, j, :] = 1.0e-4
            elif kind == 5:
                hn[i, j, :] = ho[i, j, :]                 # unchanged grid
            live = hn[i, j] > 1.5e-4
            if live.any() and kind not in (4, 5):
                hn[i, j, live] *= 1.0 + (ho[i, j].sum() - hn[i, j].sum()) / hn[i, j, live].sum()
                k = int(np.argmax(hn[i, j]))
                hn[i, j, k] += ho[i, j].sum() - hn[i, j].sum()
    return (np.asfortranarray(ho), np.asfortranarray(hn), np.asfortranarray(c))


@pytest.mark.parametrize("nz", [1, 3, 6, 20])
def test_remap_tracer_field_bitwise(orc, nz):
    nx, ny = 37, 23
    ho, hn, c = _field(nx, ny, nz, 7 + nz)
    d_ho, d_hn = R.DeviceArray.from_host(ho), R.DeviceArray.from_host(hn)
    for method in METHODS:
        for be in (0, 1):
            for nu in (0, 1):
                print(f"[case] remap_tracer_field {nx}x{ny}x{nz} method={method} be={be} nu={nu}", flush=True)
                bud0 = np.asfortranarray(np.random.default_rng(1).normal(size=(nx, ny, nz)))
                want_b = bud0.copy(order="F")
                want_c, want_cl = orc.remap_tracer_field(method, ho, hn, c, be, nu, budget=want_b)
                d_c, d_b = R.DeviceArray.from_host(c), R.DeviceArray.from_host(bud0)
                d_cl = R.DeviceArray.from_host(np.full((nx, ny, nz), np.nan, order="F"))
                if (method + be + nu) % 2:  # in place, as the reference
                    R.ocean_remap_tracer_field(d_ho, d_hn, d_c, method, bnd_extrap=be, nonunif=nu, budget=d_b,
                                               clamp_out=d_cl)
                    got_c = d_c.to_host()
                else:
                    d_out = R.DeviceArray((nx, ny, nz))
                    R.ocean_remap_tracer_field(d_ho, d_hn, d_c, method, bnd_extrap=be, nonunif=nu, budget=d_b,
                                               clamp_out=d_cl, c_out=d_out)
                    assert np.array_equal(d_c.to_host(), c)  # the input is left alone
                    got_c = d_out.to_host()
                for name, g, w in (("c", got_c, want_c), ("clamp", d_cl.to_host(), want_cl),
                                   ("budget", d_b.to_host(), want_b)):
                    assert np.array_equal(g, w), (name, method, be, nu, int(np.sum(g != w)))


def test_remap_tracer_field_tile_crop_zstar(orc):
    """A z-level 160 x 128 x 75 crop of the tile (partial bottom cells,
    vanished layers below the bathymetry), T and S remapped onto the z*
    grid of a free-surface anomaly, PPM default, bitwise."""
    print("[case] remap tile crop 160x128x75 z* PPM, T and S", flush=True)
    ocean = synth.tile(2, nx=160, ny=128)
    ho, hn = synth.ale_pair(ocean, seed=2)
    d_ho, d_hn = R.DeviceArray.from_host(ho), R.DeviceArray.from_host(hn)
    for name, fld in (("T", ocean.T), ("S", ocean.S)):
        bud = np.zeros_like(ho, order="F")
        want, want_cl = orc.remap_tracer_field(3, ho, hn, fld, budget=bud)
        d_c, d_b = R.DeviceArray.from_host(fld), R.DeviceArray.zeros(ho.shape)
        R.ocean_remap_tracer_field(d_ho, d_hn, d_c, 3, budget=d_b)
        got = d_c.to_host()
        assert np.array_equal(got, want), (name, int(np.sum(got != want)))
        assert np.array_equal(d_b.to_host(), bud)
        # per column the content is conserved up to what the bounding moved
        s_old = (ho * fld).sum(axis=2)
        s_new = (hn * got).sum(axis=2) - want_cl.sum(axis=2)
        assert np.max(np.abs(s_new - s_old) / np.maximum(np.abs(s_old), 1e-300)) < 1e-12


@pytest.mark.parametrize("zlevel", [0, 1])
def test_remap_faces_bitwise(orc, zlevel):
    nx, ny, nz = 29, 17, 12
    ho, hn, _ = _field(nx, ny, nz, 40 + zlevel)
    rng = np.random.default_rng(3)
    u = np.asfortranarray(rng.normal(0, 0.2, (nx + 1, ny, nz)))
    v = np.asfortranarray(rng.normal(0, 0.2, (nx, ny + 1, nz)))
    d_ho, d_hn = R.DeviceArray.from_host(ho), R.DeviceArray.from_host(hn)
    for method in METHODS:
        for ke in (0, 1):
            for be, nu in ((0, 0), (1, 1)):
                print(f"[case] faces zlevel={zlevel} method={method} ke={ke} be={be} nu={nu}", flush=True)
                d_u, d_v = R.DeviceArray.from_host(u), R.DeviceArray.from_host(v)
                R.ocean_apply_ale_remap_faces(d_ho, d_hn, d_u, d_v, method, conserve_ke=ke, zlevel_faces=zlevel,
                                              bnd_extrap=be, nonunif=nu)
                for axis, dev, arr in ((0, d_u, u), (1, d_v, v)):
                    want = orc.remap_face_velocity(axis, ho, hn, arr, method, ke, zlevel, be, nu)
                    got = dev.to_host()
                    assert np.array_equal(got, want), (axis, method, ke, int(np.sum(got != want)))


def test_remap_scan_preconditions_and_errors(orc, gpu_lib):
    nx, ny, nz = 50, 40, 30
    ho, hn, _ = _field(nx, ny, nz, 5)
    hn = hn.copy(order="F")
    hn[3, 4, 5] = -2.0
    hn[10, 11, 2] += 1e-3
    want = orc.remap_scan_preconditions(ho, hn, 1e-9)
    got = R.ocean_remap_scan_preconditions(R.DeviceArray.from_host(ho), R.DeviceArray.from_host(hn), 1e-9)
    assert got == want and got[0] >= 2 and got[2] == -2.0
    # refusals before any work: bad extents, NULL arrays, a bad face axis
    from roundabout_amd.rdb import _ffi
    assert gpu_lib.rdb_remap_tracer_field(3, 0, 4, 4, None, None, None, None, 0, 0, None, None, None) == \
        _ffi.ERR_BAD_SHAPE
    assert gpu_lib.rdb_remap_face_velocity(2, 4, 4, 4, None, None, None, 3, 0, 0, 0, 0, None) == _ffi.ERR_BAD_SHAPE
    assert "axis" in _ffi.last_error(gpu_lib)
