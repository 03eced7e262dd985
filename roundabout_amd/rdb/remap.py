"""Conservative vertical (ALE) remap on device arrays: the Python side of the
``rdb_remap_*`` entry points.

Mirrors the reference's remap interfaces:

* :func:`ocean_remap_tracer_field` -- ``ocean_remap_tracer_field``
  (src/ALE/rdb_ocean_remap.F90:236-319), the concentration-form tracer remap
  with its budget slot and the ``c_new_out`` / ``clamp_out`` hooks;
* :func:`ocean_apply_ale_remap_faces` -- ``ocean_apply_ale_remap_faces``
  (:915-965): both face-velocity passes;
* :func:`remap_column` -- ``remap_column`` (src/ALE/rdb_remap_column.F90:
  328-375) over a batch of columns;
* :func:`ocean_remap_scan_preconditions` -- (:1348-1405).

Arrays are torch tensors on the GPU in the natural layout: a field of the
reference shape ``(nx, ny, nz)`` (Fortran order, i fastest, k = 1 at the bed)
is the C-contiguous tensor of shape ``(nz, ny, nx)``; :func:`to_device` and
:func:`to_host` convert from and to Fortran-ordered numpy arrays of the
reference shape.  Everything runs on torch's current stream.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi

REMAP_PCM, REMAP_PLM, REMAP_PPM, REMAP_PPM_H4, REMAP_PQM = 1, 2, 3, 4, 5
METHODS = {"pcm": REMAP_PCM, "plm": REMAP_PLM, "ppm": REMAP_PPM, "ppm_h4": REMAP_PPM_H4, "pqm": REMAP_PQM}
OCEAN_REMAP_PRECOND_RTOL = 1.0e-9  # rdb_ocean_remap.F90:49


def parse_remap_method(name: str) -> int:
    """The namelist string of a remap method to its code (the &vcoord_nml
    ``remap_method`` strings); unknown strings are an error."""
    key = name.strip().lower()
    if key not in METHODS:
        raise ValueError(f"unknown remap method {name!r} (one of {sorted(METHODS)})")
    return METHODS[key]


def _torch():
    import torch
    return torch


def to_device(a: np.ndarray):
    """A Fortran-ordered float64 array of the reference shape as a device
    tensor in the natural layout (shape reversed)."""
    torch = _torch()
    a = np.asarray(a, dtype=np.float64)
    flat = np.ascontiguousarray(a.ravel(order="F"))
    return torch.from_numpy(flat).to("cuda").view(*reversed(a.shape))


def to_host(t) -> np.ndarray:
    """The inverse of :func:`to_device`."""
    a = t.detach().to("cpu").contiguous().numpy()
    return np.asfortranarray(a.reshape(-1).reshape(tuple(reversed(t.shape)), order="F"))


def _check(t, shape, name):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError(f"{name}: a contiguous float64 CUDA tensor is required")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != {tuple(shape)}")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _method(method) -> int:
    return parse_remap_method(method) if isinstance(method, str) else int(method)


def ocean_remap_tracer_field(h_old, h_new, c, method=REMAP_PPM, *, bnd_extrap=False, nonunif=False, budget=None,
                             c_new_out=None, clamp_out=None):
    """Remap the concentration ``c`` (nz, ny, nx) from ``h_old`` to ``h_new``
    in place; ``budget`` accumulates h_new*c_new - h_old*c_old,
    ``c_new_out`` / ``clamp_out`` receive the bounded concentration and the
    content the bounding moved."""
    lib = _ffi.load()
    nz, ny, nx = c.shape
    shp = (nz, ny, nx)
    p = [_check(h_old, shp, "h_old"), _check(h_new, shp, "h_new"), _check(c, shp, "c")]
    pb = _check(budget, shp, "budget") if budget is not None else None
    pc = _check(clamp_out, shp, "clamp_out") if clamp_out is not None else None
    _ffi.check(lib.rdb_remap_tracer_field(_method(method), nx, ny, nz, p[0], p[1], p[2], p[2], int(bnd_extrap),
                                          int(nonunif), pb, pc, _stream()), lib)
    if c_new_out is not None:
        c_new_out.copy_(c)
    return c


def remap_tracer_field_into(h_old, h_new, c_in, c_out, method=REMAP_PPM, *, bnd_extrap=False, nonunif=False,
                            budget=None, clamp_out=None):
    """:func:`ocean_remap_tracer_field` out of place: ``c_in`` is left as it
    is and ``c_out`` receives the remapped field (no copy-back)."""
    lib = _ffi.load()
    nz, ny, nx = c_in.shape
    shp = (nz, ny, nx)
    p = [_check(h_old, shp, "h_old"), _check(h_new, shp, "h_new"), _check(c_in, shp, "c_in"),
         _check(c_out, shp, "c_out")]
    pb = _check(budget, shp, "budget") if budget is not None else None
    pc = _check(clamp_out, shp, "clamp_out") if clamp_out is not None else None
    _ffi.check(lib.rdb_remap_tracer_field(_method(method), nx, ny, nz, *p, int(bnd_extrap), int(nonunif), pb, pc,
                                          _stream()), lib)
    return c_out


def ocean_apply_ale_remap_faces(h_old, h_new, u_face_x, v_face_y, method=REMAP_PPM, *, conserve_ke=False,
                                zlevel_faces=False, bnd_extrap=False, nonunif=False):
    """Remap u_face_x (nz, ny, nx+1) and v_face_y (nz, ny+1, nx) in place."""
    lib = _ffi.load()
    nz, ny, nx = h_old.shape
    ph = [_check(h_old, (nz, ny, nx), "h_old"), _check(h_new, (nz, ny, nx), "h_new")]
    for axis, u, shp in ((0, u_face_x, (nz, ny, nx + 1)), (1, v_face_y, (nz, ny + 1, nx))):
        if u is None:
            continue
        _ffi.check(lib.rdb_remap_face_velocity(axis, nx, ny, nz, ph[0], ph[1], _check(u, shp, "face velocity"),
                                               _method(method), int(conserve_ke), int(zlevel_faces),
                                               int(bnd_extrap), int(nonunif), _stream()), lib)
    return u_face_x, v_face_y


def remap_column(dz_old, dz_new, q_old, method=REMAP_PPM, *, bnd_extrap=False, nonunif=False, q_new=None):
    """remap_column over a batch: arrays (nz, ncol), column index fastest."""
    torch = _torch()
    lib = _ffi.load()
    nz, ncol = q_old.shape
    shp = (nz, ncol)
    if q_new is None:
        q_new = torch.empty_like(q_old)
    p = [_check(dz_old, shp, "dz_old"), _check(dz_new, shp, "dz_new"), _check(q_old, shp, "q_old"),
         _check(q_new, shp, "q_new")]
    _ffi.check(lib.rdb_remap_columns(_method(method), ncol, nz, *p, int(bnd_extrap), int(nonunif), _stream()), lib)
    return q_new


def ocean_remap_scan_preconditions(h_old, h_new, rel_tol=OCEAN_REMAP_PRECOND_RTOL):
    """(n_bad, worst_rel, worst_neg) over every column."""
    lib = _ffi.load()
    nz, ny, nx = h_old.shape
    n_bad, worst_rel, worst_neg = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _ffi.check(lib.rdb_remap_scan_preconditions(nx, ny, nz, _check(h_old, (nz, ny, nx), "h_old"),
                                                _check(h_new, (nz, ny, nx), "h_new"), float(rel_tol),
                                                ctypes.byref(n_bad), ctypes.byref(worst_rel), ctypes.byref(worst_neg),
                                                _stream()), lib)
    return n_bad.value, worst_rel.value, worst_neg.value
