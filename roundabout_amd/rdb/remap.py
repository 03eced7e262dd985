"""Conservative vertical (ALE) remap on device arrays: the Python side of the
``rdb_remap_*`` entry points.

Mirrors the reference's remap interfaces:

* :func:`ocean_remap_tracer_field` -- ``ocean_remap_tracer_field``
  (src/ALE/rdb_ocean_remap.F90:236-319), the concentration-form tracer remap
  with its budget slot and the ``clamp_out`` hook;
* :func:`ocean_apply_ale_remap_faces` -- ``ocean_apply_ale_remap_faces``
  (:915-965): both face-velocity passes;
* :func:`remap_column` -- ``remap_column`` (src/ALE/rdb_remap_column.F90:
  328-375) over a batch of columns;
* :func:`ocean_remap_scan_preconditions` -- (:1348-1405).

Device memory is :class:`DeviceArray`: a buffer of the HIP runtime the
library itself runs on, holding a Fortran-ordered float64 array of the
reference shape (``(nx, ny, nz)`` for a field, i fastest, k = 1 at the bed),
which is the natural layout the kernels read.  Work is enqueued on the null
stream; :func:`synchronize` waits for it.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi

REMAP_PCM, REMAP_PLM, REMAP_PPM, REMAP_PPM_H4, REMAP_PQM = 1, 2, 3, 4, 5
METHODS = {"pcm": REMAP_PCM, "plm": REMAP_PLM, "ppm": REMAP_PPM, "ppm_h4": REMAP_PPM_H4, "pqm": REMAP_PQM}
OCEAN_REMAP_PRECOND_RTOL = 1.0e-9  # rdb_ocean_remap.F90:49
_H2D, _D2H, _D2D = 1, 2, 3

_HIP = None


def _hip() -> ctypes.CDLL:
    """The HIP runtime the library was loaded against (by soname, after the
    library, so both use the same one)."""
    global _HIP
    if _HIP is None:
        _ffi.load()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipDeviceSynchronize.argtypes = []
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipEventDestroy.argtypes = [vp]
        for fn in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipEventCreate",
                   "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime", "hipEventDestroy"):
            getattr(h, fn).restype = ctypes.c_int
        _HIP = h
    return _HIP


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: HIP error {rc}")


def synchronize() -> None:
    _chk(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A float64 array of a reference shape in device memory, Fortran
    order."""

    def __init__(self, shape):
        self.shape = tuple(int(n) for n in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        p = ctypes.c_void_p()
        _chk(_hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self._p = p

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        d = cls(a.shape)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape) -> "DeviceArray":
        d = cls(shape)
        _chk(_hip().hipMemset(d._p, 0, d.nbytes), "hipMemset")
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _chk(_hip().hipMemcpy(self._p, a.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64, order="F")
        _chk(_hip().hipMemcpy(out.ctypes.data, self._p, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def copy(self) -> "DeviceArray":
        d = DeviceArray(self.shape)
        _chk(_hip().hipMemcpy(d._p, self._p, self.nbytes, _D2D), "hipMemcpy D2D")
        return d

    def free(self) -> None:
        if self._p is not None and self._p.value:
            _hip().hipFree(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown
            pass


class Timer:
    """Kernel time on the null stream from HIP events."""

    def __init__(self):
        h = _hip()
        self._a, self._b = ctypes.c_void_p(), ctypes.c_void_p()
        _chk(h.hipEventCreate(ctypes.byref(self._a)), "hipEventCreate")
        _chk(h.hipEventCreate(ctypes.byref(self._b)), "hipEventCreate")

    def start(self) -> None:
        _chk(_hip().hipEventRecord(self._a, None), "hipEventRecord")

    def stop_ms(self) -> float:
        h = _hip()
        _chk(h.hipEventRecord(self._b, None), "hipEventRecord")
        _chk(h.hipEventSynchronize(self._b), "hipEventSynchronize")
        ms = ctypes.c_float()
        _chk(h.hipEventElapsedTime(ctypes.byref(ms), self._a, self._b), "hipEventElapsedTime")
        return float(ms.value)


def parse_remap_method(name: str) -> int:
    """The namelist string of a remap method to its code (the &vcoord_nml
    ``remap_method`` strings); unknown strings are an error."""
    key = name.strip().lower()
    if key not in METHODS:
        raise ValueError(f"unknown remap method {name!r} (one of {sorted(METHODS)})")
    return METHODS[key]


def _method(method) -> int:
    return parse_remap_method(method) if isinstance(method, str) else int(method)


def _p(d: DeviceArray | None, shape, name):
    if d is None:
        return None
    if not isinstance(d, DeviceArray):
        raise TypeError(f"{name}: a DeviceArray is required")
    if d.shape != tuple(shape):
        raise ValueError(f"{name}: shape {d.shape} != {tuple(shape)}")
    return d.ptr


def ocean_remap_tracer_field(h_old, h_new, c, method=REMAP_PPM, *, bnd_extrap=False, nonunif=False, budget=None,
                             clamp_out=None, c_out=None):
    """Remap the concentration ``c`` (nx, ny, nz) from ``h_old`` to ``h_new``:
    in place, as the reference, or into ``c_out``; ``budget`` accumulates
    h_new*c_new - h_old*c_old, ``clamp_out`` receives the content the
    bounding moved."""
    lib = _ffi.load()
    nx, ny, nz = c.shape
    shp = (nx, ny, nz)
    out = c if c_out is None else c_out
    _ffi.check(lib.rdb_remap_tracer_field(_method(method), nx, ny, nz, _p(h_old, shp, "h_old"),
                                          _p(h_new, shp, "h_new"), _p(c, shp, "c"), _p(out, shp, "c_out"),
                                          int(bnd_extrap), int(nonunif), _p(budget, shp, "budget"),
                                          _p(clamp_out, shp, "clamp_out"), None), lib)
    return out


def ocean_apply_ale_remap_faces(h_old, h_new, u_face_x, v_face_y, method=REMAP_PPM, *, conserve_ke=False,
                                zlevel_faces=False, bnd_extrap=False, nonunif=False):
    """Remap u_face_x (nx+1, ny, nz) and v_face_y (nx, ny+1, nz) in place."""
    lib = _ffi.load()
    nx, ny, nz = h_old.shape
    ph = [_p(h_old, (nx, ny, nz), "h_old"), _p(h_new, (nx, ny, nz), "h_new")]
    for axis, u, shp in ((0, u_face_x, (nx + 1, ny, nz)), (1, v_face_y, (nx, ny + 1, nz))):
        if u is None:
            continue
        _ffi.check(lib.rdb_remap_face_velocity(axis, nx, ny, nz, ph[0], ph[1], _p(u, shp, "face velocity"),
                                               _method(method), int(conserve_ke), int(zlevel_faces),
                                               int(bnd_extrap), int(nonunif), None), lib)
    return u_face_x, v_face_y


def remap_column(dz_old, dz_new, q_old, method=REMAP_PPM, *, bnd_extrap=False, nonunif=False, q_new=None):
    """remap_column over a batch of columns: arrays (ncol, nz), column index
    fastest; q_new may be q_old."""
    lib = _ffi.load()
    ncol, nz = q_old.shape
    if q_new is None:
        q_new = DeviceArray((ncol, nz))
    shp = (ncol, nz)
    _ffi.check(lib.rdb_remap_columns(_method(method), ncol, nz, _p(dz_old, shp, "dz_old"), _p(dz_new, shp, "dz_new"),
                                     _p(q_old, shp, "q_old"), _p(q_new, shp, "q_new"), int(bnd_extrap),
                                     int(nonunif), None), lib)
    return q_new


def ocean_remap_scan_preconditions(h_old, h_new, rel_tol=OCEAN_REMAP_PRECOND_RTOL):
    """(n_bad, worst_rel, worst_neg) over every column."""
    lib = _ffi.load()
    nx, ny, nz = h_old.shape
    n_bad, worst_rel, worst_neg = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _ffi.check(lib.rdb_remap_scan_preconditions(nx, ny, nz, _p(h_old, (nx, ny, nz), "h_old"),
                                                _p(h_new, (nx, ny, nz), "h_new"), float(rel_tol),
                                                ctypes.byref(n_bad), ctypes.byref(worst_rel), ctypes.byref(worst_neg),
                                                None), lib)
    return n_bad.value, worst_rel.value, worst_neg.value
