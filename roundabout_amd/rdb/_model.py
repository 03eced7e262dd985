"""``ColumnPhysics``: the Python driver over librdb_colphys.so.

Mirrors the lifecycle of the reference's ``rdb.Model`` (python/rdb/_model.py:
one live handle per process, explicit ``close()``/context manager, a
``destroy`` that is safe to call twice, errors raised with the ring's
message) for the vertical-mixing column path.  Arrays cross the boundary as
Fortran-ordered float64 numpy arrays of the reference layout.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

LAYER_FIELDS = ("h_layer", "temperature", "salinity", "rho_layer", "heat_budget_vdiff", "salt_budget_vdiff")
IFACE_FIELDS = ("kv", "kt", "ks", "gamma_t", "gamma_s", "kd_epbl")
COLUMN_FIELDS = ("stress_shelf", "Q_heat", "Q_salt", "bl_depth", "stress_mag", "b0", "kbl", "coriolis", "wet_t",
                 # EPBL state (rdb_ocean_epbl.F90:283-312), kept when EPBL is on
                 "mld", "la", "b0_epbl", "tke_wind", "tke_conv", "tke_forcing", "tke_mixing", "tke_mech_decay",
                 "tke_conv_decay")
# face-located inputs of the momentum solve's folds (rdb_colphys_vdiff_momentum)
U_FACE_2D = ("lambda_bot_u", "lambda_top_u", "cover_u", "kv_bbl_u", "bbl_thick_u", "k_top_u", "k_bot_u")
V_FACE_2D = ("lambda_bot_v", "lambda_top_v", "cover_v", "kv_bbl_v", "bbl_thick_v", "k_top_v", "k_bot_v")


class HostField:
    """A window onto one host mirror of the library (rdb_colphys_get_*_ptr):
    it remembers the outer step count its contents were last known good at
    (the generation) and, on access, calls rdb_colphys_refresh_host when the
    handle has stepped since -- the lazy, silent re-sync of the reference's
    python/rdb Field (python/rdb/_field.py).  Entry points that change the
    state without a step (setters, the momentum solve) keep the step count:
    call ``refresh_host`` after them, as with the reference."""

    def __init__(self, cp, name):
        self._cp = cp
        self.name = name
        self._arr, self._gen = cp.host_pointer(name)

    @property
    def generation(self) -> int:
        return self._gen

    @property
    def values(self) -> np.ndarray:
        now = self._cp.step_count
        if now != self._gen:
            self._cp.refresh_host()
            self._gen = now
        return self._arr


class ColumnPhysics:
    """One column-physics plan on the GPU (``rdb_colphys_create``)."""

    def __init__(self, nx: int, ny: int, nz: int, wet, *, n_passive: int = 0, config=None,
                 closures=(), k_top=None, sort_key=None, lib=None, namelist=None):
        """``namelist``: Fortran namelist text (&ocean_vmix_nml, &ocean_eos_nml,
        &ocean_ic_nml) to create from instead of ``config``, as
        rdb_ocean_create_from_string does (rdb_colphys_create_from_string)."""
        self._lib = lib or _ffi.load()
        if namelist is not None:
            if config is not None or closures:
                raise ValueError("give either namelist text or config/closures")
            text = namelist.encode() if isinstance(namelist, str) else bytes(namelist)
            cfg = _ffi.Config()
            _ffi.check(self._lib.rdb_colphys_config_from_string(text, len(text), ctypes.byref(cfg)), self._lib)
        else:
            cfg = config if config is not None else _ffi.default_config()
        for c in closures:
            c.apply(cfg)
        self.config = cfg
        self.nx, self.ny, self.nz, self.n_passive = int(nx), int(ny), int(nz), int(n_passive)
        wet = np.asfortranarray(wet, dtype=np.float64)
        if wet.shape != (nx, ny):
            raise ValueError(f"wet mask shape {wet.shape} != ({nx}, {ny})")
        kt = None if k_top is None else np.asfortranarray(k_top, dtype=np.int32)
        sk = None if sort_key is None else np.asfortranarray(sort_key, dtype=np.int32)
        handle = ctypes.c_void_p()
        st = self._lib.rdb_colphys_create(
            ctypes.byref(cfg), self.nx, self.ny, self.nz, self.n_passive,
            wet.ctypes.data_as(_dp),
            None if kt is None else kt.ctypes.data_as(_ip),
            None if sk is None else sk.ctypes.data_as(_ip),
            ctypes.byref(handle))
        _ffi.check(st, self._lib)
        self._h = handle

    # ------------------------------------------------------------- factory
    @classmethod
    def from_ocean(cls, ocean, config=None, closures=(), sort: bool = True, lib=None,
                   namelist=None) -> "ColumnPhysics":
        """Create a plan for a ``roundabout_amd.synth.Ocean`` and upload it."""
        cp = cls(ocean.nx, ocean.ny, ocean.nz, ocean.wet, n_passive=ocean.n_passive, config=config,
                 closures=closures, k_top=ocean.k_top, sort_key=ocean.n_live if sort else None, lib=lib,
                 namelist=namelist)
        cp.upload(ocean)
        return cp

    def upload(self, ocean) -> None:
        self.set_field("h_layer", ocean.h)
        self.set_field("u_face_x_layer", ocean.u)
        self.set_field("v_face_y_layer", ocean.v)
        self.set_field("temperature", ocean.T)
        self.set_field("salinity", ocean.S)
        for i, p in enumerate(ocean.passive):
            self.set_field(f"tracer_{i + 1}", p)
        self.set_field("tau_x", ocean.tau_x)
        self.set_field("tau_y", ocean.tau_y)
        self.set_field("Q_heat", ocean.Q_heat)
        self.set_field("Q_salt", ocean.Q_salt)
        self.set_field("bl_depth", ocean.bl_depth)
        self.set_field("coriolis", ocean.coriolis)

    # ------------------------------------------------------------ lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value is not None:
            self._lib.rdb_colphys_destroy(ctypes.byref(self._h))
        self._h = None

    destroy = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _handle(self):
        if getattr(self, "_h", None) is None or self._h.value is None:
            raise _ffi.RdbError(_ffi.ERR_BAD_HANDLE, "handle is closed")
        return self._h

    # ---------------------------------------------------------------- fields
    def shape_of(self, name: str) -> tuple[int, int, int]:
        nx, ny, nz = self.nx, self.ny, self.nz
        if name in LAYER_FIELDS or name.startswith("tracer_"):
            return nx, ny, nz
        if name in IFACE_FIELDS:
            return nx, ny, nz + 1
        if name == "u_face_x_layer":
            return nx + 1, ny, nz
        if name == "v_face_y_layer":
            return nx, ny + 1, nz
        if name in ("tau_x",) + U_FACE_2D:
            return nx + 1, ny, 1
        if name in ("tau_y",) + V_FACE_2D:
            return nx, ny + 1, 1
        if name == "visc_rem_u":
            return nx + 1, ny, nz
        if name == "visc_rem_v":
            return nx, ny + 1, nz
        if name == "kv_corner":
            return nx + 1, ny + 1, nz + 1
        return nx, ny, 1  # 2-D column fields; unknown names get NOT_FOUND from the library

    def set_field(self, name: str, data) -> None:
        n1, n2, n3 = self.shape_of(name)
        arr = np.asfortranarray(data, dtype=np.float64)
        if arr.size != n1 * n2 * n3:
            raise ValueError(f"{name}: {arr.shape} does not hold {(n1, n2, n3)}")
        st = self._lib.rdb_colphys_set_field(self._handle(), name.encode(), arr.ctypes.data_as(_dp), n1, n2, n3)
        _ffi.check(st, self._lib)

    def get_field(self, name: str) -> np.ndarray:
        n1, n2, n3 = self.shape_of(name)
        out = np.empty((n1, n2, n3), order="F")
        st = self._lib.rdb_colphys_get_field(self._handle(), name.encode(), out.ctypes.data_as(_dp), n1, n2, n3)
        _ffi.check(st, self._lib)
        if name in COLUMN_FIELDS or name in ("tau_x", "tau_y") + U_FACE_2D + V_FACE_2D:
            out = out[:, :, 0]
        if name == "kbl":
            out = out.astype(np.int32)
        return out

    # -------------------------------------------- rdb_ocean.h host surface
    _PTR3 = ("h_layer", "u_face_x_layer", "v_face_y_layer", "rho_layer", "kv", "kt", "ks")
    _PTR2 = {"wet_t": "wet_t", "tau_x": "tau_x", "tau_y": "tau_y", "Q_heat": "q_heat", "Q_salt": "q_salt"}

    def host_pointer(self, name: str):
        """(read-only numpy view of the host mirror, generation) of a field,
        through the rdb_colphys_get_*_ptr getters; tracers by registry name."""
        p = ctypes.c_void_p()
        n = [ctypes.c_int() for _ in range(4)]
        h = self._handle()
        if name in self._PTR3:
            st = getattr(self._lib, f"rdb_colphys_get_{name}_ptr")(h, ctypes.byref(p), *[ctypes.byref(x) for x in n])
            shape = (n[0].value, n[1].value, n[2].value)
            gen = n[3].value
        elif name in self._PTR2:
            st = getattr(self._lib, f"rdb_colphys_get_{self._PTR2[name]}_ptr")(
                h, ctypes.byref(p), ctypes.byref(n[0]), ctypes.byref(n[1]), ctypes.byref(n[3]))
            shape = (n[0].value, n[1].value)
            gen = n[3].value
        else:
            b = name.encode()
            st = self._lib.rdb_colphys_get_tracer_ptr(h, b, len(b), ctypes.byref(p), *[ctypes.byref(x) for x in n])
            shape = (n[0].value, n[1].value, n[2].value)
            gen = n[3].value
        _ffi.check(st, self._lib)
        count = int(np.prod(shape))
        arr = np.ndarray(shape, dtype=np.float64, buffer=(ctypes.c_double * count).from_address(p.value),
                         order="F")
        arr.flags.writeable = False
        return arr, gen

    def field(self, name: str) -> "HostField":
        """A lazily synchronised view of a field (see HostField)."""
        return HostField(self, name)

    def refresh_host(self) -> None:
        _ffi.check(self._lib.rdb_colphys_refresh_host(self._handle()), self._lib)

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        _ffi.check(self._lib.rdb_colphys_get_time(self._handle(), ctypes.byref(t)), self._lib)
        return t.value

    @property
    def step_count(self) -> int:
        n = ctypes.c_int()
        _ffi.check(self._lib.rdb_colphys_get_step_count(self._handle(), ctypes.byref(n)), self._lib)
        return n.value

    @property
    def grid_info(self) -> tuple[int, int, int, int]:
        n = [ctypes.c_int() for _ in range(4)]
        _ffi.check(self._lib.rdb_colphys_get_grid_info(self._handle(), *[ctypes.byref(x) for x in n]), self._lib)
        return tuple(x.value for x in n)

    def tracer_names(self) -> list[str]:
        cnt = ctypes.c_int()
        _ffi.check(self._lib.rdb_colphys_get_tracer_count(self._handle(), ctypes.byref(cnt)), self._lib)
        out = []
        for i in range(cnt.value):
            buf = ctypes.create_string_buffer(64)
            self._lib.rdb_colphys_list_tracers(self._handle(), i, buf, 64)
            out.append(buf.value.decode())
        return out

    def set_budgets(self, enable: bool) -> None:
        """Keep heat/salt_budget_vdiff (fields of the same names)."""
        _ffi.check(self._lib.rdb_colphys_set_budgets(self._handle(), int(bool(enable))), self._lib)

    # --------------------------------------------------------------- compute
    def eos(self) -> None:
        _ffi.check(self._lib.rdb_colphys_eos(self._handle()), self._lib)

    def coefficients(self, dt: float) -> int:
        st = self._lib.rdb_colphys_coefficients(self._handle(), float(dt))
        if st not in (_ffi.OK, _ffi.GUARD_TRIPPED):
            _ffi.check(st, self._lib)
        return st

    def surface_flux(self, dt: float) -> None:
        _ffi.check(self._lib.rdb_colphys_surface_flux(self._handle(), float(dt)), self._lib)

    def diffuse_tracers(self, dt: float) -> None:
        _ffi.check(self._lib.rdb_colphys_diffuse_tracers(self._handle(), float(dt)), self._lib)

    def nonlocal_tendencies(self, dt: float) -> None:
        _ffi.check(self._lib.rdb_colphys_nonlocal(self._handle(), float(dt)), self._lib)

    def step(self, dt: float, n_steps: int = 1) -> int:
        st = self._lib.rdb_colphys_step(self._handle(), float(dt), int(n_steps))
        if st not in (_ffi.OK, _ffi.GUARD_TRIPPED):
            _ffi.check(st, self._lib)
        return st

    def vdiff_momentum(self, dt: float, **overrides) -> None:
        """Backward-Euler vertical viscosity of the face velocities
        (vdiff_apply_momentum); keyword arguments override
        rdb_colphys_momentum_config (use_kv_source=1 by default)."""
        cfg = _ffi.MomentumConfig()
        _ffi.check(self._lib.rdb_colphys_momentum_config_default(ctypes.byref(cfg)), self._lib)
        for k, v in overrides.items():
            if not hasattr(cfg, k):
                raise AttributeError(f"rdb_colphys_momentum_config has no field {k!r}")
            setattr(cfg, k, v)
        _ffi.check(self._lib.rdb_colphys_vdiff_momentum(self._handle(), float(dt), ctypes.byref(cfg)), self._lib)

    def synchronize(self) -> None:
        _ffi.check(self._lib.rdb_colphys_synchronize(self._handle()), self._lib)

    def autotune(self, dt: float) -> None:
        _ffi.check(self._lib.rdb_colphys_autotune(self._handle(), float(dt)), self._lib)

    def set_precision(self, mode: int) -> None:
        _ffi.check(self._lib.rdb_colphys_set_precision(self._handle(), int(mode)), self._lib)

    def set_closure(self, mode: int) -> None:
        _ffi.check(self._lib.rdb_colphys_set_closure(self._handle(), int(mode)), self._lib)

    def heat_salt_totals(self) -> tuple[float, float]:
        heat, salt = ctypes.c_double(), ctypes.c_double()
        _ffi.check(self._lib.rdb_colphys_heat_salt_totals(self._handle(), ctypes.byref(heat), ctypes.byref(salt)),
                   self._lib)
        return heat.value, salt.value

    def set_timing(self, enable: bool) -> None:
        _ffi.check(self._lib.rdb_colphys_set_timing(self._handle(), int(bool(enable))), self._lib)

    # ------------------------------------------------------- learned closure
    def mlp_collect(self, dt: float, n_steps: int, stride: int = 1) -> int:
        """Advance n_steps plain steps, recording training columns; returns
        the number of samples collected so far."""
        n = ctypes.c_longlong()
        _ffi.check(self._lib.rdb_colphys_mlp_collect(self._handle(), float(dt), int(n_steps), int(stride),
                                                     ctypes.byref(n)), self._lib)
        return n.value

    def mlp_train(self, **overrides) -> tuple[list, list]:
        """Train the network (Adam on the GPU); returns per-epoch training and
        held-out losses.  Keyword arguments override rdb_colphys_mlp_config."""
        cfg = _ffi.MlpConfig()
        _ffi.check(self._lib.rdb_colphys_mlp_config_default(ctypes.byref(cfg)), self._lib)
        for k, v in overrides.items():
            setattr(cfg, k, v)
        n = max(cfg.epochs, 1)
        tl = (ctypes.c_float * n)()
        vl = (ctypes.c_float * n)()
        done = ctypes.c_int()
        _ffi.check(self._lib.rdb_colphys_mlp_train(self._handle(), ctypes.byref(cfg), tl, vl, n, ctypes.byref(done)),
                   self._lib)
        return list(tl)[:done.value], list(vl)[:done.value]

    def mlp_save(self, path) -> None:
        _ffi.check(self._lib.rdb_colphys_mlp_save(self._handle(), str(path).encode()), self._lib)

    def mlp_load(self, path) -> None:
        _ffi.check(self._lib.rdb_colphys_mlp_load(self._handle(), str(path).encode()), self._lib)

    def set_guard(self, mode: int) -> None:
        _ffi.check(self._lib.rdb_colphys_set_guard(self._handle(), int(mode)), self._lib)

    def stats(self) -> dict:
        c = _ffi.Counters()
        _ffi.check(self._lib.rdb_colphys_stats(self._handle(), ctypes.byref(c)), self._lib)
        return {f: getattr(c, f) for f, _ in _ffi.Counters._fields_}


class ColumnEnsemble(ColumnPhysics):
    """Several independent members stepped together
    (``rdb_colphys_ensemble_create``): one persistent launch per step over
    every member's wet columns.  Fields are reached per member; stepping,
    the learned closure, totals and counters cover all members."""

    def __init__(self, grids, *, n_passive: int = 0, config=None, closures=(), sort_keys=None, lib=None):
        """``grids``: sequence of (nx, ny, nz, wet) tuples."""
        self._lib = lib or _ffi.load()
        cfg = config if config is not None else _ffi.default_config()
        for c in closures:
            c.apply(cfg)
        self.config = cfg
        self.n_passive = int(n_passive)
        self.dims = [(int(nx), int(ny), int(nz)) for nx, ny, nz, _ in grids]
        n = len(self.dims)
        wets = []
        for (nx, ny, _), (_, _, _, wet) in zip(self.dims, grids):
            w = np.asfortranarray(wet, dtype=np.float64)
            if w.shape != (nx, ny):
                raise ValueError(f"wet mask shape {w.shape} != ({nx}, {ny})")
            wets.append(w)
        keys = None
        if sort_keys is not None:
            keys = [np.asfortranarray(k, dtype=np.int32) for k in sort_keys]
        self._keep = (wets, keys)  # the library copies them; keep until create returns
        IntN = ctypes.c_int * n
        wet_ptrs = (_dp * n)(*[w.ctypes.data_as(_dp) for w in wets])
        key_ptrs = (_ip * n)(*[k.ctypes.data_as(_ip) for k in keys]) if keys is not None else None
        handle = ctypes.c_void_p()
        st = self._lib.rdb_colphys_ensemble_create(
            ctypes.byref(cfg), n, IntN(*[d[0] for d in self.dims]), IntN(*[d[1] for d in self.dims]),
            IntN(*[d[2] for d in self.dims]), self.n_passive, wet_ptrs, key_ptrs, ctypes.byref(handle))
        _ffi.check(st, self._lib)
        self._h = handle
        self._keep = None

    @classmethod
    def from_oceans(cls, oceans, config=None, closures=(), sort: bool = True, lib=None) -> "ColumnEnsemble":
        n_passive = {o.n_passive for o in oceans}
        if len(n_passive) != 1:
            raise ValueError("ensemble members must carry the same number of passive tracers")
        ens = cls([(o.nx, o.ny, o.nz, o.wet) for o in oceans], n_passive=n_passive.pop(), config=config,
                  closures=closures, sort_keys=[o.n_live for o in oceans] if sort else None, lib=lib)
        for m, o in enumerate(oceans):
            ens.upload_member(m, o)
        return ens

    @property
    def n_members(self) -> int:
        return len(self.dims)

    def member_shape(self, member: int, name: str) -> tuple[int, int, int]:
        if not 0 <= member < len(self.dims):
            return 1, 1, 1  # the library reports the unknown member (RDB_COLPHYS_ERR_NOT_FOUND)
        nx, ny, nz = self.dims[member]
        self.nx, self.ny, self.nz = nx, ny, nz
        return ColumnPhysics.shape_of(self, name)

    def member_set_field(self, member: int, name: str, data) -> None:
        n1, n2, n3 = self.member_shape(member, name)
        arr = np.asfortranarray(data, dtype=np.float64)
        if arr.size != n1 * n2 * n3:
            raise ValueError(f"{name}: {arr.shape} does not hold {(n1, n2, n3)}")
        st = self._lib.rdb_colphys_member_set_field(self._handle(), int(member), name.encode(), arr.ctypes.data_as(_dp),
                                                    n1, n2, n3)
        _ffi.check(st, self._lib)

    def member_get_field(self, member: int, name: str) -> np.ndarray:
        n1, n2, n3 = self.member_shape(member, name)
        out = np.empty((n1, n2, n3), order="F")
        st = self._lib.rdb_colphys_member_get_field(self._handle(), int(member), name.encode(),
                                                    out.ctypes.data_as(_dp), n1, n2, n3)
        _ffi.check(st, self._lib)
        if name in COLUMN_FIELDS or name in ("tau_x", "tau_y"):
            out = out[:, :, 0]
        if name == "kbl":
            out = out.astype(np.int32)
        return out

    def upload_member(self, member: int, ocean) -> None:
        for name, data in (("h_layer", ocean.h), ("u_face_x_layer", ocean.u), ("v_face_y_layer", ocean.v),
                           ("temperature", ocean.T), ("salinity", ocean.S), ("tau_x", ocean.tau_x),
                           ("tau_y", ocean.tau_y), ("Q_heat", ocean.Q_heat), ("Q_salt", ocean.Q_salt),
                           ("bl_depth", ocean.bl_depth), ("coriolis", ocean.coriolis)):
            self.member_set_field(member, name, data)
        for i, p in enumerate(ocean.passive):
            self.member_set_field(member, f"tracer_{i + 1}", p)

    def upload(self, oceans) -> None:
        for m, o in enumerate(oceans):
            self.upload_member(m, o)

    # the single-grid field calls go to the library, which refuses them on an
    # ensemble handle (RDB_COLPHYS_ERR_BAD_HANDLE): use the member_* forms
    def set_field(self, name: str, data) -> None:
        arr = np.asfortranarray(data, dtype=np.float64)
        n = arr.shape + (1,) * (3 - arr.ndim)
        st = self._lib.rdb_colphys_set_field(self._handle(), name.encode(), arr.ctypes.data_as(_dp), *n[:3])
        _ffi.check(st, self._lib)

    def get_field(self, name: str) -> np.ndarray:
        out = np.empty((1, 1, 1), order="F")
        st = self._lib.rdb_colphys_get_field(self._handle(), name.encode(), out.ctypes.data_as(_dp), 1, 1, 1)
        _ffi.check(st, self._lib)
        return out
