"""Vertical-closure knob objects, after the reference's ``python/rdb/closures.py:21-85``.

Each object writes only its own knobs into an ``rdb_colphys_config``
(``roundabout_amd.rdb._ffi.Config``).  Names, keywords and defaults follow
the reference's curated Python objects; the namelist-level defaults (used
when no object is passed) are those of ``src/core/rdb_config.F90:1872-1893``.
"""

from __future__ import annotations


class PacanowskiPhilander:
    """PP81 interior closure.

    ``PacanowskiPhilander(nu0=1e-2, nu_bg=1e-3, kappa_bg=1e-4, alpha=5.0,
    shear2_floor=1e-9, enabled=True)``.  The backgrounds also seed the
    assembly floors (kv_bg = nu_bg, kt_bg = ks_bg = kappa_bg), as the
    reference's configure pass does (src/core/rdb_config.F90:1874-1881).
    ``enabled=False`` is refused by the column engine: the reference then
    skips the whole coefficient build, which is not this path.
    """

    def __init__(self, nu0=1e-2, nu_bg=1e-3, kappa_bg=1e-4, alpha=5.0, shear2_floor=1e-9, enabled=True):
        self.nu0, self.nu_bg, self.kappa_bg = nu0, nu_bg, kappa_bg
        self.alpha, self.shear2_floor = alpha, shear2_floor
        self.enabled = bool(enabled)

    def apply(self, config):
        if not self.enabled:
            raise ValueError("PacanowskiPhilander(enabled=False): use_closure=.false. has no column-physics path")
        config.pp81_nu0 = float(self.nu0)
        config.pp81_nu_bg = float(self.nu_bg)
        config.pp81_kappa_bg = float(self.kappa_bg)
        config.pp81_alpha = float(self.alpha)
        config.shear2_floor = float(self.shear2_floor)
        config.kv_bg = float(self.nu_bg)
        config.kt_bg = float(self.kappa_bg)
        config.ks_bg = float(self.kappa_bg)


class KPP:
    """KPP boundary-layer overlay.

    ``KPP(ri_crit=0.3, cs_nonlocal=6.3, c_vt2=1.8, shortwave_method="mxl_sw",
    enabled=True)``.  Only ``shortwave_method="mxl_sw"`` with no penetrating
    shortwave is supported (the method is then inert, rdb_ocean_vmix.F90:343-348).
    """

    def __init__(self, ri_crit=0.3, cs_nonlocal=6.3, c_vt2=1.8, shortwave_method="mxl_sw", enabled=True):
        self.ri_crit, self.cs_nonlocal, self.c_vt2 = ri_crit, cs_nonlocal, c_vt2
        self.shortwave_method = shortwave_method
        self.enabled = bool(enabled)

    def apply(self, config):
        config.use_kpp = 1 if self.enabled else 0
        if not self.enabled:
            return
        if self.shortwave_method not in ("mxl_sw", "all_sw", "lv1_sw"):
            raise ValueError(f"unknown KPP shortwave method {self.shortwave_method!r}")
        config.ri_crit = float(self.ri_crit)
        config.cs_nonlocal = float(self.cs_nonlocal)
        config.c_vt2 = float(self.c_vt2)


class VmixAssembly:
    """The ``vmix_assemble`` gate: floors, ceilings and the debug guard
    (rdb_ocean_vmix.F90:352-380)."""

    def __init__(self, kv_bg=None, kt_bg=None, ks_bg=None, kv_max=None, kd_max=None, guard=False):
        self.kv_bg, self.kt_bg, self.ks_bg = kv_bg, kt_bg, ks_bg
        self.kv_max, self.kd_max, self.guard = kv_max, kd_max, bool(guard)

    def apply(self, config):
        for name in ("kv_bg", "kt_bg", "ks_bg", "kv_max", "kd_max"):
            val = getattr(self, name)
            if val is not None:
                setattr(config, name, float(val))
        config.vmix_guard = 1 if self.guard else 0
