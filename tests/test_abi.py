"""C-ABI library checks that need no GPU: it loads, exports every symbol the
public header declares, and its defaults/error ring behave like the reference
ABI (include/rdb_ocean.h conventions)."""

from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "rdb_colphys.h").read_text()
    return sorted(set(re.findall(r"^\s*int\s+(rdb_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(gpu_lib):
    syms = declared_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(gpu_lib, name), name


def test_python_prototypes_cover_header():
    from roundabout_amd.rdb._ffi import PROTOTYPES
    assert sorted(n for n, _, _ in PROTOTYPES) == declared_symbols()


def test_config_defaults_match_reference(gpu_lib):
    from roundabout_amd.rdb import _ffi
    c = _ffi.Config()
    assert gpu_lib.rdb_colphys_config_default(ctypes.byref(c)) == 0
    d = _ffi.default_config()
    for f, _ in _ffi.Config._fields_:
        assert getattr(c, f) == getattr(d, f), f
    # src/core/rdb_config.F90:1872-1893, rdb_eos.F90:537-554, rdb_constants.F90:44
    assert (c.pp81_nu0, c.pp81_nu_bg, c.pp81_kappa_bg, c.pp81_alpha) == (1e-2, 1e-4, 1e-5, 5.0)
    assert (c.shear2_floor, c.ri_crit, c.c_vt2, c.cs_nonlocal) == (1e-10, 0.3, 1.8, 6.3)
    assert (c.alpha_T, c.beta_S, c.rho0, c.cp) == (1.7e-4, 7.6e-4, 1035.0, 3992.0)
    assert gpu_lib.rdb_colphys_working_precision() == 8


def test_error_paths_without_device(gpu_lib):
    from roundabout_amd.rdb import _ffi
    # NULL / destroyed handles: idempotent destroy, BAD_HANDLE elsewhere
    h = ctypes.c_void_p()
    assert gpu_lib.rdb_colphys_destroy(ctypes.byref(h)) == 0
    assert gpu_lib.rdb_colphys_destroy(None) == 0
    assert gpu_lib.rdb_colphys_step(None, 1.0, 1) == _ffi.ERR_BAD_HANDLE
    assert "bad handle" in _ffi.last_error(gpu_lib)
    # refused knobs: a horizontal smoother, and the configure-time checks of
    # the contributors (rdb_ocean_setup_vmix.F90:552-719, 1200-1330)
    wet = np.ones((2, 2))
    for knob, val, msg in (("kd_smooth_iterations", 1, "kd_smooth_iterations"),
                           ("ddiff_strat_param_max", 1.0, "strat_param_max must be > 1"),
                           ("conv_prandtl", 0.0, "prandtl_conv must be > 0"),
                           ("epbl_translay_scale", 1.0, "translay_scale must be in [0,1)"),
                           ("epbl_combine", 3, "max_shear")):
        cfg = _ffi.default_config()
        setattr(cfg, knob, val)
        cfg.ddiff_enable = int(knob.startswith("ddiff"))
        cfg.conv_enable = int(knob.startswith("conv"))
        cfg.epbl_enable = int(knob.startswith("epbl"))
        st = gpu_lib.rdb_colphys_create(ctypes.byref(cfg), 2, 2, 3, 0,
                                        wet.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None,
                                        ctypes.byref(h))
        assert st == _ffi.ERR_CONFIG_VALIDATE and h.value is None, knob
        assert msg in _ffi.last_error(gpu_lib), knob
    # snprintf semantics of the ring
    buf = ctypes.create_string_buffer(4)
    n = gpu_lib.rdb_colphys_last_error(0, buf, 4)
    assert n > 3 and len(buf.value) == 3
    assert gpu_lib.rdb_colphys_last_error(10_000, buf, 4) == 0


def test_config_mirror_matches_the_c_layout(tmp_path):
    # sizeof and every field offset of rdb_colphys_config, from gcc, against
    # the ctypes mirror (the Fortran bind(c) type lists the same fields)
    import subprocess
    from roundabout_amd.rdb import _ffi
    names = [n for n, _ in _ffi.Config._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("%s %zu\\n", "{n}", offsetof(rdb_colphys_config, {n}));' for n in names)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "rdb_colphys.h"\nint main(void) {\n'
                   '  printf("sizeof %zu\\n", sizeof(rdb_colphys_config));\n' + body + "\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    assert int(out["sizeof"]) == ctypes.sizeof(_ffi.Config)
    for n in names:
        assert int(out[n]) == getattr(_ffi.Config, n).offset, n
    ftext = (ROOT / "roundabout_amd" / "fortran" / "rdb_colphys.F90").read_text().lower()
    block = ftext[ftext.index("type, bind(c), public :: rdb_colphys_config_t"):ftext.index("end type rdb_colphys_config_t")]
    fnames = [w.strip() for line in block.splitlines()[1:] if "::" in line and not line.strip().startswith("!")
              for w in line.split("::", 1)[1].split(",")]
    assert fnames == [n.lower() for n in names]


def test_missing_library_fails_loudly(tmp_path):
    from roundabout_amd.rdb import _ffi
    with pytest.raises(_ffi.LibraryNotFoundError):
        _ffi.load(tmp_path / "librdb_colphys.so")
