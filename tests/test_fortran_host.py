"""The Fortran host interface (ISO_C_BINDING module roundabout_amd/fortran/
rdb_colphys.F90) reaches every entry point it binds: without a GPU the ABI's
defaults, refused knobs and error ring; on the GPU a fused step through
Fortran equals the staged entry points bitwise, and a two-member ensemble
built through the Fortran bindings steps each member to the same bits."""

from __future__ import annotations

import subprocess

import pytest

from conftest import ROOT

EXE = ROOT / "roundabout_amd" / "fortran" / "colphys_fortran_check"


def _exe():
    if not EXE.exists():
        from roundabout_amd import build
        if build.build_fortran() is None:
            pytest.skip("no Fortran compiler in this image")
    return str(EXE)


def test_fortran_interface_cpu(gpu_lib):
    r = subprocess.run([_exe(), "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran-abi-cpu: ok" in r.stdout


@pytest.mark.gpu
def test_fortran_interface_gpu(gpu_lib):
    print("[case] Fortran host: 16x12x8, 3 fused steps vs staged entry points and a "
          "two-member ensemble", flush=True)
    r = subprocess.run([_exe(), "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran-abi-gpu: ok" in r.stdout
