"""numpy oracle of the learned closure's guard (which columns fall back to
the plain PP81 + KPP path).

TEST INFRASTRUCTURE ONLY (the checker of the guard flags the feature kernel
computes).  The guard has no counterpart in the reference; it is restated
here from its definition:

* static instability, N^2 < 0 with N^2 = -g (rho_upper - rho_lower) /
  (rho0 dz) (the sign of rdb_ocean_vmix.F90:1050-1056): the upper of two
  adjacent live layers (h > H_VANISHED) is denser than the lower one, with
  rho from the Wright EOS (the oracle's FP64 restatement);
    - strict mode: at any such pair of the column,
    - surface mode: at the pair formed by the first live layer below the
      surface and the layer under it;
* inputs out of range: a feature outside [fmin, fmax] of the training set
  (checked here only for features that are plain single-precision copies
  of a column input, such as Q_heat).
"""

from __future__ import annotations

import numpy as np

from . import oracle as O

HV = 1.5e-4
STRICT, SURFACE, RANGE = 0, 1, 2


def unstable_pairs(h, T, S, rho0=1035.0, p_ref=0.0):
    """Boolean (nx, ny, nz-1): pair (k, k+1) live and rho[k+1] > rho[k]."""
    rho = O.eos_wright(np.asfortranarray(h), np.asfortranarray(S), np.asfortranarray(T), rho0, p_ref)
    live = h > HV
    return live[:, :, :-1] & live[:, :, 1:] & (rho[:, :, 1:] > rho[:, :, :-1])


def stability_flags(mode, h, T, S, rho0=1035.0, p_ref=0.0):
    """Per column (nx, ny): the N^2 part of the guard for ``mode``."""
    nx, ny, nz = h.shape
    if mode == RANGE:
        return np.zeros((nx, ny), dtype=bool)
    un = unstable_pairs(h, T, S, rho0, p_ref)
    if mode == STRICT:
        return un.any(axis=2)
    live = h > HV
    out = np.zeros((nx, ny), dtype=bool)
    for i in range(nx):
        for j in range(ny):
            ks = np.nonzero(live[i, j])[0]
            if len(ks) >= 2:
                top = ks[-1]
                if top >= 1 and live[i, j, top - 1]:
                    out[i, j] = bool(un[i, j, top - 1])
    return out


def range_flags(values, lo, hi):
    """A column input compared in single precision with its feature range."""
    v = np.asarray(values, dtype=np.float32)
    return (v < np.float32(lo)) | (v > np.float32(hi))
