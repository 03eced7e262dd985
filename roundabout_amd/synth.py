"""Seeded synthetic ocean columns for the vertical-mixing workloads.

All arrays follow the reference state layout (Fortran column-major, ``i``
fastest, ``k = 0`` the bed layer, ``k = nz - 1`` the surface; see
``src/core/rdb_multilayer_state.F90:105-339``): numpy arrays of shape
``(nx, ny, nz)`` in ``order="F"``.  Face arrays are ``(nx+1, ny, nz)`` (u) and
``(nx, ny+1, nz)`` (v); wind stress faces ``(nx+1, ny)`` / ``(nx, ny+1)``.

Workloads (names used by bench.py and the tests):

* ``small``  -- 64 x 64 columns x 20 z-star layers, 10 % land by a seeded
  uniform mask, depth U(2000, 5000) m split into equal layers.
* ``tile``   -- 1440 x 1080 columns x 75 layers, smoothed Gaussian-random-field
  bathymetry (50-5500 m, 30 % land), z-level partial cells with fillers
  (h = 1e-4 m <= H_VANISHED) below the bed, so the deepest live layer index
  runs from 1 to 75.
* ``thin``   -- ``tile`` plus 5 % of wet columns carrying thin live layers
  (h in [1e-3, 1e-1] m) to stress the conditioning of the tridiagonal solve.

Land follows the reference land contract (``h = H_VANISHED``, c = 0,
``src/core/ocean/state/rdb_ocean_state_seed.F90:1842-1880``); velocity faces
touching land or a filler layer are zero.

Physical draws: T = 2 + A exp(-z/500 m) + N(0, 0.1) with
A ~ U(10, 28) per column; S = 35 + N(0, 0.3); u, v ~ N(0, 0.1) m/s per layer
and face; |tau| lognormal with median 0.08 N/m^2 split by a uniform angle;
Q_heat ~ N(-50, 100) W/m^2; Q_salt = 0; f on a 20-60 N beta-plane; passive
tracers U(0, 1).
"""

from __future__ import annotations

import dataclasses

import numpy as np

H_VANISHED = 1.5e-4
FILLER_H = 1.0e-4
OMEGA = 7.2921e-5


@dataclasses.dataclass
class Ocean:
    nx: int
    ny: int
    nz: int
    n_passive: int
    h: np.ndarray
    u: np.ndarray
    v: np.ndarray
    T: np.ndarray
    S: np.ndarray
    passive: list
    tau_x: np.ndarray
    tau_y: np.ndarray
    Q_heat: np.ndarray
    Q_salt: np.ndarray
    wet: np.ndarray
    k_top: np.ndarray
    n_live: np.ndarray  # live-layer count per column (0 on land)
    coriolis: np.ndarray
    bl_depth: np.ndarray

    @property
    def n_wet(self) -> int:
        return int((self.wet > 0.5).sum())

    def copy(self) -> "Ocean":
        kw = {}
        for f in dataclasses.fields(self):
            val = getattr(self, f.name)
            if isinstance(val, np.ndarray):
                kw[f.name] = val.copy(order="F")
            elif isinstance(val, list):
                kw[f.name] = [a.copy(order="F") for a in val]
            else:
                kw[f.name] = val
        return Ocean(**kw)


def _f3(a: np.ndarray) -> np.ndarray:
    """(nz, ny, nx) C-order draw -> (nx, ny, nz) Fortran-order view."""
    return np.asfortranarray(a.transpose(2, 1, 0))


def _grf_depth(rng: np.random.Generator, nx: int, ny: int, land_frac: float,
               dmin: float, dmax: float, corr_cells: float) -> tuple[np.ndarray, np.ndarray]:
    """Smoothed Gaussian random field -> (depth, wet) with the given land share.

    White noise filtered by a Gaussian of width ``corr_cells`` in spectral
    space; the lowest ``land_frac`` of the field is land, the rest is mapped by
    rank onto ``[dmin, dmax]`` metres.
    """
    noise = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    filt = np.exp(-0.5 * (2.0 * np.pi * corr_cells) ** 2 * (kx * kx + ky * ky))
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    flat = field.ravel()
    order = np.argsort(flat, kind="stable")
    n = flat.size
    n_land = int(round(land_frac * n))
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    wet_flat = rank >= n_land
    depth_flat = np.zeros(n)
    nw = n - n_land
    depth_flat[wet_flat] = dmin + (dmax - dmin) * (rank[wet_flat] - n_land + 0.5) / max(nw, 1)
    depth = depth_flat.reshape(ny, nx).T.copy(order="F")
    wet = wet_flat.reshape(ny, nx).T.astype(np.float64).copy(order="F")
    return depth, wet


def make_ocean(nx: int, ny: int, nz: int, seed: int = 1, *, land_frac: float = 0.1,
               topo: str = "zstar", dmin: float = 2000.0, dmax: float = 5000.0,
               n_passive: int = 2, thin_frac: float = 0.0, corr_cells: float = 12.0) -> Ocean:
    """Generate one synthetic ocean (see module docstring)."""
    rng = np.random.default_rng(seed)
    if topo == "zstar":
        wet = (rng.random((ny, nx)) >= land_frac).T.astype(np.float64).copy(order="F")
        depth = rng.uniform(dmin, dmax, size=(ny, nx)).T.copy(order="F")
    elif topo == "zlevel":
        depth, wet = _grf_depth(rng, nx, ny, land_frac, dmin, dmax, corr_cells)
    else:
        raise ValueError(f"unknown topo {topo!r}")
    wetb = wet > 0.5

    # Layer thickness, surface-first index m = nz-1-k.
    h_sf = np.empty((nx, ny, nz), order="F")  # indexed [i, j, m], m = 0 surface
    if topo == "zstar":
        h_sf[...] = (depth / nz)[:, :, None]
        n_live = np.where(wetb, nz, 0).astype(np.int32)
    else:
        dz = dmax / nz
        n_live = np.ceil(depth / dz).astype(np.int32)
        n_live = np.clip(n_live, 1, nz)
        m = np.arange(nz)[None, None, :]
        top = m * dz
        part = depth[:, :, None] - top
        # partial bottom cell of at least 20 % of a full layer
        h_live = np.minimum(dz, np.maximum(part, 0.2 * dz))
        h_sf[...] = np.where(m < n_live[:, :, None], h_live, FILLER_H)
        n_live = np.where(wetb, n_live, 0).astype(np.int32)
    h_sf[~wetb, :] = H_VANISHED
    if thin_frac > 0.0:
        cand = np.flatnonzero((n_live.ravel(order="F") >= 3))
        n_thin = int(round(thin_frac * cand.size))
        pick = rng.choice(cand, size=n_thin, replace=False)
        hflat = h_sf.reshape(nx * ny, nz, order="F")
        nl = n_live.ravel(order="F")
        for c in pick:
            # one to three thin live layers inside the column
            cnt = int(rng.integers(1, 4))
            for _ in range(cnt):
                mm = int(rng.integers(0, nl[c]))
                hflat[c, mm] = 10.0 ** rng.uniform(-3.0, -1.0)
        h_sf = hflat.reshape(nx, ny, nz, order="F")
    h = np.asfortranarray(h_sf[:, :, ::-1])  # k = 0 bed
    return _with_state(rng, h, wet, n_live, n_passive)


def _with_state(rng: np.random.Generator, h: np.ndarray, wet: np.ndarray, n_live: np.ndarray,
                n_passive: int) -> Ocean:
    """Draw T, S, passives, face velocities and forcing for a given geometry."""
    nx, ny, nz = h.shape
    wetb = wet > 0.5
    h_sf = h[:, :, ::-1]
    # Layer-centre depth from the surface (live layers only matter).
    cum = np.cumsum(h_sf, axis=2) - 0.5 * h_sf
    z_c = np.asfortranarray(cum[:, :, ::-1])
    amp = rng.uniform(10.0, 28.0, size=(ny, nx)).T
    T = 2.0 + amp[:, :, None] * np.exp(-z_c / 500.0) + 0.1 * _f3(rng.standard_normal((nz, ny, nx)))
    S = 35.0 + 0.3 * _f3(rng.standard_normal((nz, ny, nx)))
    passive = [_f3(rng.random((nz, ny, nx))) for _ in range(n_passive)]

    live = h > H_VANISHED  # (nx, ny, nz)
    u = 0.1 * _f3(rng.standard_normal((nz, ny, nx + 1)))
    v = 0.1 * _f3(rng.standard_normal((nz, ny + 1, nx)))
    # zero faces that touch land or a filler layer on either side
    lu = np.zeros((nx + 2, ny, nz), dtype=bool)
    lu[1:-1] = live
    u *= (lu[:-1] & lu[1:])
    lv = np.zeros((nx, ny + 2, nz), dtype=bool)
    lv[:, 1:-1] = live
    v *= (lv[:, :-1] & lv[:, 1:])

    T[~wetb, :] = 0.0
    S[~wetb, :] = 0.0
    for p in passive:
        p[~wetb, :] = 0.0

    mag_x = 0.08 * np.exp(0.5 * rng.standard_normal((ny, nx + 1)))
    ang_x = rng.uniform(0.0, 2.0 * np.pi, size=(ny, nx + 1))
    mag_y = 0.08 * np.exp(0.5 * rng.standard_normal((ny + 1, nx)))
    ang_y = rng.uniform(0.0, 2.0 * np.pi, size=(ny + 1, nx))
    tau_x = np.asfortranarray((mag_x * np.cos(ang_x)).T)
    tau_y = np.asfortranarray((mag_y * np.sin(ang_y)).T)
    Q_heat = np.asfortranarray(rng.normal(-50.0, 100.0, size=(ny, nx)).T)
    Q_salt = np.zeros((nx, ny), order="F")
    lat = np.deg2rad(20.0 + 40.0 * (np.arange(ny) + 0.5) / ny)
    cor = np.asfortranarray(np.broadcast_to(2.0 * OMEGA * np.sin(lat)[None, :], (nx, ny)).copy())
    k_top = np.full((nx, ny), nz, dtype=np.int32, order="F")
    bl = np.zeros((nx, ny), order="F")
    return Ocean(nx, ny, nz, n_passive, np.asfortranarray(h), np.asfortranarray(u),
                 np.asfortranarray(v), np.asfortranarray(T), np.asfortranarray(S),
                 [np.asfortranarray(p) for p in passive], tau_x, tau_y, Q_heat, Q_salt,
                 np.asfortranarray(wet), k_top, np.asfortranarray(n_live), cor, bl)


def redraw(o: Ocean, seed: int) -> Ocean:
    """Another state on the same grid: the geometry (h, wet mask, live
    layers) of ``o`` with T, S, passives, velocities and forcing drawn anew
    from the same distributions with an independent seed."""
    return _with_state(np.random.default_rng(seed), o.h.copy(order="F"), o.wet.copy(order="F"), o.n_live.copy(),
                       o.n_passive)


def crop(o: Ocean, i0: int, i1: int, j0: int, j1: int) -> Ocean:
    """Sub-block [i0, i1) x [j0, j1) of an ocean with its bounding faces.

    Every quantity of the path is column-local once the face averages are
    formed, so a crop evolves exactly like the same columns of the full grid.
    """
    def c3(a, di=0, dj=0):
        return np.asfortranarray(a[i0:i1 + di, j0:j1 + dj, ...].copy())
    return Ocean(i1 - i0, j1 - j0, o.nz, o.n_passive, c3(o.h), c3(o.u, di=1), c3(o.v, dj=1), c3(o.T), c3(o.S),
                 [c3(p) for p in o.passive], c3(o.tau_x, di=1), c3(o.tau_y, dj=1), c3(o.Q_heat), c3(o.Q_salt),
                 c3(o.wet), c3(o.k_top), c3(o.n_live), c3(o.coriolis), c3(o.bl_depth))


def small(seed: int = 1) -> Ocean:
    """64 x 64 x 20 z-star ocean, 10 % land (the CPU-runnable case)."""
    return make_ocean(64, 64, 20, seed, land_frac=0.1, topo="zstar", dmin=2000.0, dmax=5000.0)


def tile(seed: int = 1, nx: int = 1440, ny: int = 1080, nz: int = 75, thin_frac: float = 0.0) -> Ocean:
    """1440 x 1080 x 75 z-level tile with GRF bathymetry, 30 % land."""
    return make_ocean(nx, ny, nz, seed, land_frac=0.3, topo="zlevel", dmin=50.0, dmax=5500.0,
                      thin_frac=thin_frac)


def thin(seed: int = 1, nx: int = 1440, ny: int = 1080, nz: int = 75) -> Ocean:
    """``tile`` plus 5 % thin-layer columns."""
    return tile(seed, nx, ny, nz, thin_frac=0.05)


def ale_pair(o: Ocean, seed: int = 1, eta_amp: float = 1.7, jitter: float = 0.05) -> tuple[np.ndarray, np.ndarray]:
    """A source/target thickness pair for one ALE remap on the grid of `o`.

    h_old: the Lagrangian thicknesses after a step -- every live layer
    perturbed by up to +-jitter/2 and the column rescaled to carry a
    free-surface anomaly eta (a smooth field of amplitude eta_amp metres).
    h_new: the z* target of the same column total -- the rest thicknesses
    stretched by (H + eta) / H; the few ulp of mismatch the arithmetic
    leaves go into the top live target layer.  Vanished layers (bed fillers,
    land) keep their thickness in both.  Fortran-ordered (nx, ny, nz)."""
    rng = np.random.default_rng(seed)
    h0 = np.asarray(o.h)
    live = h0 > 1.5e-4
    H = h0.sum(axis=2)
    H_live = np.where(live, h0, 0.0).sum(axis=2)
    eta = (eta_amp * np.sin(np.linspace(0.0, 3.0, o.nx))[:, None]
           * np.cos(np.linspace(0.0, 2.0, o.ny))[None, :])
    eta = np.where(H_live > 0.0, eta, 0.0)
    ho = h0 * (1.0 + jitter * (rng.random(h0.shape) - 0.5))
    ho = np.where(live, ho, h0)
    ho_live = np.where(live, ho, 0.0).sum(axis=2)
    ho = np.where(live, ho * ((H_live + eta) / np.maximum(ho_live, 1e-300))[:, :, None], h0)
    hn = np.where(live, h0 * ((H + eta) / np.maximum(H, 1e-300))[:, :, None], h0)
    top = np.argmax(np.where(live, np.arange(o.nz)[None, None, :], -1), axis=2)
    ii, jj = np.meshgrid(np.arange(o.nx), np.arange(o.ny), indexing="ij")
    hn[ii, jj, top] += ho.sum(axis=2) - hn.sum(axis=2)
    return np.asfortranarray(ho), np.asfortranarray(hn)


WORKLOADS = {"small": small, "tile": tile, "thin": thin}
