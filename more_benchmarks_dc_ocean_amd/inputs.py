"""Seeded synthetic Redi inputs and the binary exchange format.

A ``RediCase`` holds one ocean domain in the reference's layout: every 3-D
field is stored C-contiguous with shape ``(nz, ny, nx)`` -- byte-identical to
the Fortran column-major ``(nx, ny, nz)`` arrays of the reference state types
(redi/multilayer_cgrid_state.F90:18-28, redi/ocean_metrics.F90:14-22).  Native
layer ``k = 0`` is the bed and ``k = nz-1`` the surface
(redi/multilayer_cgrid_state.F90:23).  ``nx``/``ny`` include ``nghost`` ghost
cells on each side (redi/grid.F90:17-25).

The generators build the domains of the benchmark workloads:

* ``case_seamount``  64x64x25, linear EOS, 2 tracers, 3 Gaussian seamounts;
* ``case_front``  473x297x75, Roquet SpV EOS, 4 tracers, fractal bathymetry,
  one tanh front;
* ``case_two_fronts``  1440x1080x75, as case_front with two fronts and mesoscale noise;
* ``case_domain`` one batched/served domain of arbitrary size;
* ``case_analytic``  the reference's own benchmark state: all wet, uniform
  layers, analytic T/S, linear EOS, 2 tracers (473x297x30 by default);
* ``case_gabight``  case_front's T/S, front and EOS over the real 0.1 degree
  Great Australian Bight bathymetry the reference ships for its ePBL kernel,
  resampled to any size.

All randomness comes from ``numpy.random.default_rng`` seeded with
``(config, domain, field)`` so a case is reproducible bit-for-bit on any host
with the same numpy.  ``write_input`` / ``read_output`` implement the "REDI" /
"REDO" v1 files read and written by ``oracle/ref_driver.F90``;
``write_state_dump`` / ``read_state_dump`` the reference drivers' raw dumps.
"""
from __future__ import annotations

import dataclasses
import math
import os
import struct
from typing import List, Optional, Sequence

import numpy as np

EOS_LINEAR = 1
EOS_WRIGHT97 = 2
EOS_ROQUET_SPV = 4

MAGIC_IN = 0x49444552  # b"REDI" little-endian
MAGIC_OUT = 0x4F444552  # b"REDO"


@dataclasses.dataclass
class RediCase:
    """One Redi domain plus every scalar the reference step reads."""

    nx_phys: int
    ny_phys: int
    nz: int
    h: np.ndarray  # (nz, ny, nx) layer thickness, m
    hTr: np.ndarray  # (nt, nz, ny, nx) thickness-weighted tracers
    dy_cu: np.ndarray  # (ny, nx+1)
    dx_cv: np.ndarray  # (ny+1, nx)
    idxCu: np.ndarray  # (ny, nx+1)
    idyCv: np.ndarray  # (ny+1, nx)
    areaT: np.ndarray  # (ny, nx)
    wet_u: np.ndarray  # (ny, nx+1)
    wet_v: np.ndarray  # (ny+1, nx)
    khtr_u_ext: np.ndarray  # (ny, nx+1)
    khtr_v_ext: np.ndarray  # (ny+1, nx)
    nghost: int = 3
    idx_T: int = 1  # 1-based tracer index of temperature (0 = none)
    idx_S: int = 2
    eos_variant: int = EOS_LINEAR
    nsteps: int = 1
    walls: tuple = (1, 1, 1, 1)  # west, east, south, north (1 = OBC_WALL)
    use_ext: int = 1
    enable: int = 1
    dt: float = 900.0
    khtr: float = 1000.0
    rho0: float = 1035.0
    T_ref: float = 10.0
    S_ref: float = 35.0
    alpha_T: float = 1.7e-4
    beta_S: float = 7.6e-4
    p_ref: float = 0.0
    name: str = "case"
    land_fraction: float = 0.0

    @property
    def nx(self) -> int:
        return self.nx_phys + 2 * self.nghost

    @property
    def ny(self) -> int:
        return self.ny_phys + 2 * self.nghost

    @property
    def nt(self) -> int:
        return int(self.hTr.shape[0])

    @property
    def ns(self) -> int:
        return 2 * self.nz + 2

    @property
    def interior_cells(self) -> int:
        return self.nx_phys * self.ny_phys * self.nz

    def copy(self) -> "RediCase":
        c = dataclasses.replace(self)
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray):
                setattr(c, f.name, v.copy())
        return c

    def vanished_layers(self) -> int:
        """Wet-column layers thinner than 1e-10 m (reference H_VANISHED)."""
        wet = self.h.sum(axis=0) > 0
        return int(((self.h < 1e-10) & wet[None]).sum())


# --------------------------------------------------------------------------
# vertical grids and profiles
# --------------------------------------------------------------------------

def stretched_thickness(nz: int, total: float, h_top: float = 10.0) -> np.ndarray:
    """Top-down layer thicknesses h'_k = h_top * r**(k-1), sum == total.

    The ratio r is found by bisection; the result is rescaled so the column
    sums to ``total`` (10 m at the surface, r = 1.189682 for 25 layers over
    4000 m).
    """
    if nz == 1:
        return np.array([total], dtype=np.float64)
    if h_top * nz >= total:
        return np.full(nz, total / nz)
    lo, hi = 1.0 + 1e-12, 2.0
    while h_top * (hi ** nz - 1.0) / (hi - 1.0) < total:
        hi *= 1.5
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if h_top * (mid ** nz - 1.0) / (mid - 1.0) < total:
            lo = mid
        else:
            hi = mid
    r = 0.5 * (lo + hi)
    h = h_top * r ** np.arange(nz, dtype=np.float64)
    return h * (total / h.sum())


def _rng(config: int, domain: int, field: int) -> np.random.Generator:
    return np.random.default_rng([int(config), int(domain), int(field)])


_FFT_WORKERS = max(1, min(16, os.cpu_count() or 1))


def _smooth_noise(rng: np.random.Generator, shape, sigma: float, workers: int = 1) -> np.ndarray:
    """Unit-std Gaussian-filtered white noise (periodic FFT filter)."""
    from scipy.fft import rfft2, irfft2
    ny, nx = shape
    w = rng.standard_normal((ny, nx))
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    filt = np.exp(-2.0 * (math.pi * sigma) ** 2 * (kx * kx + ky * ky))
    f = irfft2(rfft2(w, workers=workers) * filt, s=(ny, nx), workers=workers)
    f -= f.mean()
    sd = f.std()
    return f / sd if sd > 0 else f


def _fractal_field(rng: np.random.Generator, shape) -> np.ndarray:
    """2-D field with power spectrum ~ |k|^-2 (spectral slope -2)."""
    from scipy.fft import rfft2, irfft2
    ny, nx = shape
    w = rng.standard_normal((ny, nx))
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    amp[k > 0] = 1.0 / k[k > 0]  # |F|^2 ~ k^-2
    f = irfft2(rfft2(w, workers=_FFT_WORKERS) * amp, s=(ny, nx), workers=_FFT_WORKERS)
    return (f - f.min()) / (f.max() - f.min())


def _metrics(nx: int, ny: int, dx: float, dy: float, ocean: np.ndarray, kappa: float):
    """Uniform C-grid metrics and wet masks (a face is wet iff both cells are
    ocean; the outermost faces of the array have no second cell)."""
    dy_cu = np.full((ny, nx + 1), dy)
    idxCu = np.full((ny, nx + 1), 1.0 / dx)
    dx_cv = np.full((ny + 1, nx), dx)
    idyCv = np.full((ny + 1, nx), 1.0 / dy)
    areaT = np.full((ny, nx), dx * dy)
    wet_u = np.zeros((ny, nx + 1))
    wet_u[:, 1:nx] = (ocean[:, :-1] & ocean[:, 1:]).astype(np.float64)
    wet_v = np.zeros((ny + 1, nx))
    wet_v[1:ny, :] = (ocean[:-1, :] & ocean[1:, :]).astype(np.float64)
    ku = np.full((ny, nx + 1), kappa)
    kv = np.full((ny + 1, nx), kappa)
    return dict(dy_cu=dy_cu, dx_cv=dx_cv, idxCu=idxCu, idyCv=idyCv, areaT=areaT,
                wet_u=wet_u, wet_v=wet_v, khtr_u_ext=ku, khtr_v_ext=kv)


def _columns(depth: np.ndarray, h_ref_td: np.ndarray, total: float, vanished: bool = False):
    """Native (bottom-up) thickness and layer-centre depth fields.

    z*-scaled columns: h_k = h_ref_k * depth / total.  With ``vanished``
    the columns are z-coordinate instead: layers wholly below the local
    depth have zero thickness and the layer holding the bed is cut."""
    nz = h_ref_td.shape[0]
    ny, nx = depth.shape
    h = np.empty((nz, ny, nx))
    zc = np.empty((nz, ny, nx))
    scale = depth / total  # 0 on land
    top = np.zeros((ny, nx))
    for kt in range(nz):  # top-down
        if vanished:
            hk = np.clip(depth - top, 0.0, h_ref_td[kt])
        else:
            hk = h_ref_td[kt] * scale
        k = nz - 1 - kt
        h[k] = hk
        zc[k] = top + 0.5 * hk
        top = top + hk
    return h, zc


def _layer_tracers(cfg: int, seed: int, h: np.ndarray, zc: np.ndarray, sigma: float,
                   noise_T: float, noise_S: float, fronts2d, nt: int) -> np.ndarray:
    """Thickness-weighted tracers, one native layer per pool task.

    T = 2 + 23 exp(-z/800) + noise_T xi_T - sum(front) exp(-z/1000),
    S = 34.5 + 0.8 exp(-z/500) + noise_S xi_S, tracers 3.. are U[0,1];
    xi are per-layer unit-std smoothed noise fields seeded by
    (cfg, seed, 100+k) and (cfg, seed, 300+k); U[0,1] by (cfg, seed,
    10000*t+k)."""
    from concurrent.futures import ThreadPoolExecutor
    nz, ny, nx = h.shape
    hTr = np.empty((nt, nz, ny, nx))

    def one(k):
        z, hk = zc[k], h[k]
        xiT = _smooth_noise(_rng(cfg, seed, 100 + k), (ny, nx), sigma)
        xiS = _smooth_noise(_rng(cfg, seed, 300 + k), (ny, nx), sigma)
        T = 2.0 + 23.0 * np.exp(-z / 800.0) + noise_T * xiT
        S = 34.5 + 0.8 * np.exp(-z / 500.0) + noise_S * xiS
        if fronts2d:
            e = np.exp(-z / 1000.0)
            for p in fronts2d:
                T = T - p * e
        hTr[0, k] = T * hk
        if nt > 1:
            hTr[1, k] = S * hk
        for t in range(2, nt):
            hTr[t, k] = _rng(cfg, seed, 10000 * t + k).uniform(0.0, 1.0, size=(ny, nx)) * hk

    with ThreadPoolExecutor(_FFT_WORKERS) as ex:
        list(ex.map(one, range(nz)))
    return hTr


# --------------------------------------------------------------------------
# configurations
# --------------------------------------------------------------------------

def case_seamount(nx_phys: int = 64, ny_phys: int = 64, nz: int = 25, nt: int = 2,
              nsteps: int = 10, seed: int = 0, nghost: int = 3) -> RediCase:
    """Linear-EOS seamount domain (64x64x25, dx=dy=10 km, 4000 m deep).

    Layers: 10 m at the surface growing geometrically to 4000 m total.
    Land: three seeded Gaussian seamounts; a column whose seamount height
    exceeds 60% of the depth is land (h = hTr = 0), all other columns are
    full depth.  T = 2 + 23 exp(-z/800) + 0.5 xi_T, S = 34.5 +
    0.8 exp(-z/500) + 0.05 xi_S with xi per layer, Gaussian-filtered (sigma 3
    cells) and renormalised to unit std.  Linear EOS with alpha=2e-4,
    beta=7.6e-4 fed literally into alpha_T/beta_S; kappa = 1000 m^2/s.
    """
    cfg = 0
    nx, ny = nx_phys + 2 * nghost, ny_phys + 2 * nghost
    D, dx = 4000.0, 10_000.0
    r = _rng(cfg, seed, 0)
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    height = np.zeros((ny, nx))
    for _ in range(3):
        cx = nghost + r.uniform(0.15, 0.85) * nx_phys
        cy = nghost + r.uniform(0.15, 0.85) * ny_phys
        w = r.uniform(0.06, 0.12) * min(nx_phys, ny_phys) + 1.0
        a = r.uniform(0.7, 1.0) * D
        height += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2.0 * w * w))
    ocean = ~(height > 0.6 * D)
    depth = np.where(ocean, D, 0.0)
    h_ref = stretched_thickness(nz, D)
    h, zc = _columns(depth, h_ref, D)
    hTr = _layer_tracers(cfg, seed, h, zc, 3.0, 0.5, 0.05, [], nt)
    m = _metrics(nx, ny, dx, dx, ocean, 1000.0)
    return RediCase(nx_phys=nx_phys, ny_phys=ny_phys, nz=nz, nghost=nghost,
                    h=h, hTr=hTr, eos_variant=EOS_LINEAR,
                    alpha_T=2.0e-4, beta_S=7.6e-4, khtr=1000.0, nsteps=nsteps,
                    name=f"seamount_{nx_phys}x{ny_phys}x{nz}_s{seed}",
                    land_fraction=float(1.0 - ocean[nghost:-nghost, nghost:-nghost].mean()),
                    **m)


def case_fractal(cfg: int, nx_phys: int, ny_phys: int, nz: int, nt: int,
                 nsteps: int, seed: int, land: float, fronts: Sequence[tuple],
                 noise_sigma: float, noise_T: float, noise_S: float,
                 eos_variant: int = EOS_ROQUET_SPV, nghost: int = 3,
                 dx: float = 8000.0, dy: float = 11100.0,
                 vanished: bool = False) -> RediCase:
    """Fractal-bathymetry domain (front, two-front, batched and served workloads).

    Bathymetry: seeded 2-D fractal (spectral slope -2) scaled to 0-5500 m,
    the lowest ``land`` fraction of cells is land; ocean depth is floored at
    50 m.  Layers are the 75-layer (or nz-layer) 10 m-top stretched grid over
    5500 m, z*-scaled per column (no vanished layers unless ``vanished``, in
    which case layers below the local depth collapse to zero thickness).
    T/S as case_seamount with noise (sigma, amplitude) as given plus tanh fronts
    ``(axis, position fraction, half-amplitude degC, half-width m)`` decaying
    as exp(-z/1000 m).  Tracers beyond T and S are i.i.d. U[0,1] concentrations.
    """
    nx, ny = nx_phys + 2 * nghost, ny_phys + 2 * nghost
    DMAX = 5500.0
    f = _fractal_field(_rng(cfg, seed, 1), (ny, nx))
    q = np.quantile(f, land) if land > 0 else f.min() - 1.0
    ocean = f > q
    depth = np.where(ocean, np.maximum(50.0, DMAX * (f - q) / (f.max() - q)), 0.0)
    h_ref = stretched_thickness(nz, DMAX)
    h, zc = _columns(depth, h_ref, DMAX, vanished=vanished)
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    rf = _rng(cfg, seed, 2)
    fronts2d = []
    for axis, pos, amp, width in fronts:
        jit = rf.uniform(-0.05, 0.05)
        if axis == "y":
            c = (pos + jit) * ny * dy
            prof = np.tanh((yy * dy - c) / width)
        else:
            c = (pos + jit) * nx * dx
            prof = np.tanh((xx * dx - c) / width)
        fronts2d.append(amp * prof)
    hTr = _layer_tracers(cfg, seed, h, zc, noise_sigma, noise_T, noise_S, fronts2d, nt)
    m = _metrics(nx, ny, dx, dy, ocean, 1000.0)
    return RediCase(nx_phys=nx_phys, ny_phys=ny_phys, nz=nz, nghost=nghost, h=h,
                    hTr=hTr, eos_variant=eos_variant,
                    alpha_T=2.0e-4, beta_S=7.6e-4, khtr=1000.0, nsteps=nsteps,
                    name=f"fractal{cfg}_{nx_phys}x{ny_phys}x{nz}_s{seed}",
                    land_fraction=float(1.0 - ocean[nghost:-nghost, nghost:-nghost].mean()),
                    **m)


def case_front(nx_phys: int = 473, ny_phys: int = 297, nz: int = 75, nt: int = 4,
              nsteps: int = 100, seed: int = 0, **kw) -> RediCase:
    """473x297x75 Roquet domain, ~35% land, one meridional 2 degC front over
    200 km (tanh half-width 100 km), noise as case_seamount."""
    return case_fractal(1, nx_phys, ny_phys, nz, nt, nsteps, seed, land=0.35,
                        fronts=[("y", 0.5, 1.0, 100_000.0)], noise_sigma=3.0,
                        noise_T=0.5, noise_S=0.05, **kw)


def case_two_fronts(nx_phys: int = 1440, ny_phys: int = 1080, nz: int = 75, nt: int = 4,
              nsteps: int = 50, seed: int = 0, **kw) -> RediCase:
    """1440x1080x75 Roquet domain, ~30% land, two fronts (meridional at 1/3,
    zonal at 2/3) and mesoscale noise (sigma 5 cells, 0.3 degC)."""
    return case_fractal(2, nx_phys, ny_phys, nz, nt, nsteps, seed, land=0.30,
                        fronts=[("y", 0.33, 1.0, 100_000.0), ("x", 0.67, 0.75, 100_000.0)],
                        noise_sigma=5.0, noise_T=0.3, noise_S=0.03, **kw)


def case_analytic(nx_phys: int = 473, ny_phys: int = 297, nz: int = 30, nsteps: int = 30,
                  nghost: int = 3) -> RediCase:
    """The state the reference's own Redi benchmarks time
    (legacy_testing/redi_benchmark/redi_bench.F90:60-130, 322-359; the same
    state as ideal_benchmark/wrap/rk2_redi.F90:77-137, 204-235): every face
    wet, walls on all four edges, uniform 8 km x 11.1 km metrics, uniform
    layers over 4000 m, layer-centre T = 2 + 16 exp(-z/800) + 3 fx + 3 fy and
    S = 34.5 + 0.7 (1 - exp(-z/1500)) - 0.2 fx with fx, fy the array
    position in [0, 1] (ghosts included), the linear EOS at its defaults,
    khtr = 100 m^2/s through the face arrays, dt = 900 s."""
    nx, ny = nx_phys + 2 * nghost, ny_phys + 2 * nghost
    hh = 4000.0 / nz
    z = (np.arange(nz, dtype=np.float64) + 0.5) * hh          # native k = 0 at the bed:
    z = z[::-1].copy()                                        # depth (nz - 1 - k + 0.5) hh
    fx = np.arange(nx, dtype=np.float64) / (nx - 1)
    fy = np.arange(ny, dtype=np.float64) / (ny - 1)
    zz = z[:, None, None]
    T = 2.0 + 16.0 * np.exp(-zz / 800.0) + 3.0 * fx[None, None, :] + 3.0 * fy[None, :, None]
    S = (34.5 + 0.7 * (1.0 - np.exp(-zz / 1500.0))) - 0.2 * fx[None, None, :]
    S = np.broadcast_to(S, (nz, ny, nx))
    h = np.full((nz, ny, nx), hh)
    hTr = np.stack([T * hh, S * hh])
    ocean = np.ones((ny, nx), dtype=bool)
    m = _metrics(nx, ny, 8000.0, 11100.0, ocean, 100.0)
    m["wet_u"][:] = 1.0  # the reference sets every face wet, the array's edge faces too
    m["wet_v"][:] = 1.0
    return RediCase(nx_phys=nx_phys, ny_phys=ny_phys, nz=nz, nghost=nghost, h=h, hTr=hTr,
                    eos_variant=EOS_LINEAR, khtr=100.0, nsteps=nsteps,
                    name=f"analytic_{nx_phys}x{ny_phys}x{nz}", land_fraction=0.0, **m)


GABIGHT_NX, GABIGHT_NY = 473, 297
GABIGHT_SHA256 = "2f809c4e5c3c3b9b7326106128f2b66bb01014ed115272c7d3fd9728f6b94e14"
GABIGHT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "gabight_bathy_0p1_473x297.f64.xz")


def gabight_bathymetry() -> np.ndarray:
    """The reference's real bathymetry (epbl/gabight_bathy_0p1_473x297.f64,
    read at epbl/drivers/dc_main.F90:450-466): 473 x 297 little-endian FP64
    depths in m (0 on land), Fortran order b(i, j), returned as [j, i].
    Stored xz-compressed; the decompressed bytes are checked against the
    original file's SHA-256."""
    import hashlib
    import lzma
    with open(GABIGHT_PATH, "rb") as f:
        raw = lzma.decompress(f.read())
    if hashlib.sha256(raw).hexdigest() != GABIGHT_SHA256:
        raise ValueError(f"{GABIGHT_PATH}: content does not match the reference bathymetry")
    return np.frombuffer(raw, dtype="<f8").reshape(GABIGHT_NY, GABIGHT_NX).copy()


def _nint(x: np.ndarray) -> np.ndarray:
    """Fortran NINT: nearest integer, halves away from zero."""
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def resample_bathymetry(b: np.ndarray, nx_phys: int, ny_phys: int, nghost: int) -> np.ndarray:
    """Nearest-neighbour resample of b[j, i] onto an (ny_phys + 2 nghost) x
    (nx_phys + 2 nghost) array, ghosts included, by the reference's rule
    (epbl/drivers/dc_main.F90:114-124): array cell i (1-based) takes source
    column clamp(nint(((i - nghost) - 0.5) / nx_phys * NX + 0.5), 1, NX)."""
    BNY, BNX = b.shape
    nx, ny = nx_phys + 2 * nghost, ny_phys + 2 * nghost
    i = np.arange(1, nx + 1, dtype=np.float64)
    j = np.arange(1, ny + 1, dtype=np.float64)
    bi = np.clip(_nint(((i - nghost) - 0.5) / nx_phys * BNX + 0.5), 1, BNX) - 1
    bj = np.clip(_nint(((j - nghost) - 0.5) / ny_phys * BNY + 0.5), 1, BNY) - 1
    return b[np.ix_(bj, bi)]


def case_gabight(nx_phys: int = 473, ny_phys: int = 297, nz: int = 75, nt: int = 4,
                 nsteps: int = 100, seed: int = 0, nghost: int = 3,
                 eos_variant: int = EOS_ROQUET_SPV, vanished: bool = False) -> RediCase:
    """case_front's state over the real Great Australian Bight bathymetry.

    Depth: the reference's 0.1 degree gabight file, resampled as the
    reference's ePBL driver does; a column is wet where its depth exceeds
    1 m (epbl/drivers/dc_main.F90:130), else land.  Layers: the nz-layer
    10 m-top stretched grid over the deepest column, z*-scaled per column
    (or z-coordinate with vanished layers).  T/S, the meridional front and
    the passive tracers as case_front; 8 km x 11.1 km metrics."""
    cfg = 5
    nx, ny = nx_phys + 2 * nghost, ny_phys + 2 * nghost
    dx, dy = 8000.0, 11100.0
    raw = resample_bathymetry(gabight_bathymetry(), nx_phys, ny_phys, nghost)
    ocean = raw > 1.0
    depth = np.where(ocean, raw, 0.0)
    dmax = float(depth.max())
    h_ref = stretched_thickness(nz, dmax)
    h, zc = _columns(depth, h_ref, dmax, vanished=vanished)
    yy = np.mgrid[0:ny, 0:nx][0].astype(np.float64)
    jit = _rng(cfg, seed, 2).uniform(-0.05, 0.05)
    front = 1.0 * np.tanh((yy * dy - (0.5 + jit) * ny * dy) / 100_000.0)
    hTr = _layer_tracers(cfg, seed, h, zc, 3.0, 0.5, 0.05, [front], nt)
    m = _metrics(nx, ny, dx, dy, ocean, 1000.0)
    return RediCase(nx_phys=nx_phys, ny_phys=ny_phys, nz=nz, nghost=nghost, h=h,
                    hTr=hTr, eos_variant=eos_variant,
                    alpha_T=2.0e-4, beta_S=7.6e-4, khtr=1000.0, nsteps=nsteps,
                    name=f"gabight_{nx_phys}x{ny_phys}x{nz}_s{seed}",
                    land_fraction=float(1.0 - ocean[nghost:-nghost, nghost:-nghost].mean()),
                    **m)


def case_domain(cfg: int, domain: int, side_x: int, side_y: int, nz: int,
                nt: int = 4, nsteps: int = 1) -> RediCase:
    """One independent domain of the batched / served workloads (case_front
    physics at its own size and seed)."""
    return case_fractal(cfg, side_x, side_y, nz, nt, nsteps, seed=domain, land=0.35,
                        fronts=[("y", 0.5, 1.0, 100_000.0)], noise_sigma=3.0,
                        noise_T=0.5, noise_S=0.05)


# --------------------------------------------------------------------------
# binary files shared with oracle/ref_driver.F90
# --------------------------------------------------------------------------

def write_input(path: str, c: RediCase) -> None:
    hdr = [c.nx_phys, c.ny_phys, c.nz, c.nghost, c.nt, c.idx_T, c.idx_S,
           c.eos_variant, c.nsteps, *c.walls, c.use_ext, c.enable, 0]
    rp = [c.dt, c.khtr, c.rho0, c.T_ref, c.S_ref, c.alpha_T, c.beta_S, c.p_ref]
    with open(path, "wb") as f:
        f.write(struct.pack("<2i", MAGIC_IN, 1))
        f.write(struct.pack("<16i", *hdr))
        f.write(struct.pack("<8d", *rp))
        for a in (c.h, *[c.hTr[t] for t in range(c.nt)], c.dy_cu, c.dx_cv, c.idxCu,
                  c.idyCv, c.areaT, c.wet_u, c.wet_v, c.khtr_u_ext, c.khtr_v_ext):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_output(path: str) -> dict:
    """Read a "REDO" file: coefficients and hTr after step 1, final hTr."""
    with open(path, "rb") as f:
        buf = f.read()
    magic, ver, nx, ny, nz, nt, ns, nsteps = struct.unpack_from("<8i", buf, 0)
    if magic != MAGIC_OUT or ver != 1:
        raise ValueError(f"{path}: not a REDO v1 file")
    off = 32
    out = dict(nx=nx, ny=ny, nz=nz, nt=nt, ns=ns, nsteps=nsteps)

    def take(shape, dt):
        nonlocal off
        n = int(np.prod(shape))
        a = np.frombuffer(buf, dtype=dt, count=n, offset=off).reshape(shape).copy()
        off += n * np.dtype(dt).itemsize
        return a

    for p, fa, fb in (("u", ny, nx + 1), ("v", ny + 1, nx)):
        out[p + "PoL"] = take((ns, fa, fb), "<f8")
        out[p + "PoR"] = take((ns, fa, fb), "<f8")
        out[p + "KoL"] = take((ns, fa, fb), "<i4")
        out[p + "KoR"] = take((ns, fa, fb), "<i4")
        out[p + "hEff"] = take((ns - 1, fa, fb), "<f8")
    out["hTr1"] = np.stack([take((nz, ny, nx), "<f8") for _ in range(nt)])
    out["hTrN"] = np.stack([take((nz, ny, nx), "<f8") for _ in range(nt)])
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes")
    return out


def write_state_dump(path: str, *fields: np.ndarray) -> None:
    """The reference drivers' raw state dump: int32 nx, ny, nz, then each
    (nz, ny, nx) FP64 field in Fortran order (x fastest), little-endian,
    no record markers.  One field (the temperature tracer hTr) is
    redi/drivers/dc_main.F90:156-163's DC_DUMP file; h, hTr_T, hTr_S is
    legacy_testing/redi_benchmark/redi_bench.F90:379-394's."""
    nz, ny, nx = fields[0].shape
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, nz], dtype="<i4").tobytes())
        for a in fields:
            if a.shape != (nz, ny, nx):
                raise ValueError(f"field shape {a.shape} != {(nz, ny, nx)}")
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_state_dump(path: str) -> List[np.ndarray]:
    """Fields of a write_state_dump / reference dump file, each (nz, ny, nx)."""
    raw = open(path, "rb").read()
    if len(raw) < 12:
        raise ValueError(f"{path}: no header")
    nx, ny, nz = (int(v) for v in np.frombuffer(raw[:12], "<i4"))
    n = nx * ny * nz * 8
    if nx <= 0 or ny <= 0 or nz <= 0 or (len(raw) - 12) % n:
        raise ValueError(f"{path}: {len(raw) - 12} data bytes for a {nx}x{ny}x{nz} grid")
    return [np.frombuffer(raw[12 + q * n:12 + (q + 1) * n], "<f8").reshape(nz, ny, nx).copy()
            for q in range((len(raw) - 12) // n)]


def crop(c: RediCase, i0: int, j0: int, nxp: int, nyp: int) -> RediCase:
    """Sub-domain whose physical region starts at physical cell (i0, j0)
    and whose ghost ring is taken from the parent's neighbouring cells."""
    g = c.nghost
    ia, ja = i0, j0  # total-grid index of the sub-domain's first ghost cell
    nx, ny = nxp + 2 * g, nyp + 2 * g
    sl = (slice(ja, ja + ny), slice(ia, ia + nx))
    slu = (slice(ja, ja + ny), slice(ia, ia + nx + 1))
    slv = (slice(ja, ja + ny + 1), slice(ia, ia + nx))
    wet_u = c.wet_u[slu].copy(); wet_u[:, 0] = 0.0; wet_u[:, -1] = 0.0
    wet_v = c.wet_v[slv].copy(); wet_v[0, :] = 0.0; wet_v[-1, :] = 0.0
    return dataclasses.replace(
        c, nx_phys=nxp, ny_phys=nyp, h=c.h[(slice(None),) + sl].copy(),
        hTr=c.hTr[(slice(None), slice(None)) + sl].copy(),
        dy_cu=c.dy_cu[slu].copy(), idxCu=c.idxCu[slu].copy(), wet_u=wet_u,
        khtr_u_ext=c.khtr_u_ext[slu].copy(), dx_cv=c.dx_cv[slv].copy(),
        idyCv=c.idyCv[slv].copy(), wet_v=wet_v, khtr_v_ext=c.khtr_v_ext[slv].copy(),
        areaT=c.areaT[sl].copy(), name=c.name + f"_crop{i0}_{j0}_{nxp}x{nyp}")


def batch_sizes(n: int = 96, seed: int = 0, lo: int = 48, hi: int = 512,
                nzs: Sequence[int] = (25, 50, 75), cfg: int = 3) -> List[tuple]:
    """(side, nz) of the domains of a batched workload: square interior side
    log-uniform in [lo, hi] (area log-uniform in [lo^2, hi^2]), nz uniform
    over ``nzs``."""
    rng = _rng(cfg, 100_000 + seed, 0)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(int)
    nz = rng.choice(np.asarray(nzs), n)
    return [(int(s), int(k)) for s, k in zip(sides, nz)]


def case_batch(n: int = 96, seed: int = 0, nsteps: int = 1, nt: int = 4) -> List[RediCase]:
    """The batched workload: ``n`` independent domains of uneven size, each with
    its own seeded bathymetry and T/S (case_domain physics)."""
    return [case_domain(3, seed * 1000 + i, s, s, nz, nt=nt, nsteps=nsteps)
            for i, (s, nz) in enumerate(batch_sizes(n, seed))]
