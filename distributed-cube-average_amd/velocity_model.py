"""Synthetic TTI models and acquisition geometry for the benchmark workloads.

The reference ships only its default homogeneous model (src/precomp.c:156-169)
plus the random-velocity absorbing shell (src/boundary.c:7-104).  The
workloads here add seeded heterogeneous models; the absorbing shell, the
coefficient precompute and the Ricker source always go through the library's
restatements of the reference's host code (dc_hip.random_boundary,
dc_hip.precompute, dc_hip.ricker).

Conventions: arrays are float32 shaped (sz, sy, sx); the local grid is
problem + 2*absorption + 2*STENCIL per axis as in src/main.c:94-99; a Ricker
of peak frequency f uses the reference's FCUT = 3f (src/calculate_source.c).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import dc_hip

STENCIL = dc_hip.STENCIL
# |symbol| of the 8th-order second derivative at the Nyquist wavenumber:
# |K0| + 2(K1 - K2 + K3 - K4) for the weights of include/derivatives.h:34-38.
_K = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)
D2_NYQUIST = abs(_K[0]) + 2.0 * (_K[1] - _K[2] + _K[3] - _K[4])


def cfl_dt(v2px: np.ndarray, dx: float, dy: float, dz: float) -> float:
    """Isotropic leapfrog stability limit with the fastest P velocity."""
    vmax = float(np.sqrt(np.max(v2px)))
    return 2.0 / (vmax * np.sqrt(D2_NYQUIST * (1 / dx**2 + 1 / dy**2 + 1 / dz**2)))


@dataclass
class Workload:
    """Everything one run of the time loop needs."""
    name: str
    coef: list            # 10 float32 arrays, dc_precomp_vars order
    dx: float
    dy: float
    dz: float
    dt: float
    steps: int
    src_index: int        # linear index in the local arrays
    fcut: float           # reference-form Ricker cut-off (3 x peak Hz)
    receivers: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.coef[0].shape

    @property
    def sizes(self):
        s = self.coef[0].shape
        return (s[2], s[1], s[0])

    def source_table(self) -> np.ndarray:
        return dc_hip.ricker_table(self.dt, self.steps, self.fcut)

    def compute_cells(self) -> int:
        sx, sy, sz = self.sizes
        return (sx - 2 * STENCIL) * (sy - 2 * STENCIL) * (sz - 2 * STENCIL)


def _anisotropy(rng, shape):
    eps = rng.uniform(0.0, 0.3, shape).astype(np.float32)
    dlt = (rng.uniform(0.0, 1.0, shape) * np.minimum(eps, 0.2)).astype(np.float32)
    theta = rng.uniform(0.0, np.pi / 3, shape).astype(np.float32)
    phi = rng.uniform(0.0, np.pi, shape).astype(np.float32)
    return eps, dlt, theta, phi


def _index(sizes, x, y, z):
    return int(x + y * sizes[0] + z * sizes[0] * sizes[1])


def _receiver_grid(sizes, z, n, lo, hi):
    xs = np.linspace(lo[0], hi[0] - 1, n).round().astype(np.int64)
    ys = np.linspace(lo[1], hi[1] - 1, n).round().astype(np.int64)
    gx, gy = np.meshgrid(xs, ys)
    return (gx + gy * sizes[0] + z * sizes[0] * sizes[1]).ravel()


def smooth_random(seed=0, problem=64, absorption=16, d=20.0, steps=200,
                  peak_hz=20.0, cfl_frac=0.4, sigma=6.0):
    """Plain Fletcher TTI on one rank: problem^3 interior + absorbing shell.

    vpz: seeded Gaussian-filtered noise (sigma cells) mapped to
    [1500, 4500] m/s; epsilon U[0,0.3]; delta U[0, min(eps,0.2)];
    theta U[0,pi/3]; phi U[0,pi]; vsv = 0.5 vpz; the reference's random
    boundary of width `absorption`; Ricker at the centre; dt = cfl_frac*CFL;
    a 32x32 receiver grid on the first plane below the absorbing shell.
    """
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    n = problem + 2 * absorption + 2 * STENCIL
    shape = (n, n, n)
    noise = gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")
    noise = (noise - noise.min()) / (noise.max() - noise.min())
    vpz = (1500.0 + 3000.0 * noise).astype(np.float32)
    vsv = (0.5 * vpz).astype(np.float32)
    eps, dlt, theta, phi = _anisotropy(rng, shape)
    dc_hip.random_boundary((n, n, n), (n, n, n), (0, 0, 0),
                           (problem,) * 3, STENCIL, absorption, vpz, vsv, 0)
    coef = dc_hip.precompute(vpz, vsv, eps, dlt, phi, theta)
    dt = cfl_frac * cfl_dt(coef[6], d, d, d)
    sizes = (n, n, n)
    top = STENCIL + absorption
    rec = _receiver_grid(sizes, top, 32, (top, top), (n - top, n - top))
    return Workload("smooth_random", coef, d, d, d, float(np.float32(dt)),
                    steps, _index(sizes, n // 2, n // 2, n // 2), 3 * peak_hz,
                    rec, dict(seed=seed, problem=problem, absorption=absorption))


def layered_octant(seed=0, local=512, absorption=32, d=20.0, steps=1000,
                   peak_hz=15.0, cfl_frac=0.4, layers=10, noise=0.05):
    """One rank's share of a 2x2x2 split: the (0,0,0) octant.

    Global compute region (2*local)^3 = problem + 2*absorption per axis; the
    local arrays are local + 2*STENCIL per axis and carry the absorbing shell
    on their three outer faces (the reference's boundary drawn over the
    global grid, src/boundary.c).  vpz: `layers` horizontal layers from 1500
    to 4500 m/s over the global depth plus `noise` relative seeded noise;
    anisotropy as smooth_random; Ricker at the octant's inner corner region
    centre; 256x256 receivers on the first plane below the shell.
    """
    rng = np.random.default_rng(seed)
    n = local + 2 * STENCIL
    shape = (n, n, n)
    gcomp = 2 * local
    problem = gcomp - 2 * absorption
    gsize = gcomp + 2 * STENCIL
    z = np.arange(n, dtype=np.float64)
    layer = np.clip(((z - STENCIL) * layers) // gcomp, 0, layers - 1)
    v_layer = 1500.0 + (4500.0 - 1500.0) * layer / (layers - 1)
    vpz = np.empty(shape, np.float32)
    vpz[:] = v_layer[:, None, None].astype(np.float32)
    vpz *= (1.0 + noise * rng.standard_normal(shape, dtype=np.float32))
    vsv = (0.5 * vpz).astype(np.float32)
    eps, dlt, theta, phi = _anisotropy(rng, shape)
    dc_hip.random_boundary((n, n, n), (gsize,) * 3, (0, 0, 0), (problem,) * 3,
                           STENCIL, absorption, vpz, vsv, 0)
    coef = dc_hip.precompute(vpz, vsv, eps, dlt, phi, theta)
    del vpz, vsv, eps, dlt, theta, phi
    dt = cfl_frac * cfl_dt(coef[6], d, d, d)
    sizes = (n, n, n)
    top = STENCIL + absorption
    c = STENCIL + absorption + (local - absorption) // 2
    rec = _receiver_grid(sizes, top, 256, (top, top), (n - STENCIL, n - STENCIL))
    return Workload("layered_octant", coef, d, d, d, float(np.float32(dt)),
                    steps, _index(sizes, c, c, top + 16), 3 * peak_hz, rec,
                    dict(seed=seed, local=local, absorption=absorption))


def crop(w: Workload, origin, size, receiver_z=None):
    """A size^3 window of w's arrays starting at origin (x, y, z), keeping dt
    and the source (which must fall inside).  Receivers: every compute cell
    of plane receiver_z (window coordinates) when given."""
    x0, y0, z0 = origin
    sl = (slice(z0, z0 + size), slice(y0, y0 + size), slice(x0, x0 + size))
    coef = [np.ascontiguousarray(c[sl]) for c in w.coef]
    sx, sy, _ = w.sizes
    s = w.src_index
    sxg, syg, szg = s % sx, (s // sx) % sy, s // (sx * sy)
    lx, ly, lz = sxg - x0, syg - y0, szg - z0
    if not all(STENCIL <= v < size - STENCIL for v in (lx, ly, lz)):
        raise ValueError("source outside the crop's compute region")
    sizes = (size, size, size)
    rec = np.zeros(0, np.int64)
    if receiver_z is not None:
        xs = np.arange(STENCIL, size - STENCIL)
        gx, gy = np.meshgrid(xs, xs)
        rec = (gx + gy * size + receiver_z * size * size).ravel().astype(np.int64)
    return Workload(w.name + f"_crop{size}", coef, w.dx, w.dy, w.dz, w.dt,
                    w.steps, _index(sizes, lx, ly, lz), w.fcut, rec,
                    dict(w.meta, crop_origin=origin))


def random_small(shape, seed=0, d=20.0, steps=20, peak_hz=20.0, cfl_frac=0.4):
    """Unfiltered random TTI model on a small grid (parity tests)."""
    rng = np.random.default_rng(seed)
    vpz = rng.uniform(1500.0, 4500.0, shape).astype(np.float32)
    vsv = (0.5 * vpz).astype(np.float32)
    eps, dlt, theta, phi = _anisotropy(rng, shape)
    coef = dc_hip.precompute(vpz, vsv, eps, dlt, phi, theta)
    dt = cfl_frac * cfl_dt(coef[6], d, d, d)
    sizes = (shape[2], shape[1], shape[0])
    return Workload("random_small", coef, d, d, d, float(np.float32(dt)), steps,
                    _index(sizes, sizes[0] // 2, sizes[1] // 2, sizes[2] // 2),
                    3 * peak_hz)


class Survey:
    """Independent shots over one shared model (the batched workload).  The
    shared model stays in host memory; shot(k) cuts shot k's aperture and
    builds its coefficients on demand, so the shots' arrays need not all be
    held on the host at once."""

    def __init__(self, model, specs, dt, d, steps, absorption, meta):
        self.model = model      # vpz, eps, dlt, theta, phi (shared)
        self.specs = specs      # per shot: sizes, origin, source xy, peak Hz
        self.dt, self.d, self.steps = dt, d, steps
        self.absorption = absorption
        self.meta = meta

    def __len__(self):
        return len(self.specs)

    def compute_cells(self) -> int:
        return sum((sx - 2 * STENCIL) * (sy - 2 * STENCIL) * (sz - 2 * STENCIL)
                   for (sx, sy, sz), *_ in self.specs)

    def shot(self, k) -> "Workload":
        (sx, sy, sz), (x0, y0), (px, py), f = self.specs[k]
        vpz, eps, dlt, theta, phi = self.model
        sl = (slice(0, sz), slice(y0, y0 + sy), slice(x0, x0 + sx))
        v = np.ascontiguousarray(vpz[sl])
        vs = (0.5 * v).astype(np.float32)
        a = self.absorption
        problem = tuple(n - 2 * STENCIL - 2 * a for n in (sx, sy, sz))
        dc_hip.random_boundary((sx, sy, sz), (sx, sy, sz), (0, 0, 0), problem,
                               STENCIL, a, v, vs, 0)
        coef = dc_hip.precompute(v, vs, np.ascontiguousarray(eps[sl]),
                                 np.ascontiguousarray(dlt[sl]),
                                 np.ascontiguousarray(phi[sl]),
                                 np.ascontiguousarray(theta[sl]))
        top = STENCIL + a
        return Workload(f"shot{k}", coef, self.d, self.d, self.d, self.dt,
                        self.steps, _index((sx, sy, sz), px, py, top), 3 * f,
                        meta=dict(origin=(x0, y0, 0), peak_hz=f))


def layered_survey(seed=0, n_shots=64, lo=(128, 128, 128), hi=(320, 320, 256),
                   lateral=640, absorption=16, d=20.0, steps=800,
                   hz=(10.0, 25.0), cfl_frac=0.4, layers=10, noise=0.05):
    """n_shots shots on one shared layered TTI model, each over its own
    aperture.

    Shared model: lateral x lateral x (hi_z + 2*STENCIL) cells, `layers`
    horizontal layers 1500->4500 m/s over the depth with `noise` relative
    seeded noise, anisotropy as smooth_random.  Shot k: compute extents
    (nx, ny, nz) drawn uniformly from [lo, hi] per axis (seeded), an aperture
    at a seeded lateral position, starting at the surface; the reference's
    random-velocity absorbing shell (`absorption` cells) on all six faces of
    the aperture; a Ricker of peak frequency U[hz] at a random surface
    position (first plane below the shell) inside it.  One dt for all shots
    (cfl_frac of the whole model's CFL limit).
    """
    rng = np.random.default_rng(seed)
    depth = hi[2] + 2 * STENCIL
    shape = (depth, lateral, lateral)
    z = np.arange(depth, dtype=np.float64)
    layer = np.clip(((z - STENCIL) * layers) // hi[2], 0, layers - 1)
    v_layer = 1500.0 + (4500.0 - 1500.0) * layer / (layers - 1)
    vpz = np.empty(shape, np.float32)
    vpz[:] = v_layer[:, None, None].astype(np.float32)
    vpz *= (1.0 + noise * rng.standard_normal(shape, dtype=np.float32))
    eps, dlt, theta, phi = _anisotropy(rng, shape)
    vmax2 = float(np.max(vpz)) ** 2 * (1.0 + 2.0 * float(np.max(eps)))
    dt = float(np.float32(cfl_frac * cfl_dt(np.array([vmax2, 3000.0 ** 2]),
                                            d, d, d)))
    specs = []
    top = STENCIL + absorption
    for _ in range(n_shots):
        nx, ny, nz = (int(rng.integers(lo[a], hi[a] + 1)) for a in range(3))
        sx, sy, sz = nx + 2 * STENCIL, ny + 2 * STENCIL, nz + 2 * STENCIL
        x0 = int(rng.integers(0, lateral - sx + 1))
        y0 = int(rng.integers(0, lateral - sy + 1))
        px = int(rng.integers(top, sx - top))
        py = int(rng.integers(top, sy - top))
        specs.append(((sx, sy, sz), (x0, y0), (px, py),
                      float(rng.uniform(*hz))))
    return Survey((vpz, eps, dlt, theta, phi), specs, dt, d, steps, absorption,
                  dict(seed=seed, n_shots=n_shots, lo=lo, hi=hi,
                       lateral=lateral, absorption=absorption))
