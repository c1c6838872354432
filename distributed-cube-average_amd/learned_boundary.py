"""Learned absorbing layer: training data from the propagator, training on the
GPU, and the reflection score against a padded-domain reference run.

The network and its kernels live in the library (include/dc_hip.h, "learned
absorbing layer"); this module drives them:

* `training_case` builds a padded domain (a smooth random TTI model with the
  reference's random-velocity shell far outside) and a small domain, a window
  of it whose outer `width` cells form the learned layer;
* `collect` steps the padded domain with a Ricker source and, every few
  steps, lets dcx_learned_sample_step record, at random layer cells of the
  small domain, the network input and the correction that makes the small
  domain's step (which sees zeros beyond its edge) equal the padded one's;
* `train` runs Adam on those samples with the library's MFMA forward and
  backward kernels;
* `reflection` runs one source in four ways -- the padded reference run, the small
  domain with the learned layer, the small domain with nothing at its edge,
  and the small domain whose outer cells are the reference's random-velocity
  shell (src/boundary.c:7-104) -- and scores each by the energy of its
  difference from the padded run at interior receivers.

The weights that ship with the package (`learned_layer_w6.npz`) were made by
`python -m distributed-cube-average_amd.learned_boundary train`.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import dc_hip as H
from .velocity_model import STENCIL, cfl_dt

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_layer_w6.npz")
NPARAM = 64 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def init_params(seed=0) -> np.ndarray:
    """He-normal hidden layers, a small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((64, 64)) * np.sqrt(2.0 / 64)
    W2 = rng.standard_normal((64, 64)) * np.sqrt(2.0 / 64)
    W3 = rng.standard_normal((2, 64)) * 0.01
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64),
                           W3.ravel(), np.zeros(2)]).astype(np.float32)


def save_weights(path, params, meta: dict):
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), {k: z[k] for k in z.files if k != "params"}


def _smooth_noise(rng, shape, coarse=10):
    """Seeded noise, smooth over ~shape/coarse cells, in [0, 1]."""
    from scipy.ndimage import zoom
    c = rng.standard_normal((coarse,) * 3)
    f = zoom(c, [s / coarse for s in shape], order=1, mode="nearest")
    f = f[:shape[0], :shape[1], :shape[2]]
    return ((f - f.min()) / (f.max() - f.min() + 1e-30)).astype(np.float32)


def random_tti(rng, shape):
    """Smooth random TTI model: vpz in [1500, 4500] m/s, vsv = 0.5 vpz;
    epsilon U[0,0.3], delta U[0, min(eps, 0.2)], theta U[0, pi/3], phi
    U[0, pi] per cell (the anisotropy of the other workloads)."""
    vpz = (1500.0 + 3000.0 * _smooth_noise(rng, shape)).astype(np.float32)
    vsv = (0.5 * vpz).astype(np.float32)
    eps = rng.uniform(0.0, 0.3, shape).astype(np.float32)
    dlt = (rng.uniform(0.0, 1.0, shape) * np.minimum(eps, 0.2)).astype(np.float32)
    theta = rng.uniform(0.0, np.pi / 3, shape).astype(np.float32)
    phi = rng.uniform(0.0, np.pi, shape).astype(np.float32)
    return vpz, vsv, eps, dlt, theta, phi


@dataclass
class Case:
    """A padded domain and the small domain cut from its middle."""
    coef_pad: list
    pad_sizes: tuple      # array sizes (sx, sy, sz)
    model: tuple          # vpz, vsv, eps, dlt, theta, phi of the padded arrays
    origin: tuple         # padded cell of small's compute (0, 0, 0)
    inner: int            # small compute extent per axis
    width: int
    d: float
    dt: float
    src: tuple            # source cell in padded array coordinates
    fcut: float

    def small_coef(self):
        o = self.origin
        n = self.inner + 2 * STENCIL
        sl = (slice(o[2] - STENCIL, o[2] - STENCIL + n),
              slice(o[1] - STENCIL, o[1] - STENCIL + n),
              slice(o[0] - STENCIL, o[0] - STENCIL + n))
        return [np.ascontiguousarray(c[sl]) for c in self.coef_pad]

    def sponge_coef(self, absorb, extra=0):
        """The small domain grown by `extra` cells per side whose outer
        `absorb` compute cells hold the reference's random-velocity shell."""
        o = self.origin
        n = self.inner + 2 * extra + 2 * STENCIL
        lo = [o[a] - STENCIL - extra for a in range(3)]
        sl = (slice(lo[2], lo[2] + n), slice(lo[1], lo[1] + n),
              slice(lo[0], lo[0] + n))
        vpz, vsv, eps, dlt, theta, phi = [np.ascontiguousarray(m[sl]) for m in self.model]
        problem = (self.inner + 2 * extra - 2 * absorb,) * 3
        H.random_boundary((n, n, n), (n, n, n), (0, 0, 0), problem, STENCIL,
                          absorb, vpz, vsv, 0)
        return H.precompute(vpz, vsv, eps, dlt, phi, theta)

    def pad_index(self, x, y, z):
        sx, sy, _ = self.pad_sizes
        return int(x + y * sx + z * sx * sy)

    def small_index(self, x, y, z):
        """Linear index in the small domain's arrays of padded cell (x,y,z)."""
        n = self.inner + 2 * STENCIL
        o = [a - STENCIL for a in self.origin]
        return int((x - o[0]) + (y - o[1]) * n + (z - o[2]) * n * n)


def make_case(seed, inner=64, width=6, margin=48, d=20.0, cfl_frac=0.4,
              hz=(10.0, 25.0), src_margin=10, pad_absorb=24, src_depth=None):
    """Padded compute region inner + 2 * margin per axis (the reference's
    random-velocity shell, pad_absorb cells, on its outer faces); the source
    is a Ricker of random peak frequency at a random cell of the small
    domain's interior (at least width + src_margin cells from its edge), or,
    with src_depth, src_depth cells below the layer of the small domain's
    low-z face, centred in x and y."""
    rng = np.random.default_rng(seed)
    n = inner + 2 * margin + 2 * STENCIL
    shape = (n, n, n)
    model = random_tti(rng, shape)
    vpz, vsv = model[0].copy(), model[1].copy()
    problem = (inner + 2 * margin - 2 * pad_absorb,) * 3
    H.random_boundary((n, n, n), (n, n, n), (0, 0, 0), problem, STENCIL,
                      pad_absorb, vpz, vsv, 0)
    coef = H.precompute(vpz, vsv, model[2], model[3], model[5], model[4])
    dt = float(np.float32(cfl_frac * cfl_dt(coef[6], d, d, d)))
    o = STENCIL + margin
    lo, hi = width + src_margin, inner - width - src_margin
    src = tuple(int(o + rng.integers(lo, hi)) for _ in range(3))
    fcut = 3.0 * float(rng.uniform(*hz))
    if src_depth is not None:
        src = (o + inner // 2, o + inner // 2, o + width + src_depth)
    return Case(coef, (n, n, n), model, (o, o, o), inner, width, d, dt, src, fcut)


def collect(samples: H.Samples, seeds, steps=700, first=60, every=4,
            cells_per_step=1024, mlp=None, rollout=False, taper=False,
            wall=False, **case_kw):
    """Training samples from the padded runs of the given model seeds.
    rollout: the small domain runs on its own state beside the padded one
    (with `mlp` attached when given) and the samples are its states with the
    corrections that pull its steps back to the padded run's.  wall: the
    small domain runs inside the one-way wall (rollout only: a
    teacher-forced step leaves its frame at rest)."""
    if wall and not rollout:
        raise ValueError("the one-way wall is collected in rollout only")
    t0 = time.time()
    for k, seed in enumerate(seeds):
        c = make_case(seed, **case_kw)
        pad = H.DeviceData(c.coef_pad, None, c.d, c.d, c.d, c.dt)
        pad.release_host()
        small = H.DeviceData(c.small_coef(), None, c.d, c.d, c.d, c.dt)
        small.release_host()
        if rollout and mlp is not None:
            H.set_learned_layer(small, mlp, c.width)
        if wall:
            H.set_wall(small, H.WALL_ONE_WAY)
        amp = H.ricker_table(c.dt, steps + 1, c.fcut)
        si = c.pad_index(*c.src)
        ss = c.small_index(*c.src)
        pad.add_source(si, float(amp[0]))
        if rollout:
            small.add_source(ss, float(amp[0]))
        for it in range(steps):
            if it >= first and (it - first) % every == 0:
                H.learned_sample_step(pad, small, c.origin, c.width,
                                      cells_per_step, seed * 100003 + it,
                                      samples, si, float(amp[it + 1]), True,
                                      rollout=rollout, taper=taper)
            else:
                pad.step(si, float(amp[it + 1]), True)
                if rollout:
                    small.step(ss, float(amp[it + 1]), True)
        pad.synchronize()
        small.synchronize()
        pad.close()
        small.close()
        if k % 8 == 7 or k == len(seeds) - 1:
            log(f"[learned] collected {samples.count()} samples from "
                f"{k + 1} models ({time.time() - t0:.0f}s)")


def train(samples: H.Samples, iters=6000, batch=4096, lr=2e-3, seed=0,
          params=None, chunk=500, mode=H.MLP_FREE, sigma_max=0.0):
    """Adam with a step decay of the learning rate; returns (mlp, losses)."""
    mlp = H.Mlp(init_params(seed) if params is None else params, mode=mode,
                sigma_max=sigma_max)
    losses = []
    t0 = time.time()
    done = 0
    while done < iters:
        n = min(chunk, iters - done)
        rate = lr * (0.1 ** (done / iters * 2.0))
        losses.append(mlp.train(samples, n, batch, rate, seed + done))
        done += n
        log(f"[learned] iter {done}: loss {float(np.mean(losses[-1][-50:])):.4e} "
            f"(lr {rate:.2e}, {time.time() - t0:.0f}s)")
    return mlp, np.concatenate(losses)


def _receivers(c: Case, gap=4, stride=4):
    """Interior receivers: cells `gap` inside the learned layer on the six
    faces of the small domain's interior, every `stride` cells, in padded
    array coordinates."""
    o, n, w = c.origin, c.inner, c.width
    a, b = w + gap, n - 1 - w - gap
    pts = set()
    rng = range(a, b + 1, stride)
    for u in rng:
        for v in rng:
            for fixed in (a, b):
                pts.add((fixed, u, v))
                pts.add((u, fixed, v))
                pts.add((u, v, fixed))
    return np.array(sorted(pts), np.int64) + np.array(o, np.int64)


def reflection(mlp, c: Case, steps, sponge_widths=(6, 32), graph=True,
               wall=False):
    """Energy ratio of (run - padded) to padded at the interior receivers over
    `steps` steps, for the learned layer, no layer and random-velocity
    shells of the given widths (shells wider than the layer grow the domain
    outward so that the interior, the receivers and the source stay put).
    With wall, also the thin domain inside the first- and second-order
    one-way walls alone ("wall", "wall2") and with the learned layer inside
    them ("learned_wall", "learned_wall2")."""
    rec = _receivers(c)
    amp = H.ricker_table(c.dt, steps, c.fcut)

    def run(coef, lo):
        """Run a domain whose array (0,0,0) is padded cell lo; record p at
        the receivers."""
        sz = coef[0].shape
        sx, sy = sz[2], sz[1]
        idx = (rec[:, 0] - lo[0]) + (rec[:, 1] - lo[1]) * sx + \
            (rec[:, 2] - lo[2]) * sx * sy
        s = c.src
        si = (s[0] - lo[0]) + (s[1] - lo[1]) * sx + (s[2] - lo[2]) * sx * sy
        return coef, idx, int(si)

    out = {}
    # Reference: the padded domain.
    d = H.DeviceData(c.coef_pad, None, c.d, c.d, c.d, c.dt)
    _, idx, si = run(c.coef_pad, (0, 0, 0))
    padded, _ = d.run_traces(steps, si, amp, idx, graph=graph)
    d.close()
    ej = float(np.sum(padded.astype(np.float64) ** 2))
    lo_small = tuple(o - STENCIL for o in c.origin)

    def score(coef, lo, net, wall_mode=H.WALL_ZERO):
        dd = H.DeviceData(coef, None, c.d, c.d, c.d, c.dt)
        if net is not None:
            H.set_learned_layer(dd, net, c.width)
        H.set_wall(dd, wall_mode)
        _, idx2, si2 = run(coef, lo)
        tr, ms = dd.run_traces(steps, si2, amp, idx2, graph=graph)
        n = coef[0].shape
        log(f"[learned]   {'learned' if net is not None else 'plain'} "
            f"{n[2]}x{n[1]}x{n[0]}: {ms / steps:.3f} ms/step")
        dd.close()
        if not np.all(np.isfinite(tr)):
            return float("inf")
        return float(np.sum((tr.astype(np.float64) - padded) ** 2) / ej)

    small = c.small_coef()
    if mlp is not None:
        out["learned"] = score(small, lo_small, mlp)
    out["none"] = score(small, lo_small, None)
    if wall:
        for tag, mode in (("wall", H.WALL_ONE_WAY), ("wall2", H.WALL_ONE_WAY2)):
            out[tag] = score(small, lo_small, None, mode)
            if mlp is not None:
                out["learned_" + tag] = score(small, lo_small, mlp, mode)
    for w in sponge_widths:
        extra = max(0, w - c.width)
        coef = c.sponge_coef(w, extra)
        lo = tuple(o - STENCIL - extra for o in c.origin)
        out[f"sponge{w}"] = score(coef, lo, None)
    return out


_K8 = np.array([-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0])


def damping_layer_1d(freq, c, width, sigma_max, power, d=20.0, wall=0):
    """What a damping layer alone can absorb, in one dimension at normal
    incidence: u_tt = c^2 u_xx - sigma(x) u_t with the 8th-order second
    difference, leapfrog with the damping centred in time, dt 0.4 of the
    1-D limit 2 d / (c sqrt(6.5)).  A Ricker of peak `freq` starts at cell
    100; the layer, sigma = sigma_max ((k+1)/width)^power at its k-th cell,
    starts at cell 404 with zeros beyond it; the receiver is at cell 200.
    Returns the energy of (trace - reference trace) over the reference's,
    the reference having its wall 4000 cells away, over the time the pulse
    takes to travel 700 cells: the layer's first echo (508 cells) arrives
    and nothing has bounced twice (about 1300 cells).  With no layer the
    ratio is one full echo over two pulses of the same energy (the direct
    one and the echo off the fixed left end): 0.5 where the grid resolves
    the pulse.  At 3 cells per wavelength (25 Hz at 1500 m/s) dispersion
    spreads the pulse past the window and the no-layer ratio is 0.36.
    wall = 1 or 2: the cells beyond the layer are filled each step by the
    first- or second-order one-way condition of dcx_set_wall (in float64)
    instead of held at zero."""
    dt = 0.4 * 2.0 * d / (c * np.sqrt(6.5))
    steps = int(700.0 / (c * dt / d))
    t = np.arange(steps) * dt - 1.2 / freq
    a = (np.pi * freq * t) ** 2
    src = (1.0 - 2.0 * a) * np.exp(-a) * dt * dt

    def trace(n_int, w):
        n = n_int + w + 8
        sig = np.zeros(n)
        k = np.arange(w)
        sig[n_int + 4 + k] = sigma_max * ((k + 1.0) / w) ** power if w else 0.0
        s = sig[4:-4] * dt / 2.0
        up, uc = np.zeros(n), np.zeros(n)
        out = np.empty(steps)
        r = (c * dt - d) / (c * dt + d)
        L = n_int + 4 + w  # first cell beyond the layer

        def pred(a, b, c_):
            return a + r * (b - c_)

        for it in range(steps):
            lap = _K8[0] * uc[4:-4]
            for j in range(1, 5):
                lap = lap + _K8[j] * (uc[4 + j:n - 4 + j] + uc[4 - j:n - 4 - j])
            un = np.zeros(n)
            un[4:-4] = (2.0 * uc[4:-4] - (1.0 - s) * up[4:-4] +
                        (c * dt / d) ** 2 * lap) / (1.0 + s)
            un[100] += src[it]
            un[L:] = 0.0
            if wall and n_int < 4000:
                for F in range(L, n):
                    p1 = pred(uc[F - 1], un[F - 1], uc[F])
                    if wall == 2:
                        v_in1 = un[F - 1] - pred(uc[F - 2], un[F - 2], uc[F - 1])
                        v_in = uc[F - 1] - pred(up[F - 2], uc[F - 2], up[F - 1])
                        v_fn = uc[F] - pred(up[F - 1], uc[F - 1], up[F])
                        p1 += pred(v_in, v_in1, v_fn)
                    un[F] = p1
            up, uc = uc, un
            out[it] = uc[200]
        return out

    ref = trace(4000, 0)
    got = trace(400, width)
    return float(np.sum((got - ref) ** 2) / np.sum(ref ** 2))


def damping_bound(width, freqs=(10.0, 15.0, 20.0, 25.0),
                  speeds=(1500.0, 3000.0, 4500.0)):
    """The best damping_layer_1d over power-law profiles (sigma_max 25-1600
    /s, powers 1-3) for each frequency and speed: the floor a `width`-cell
    damping layer cannot go below at normal incidence."""
    out = []
    for f in freqs:
        for c in speeds:
            best = min((damping_layer_1d(f, c, width, sm, p), sm, p)
                       for sm in (25, 50, 100, 200, 400, 800, 1600)
                       for p in (1, 2, 3))
            out.append({"hz": f, "speed": c, "ratio": best[0],
                        "sigma_max": best[1], "power": best[2]})
            log(f"[learned] {width}-cell damping, {f:.0f} Hz, {c:.0f} m/s: "
                f"best {best[0]:.3e} (sigma_max {best[1]} /s, power {best[2]})")
    return out


def cost(n=384, width=6, steps=50, seed=0, walls=False):
    """Device ms per step of an n^3-compute domain (random smooth TTI model)
    without and with the learned layer: (plain, learned, layer cells); with
    walls, also a dict of ms per step inside the first- and second-order
    one-way walls (no layer)."""
    rng = np.random.default_rng(seed)
    shape = (n + 2 * STENCIL,) * 3
    vpz, vsv, eps, dlt, theta, phi = random_tti(rng, shape)
    coef = H.precompute(vpz, vsv, eps, dlt, phi, theta)
    del vpz, vsv, eps, dlt, theta, phi
    dt = float(np.float32(0.4 * cfl_dt(coef[6], 20.0, 20.0, 20.0)))
    d = H.DeviceData(coef, None, 20.0, 20.0, 20.0, dt)
    d.release_host()
    src = (n // 2 + STENCIL) * (1 + shape[0] + shape[0] ** 2)
    amp = H.ricker_table(dt, steps, 45.0)
    d.run(5, src, amp[:5], graph=True)
    plain = d.run(steps, src, amp, graph=True) / steps
    wall_ms = {}
    if walls:
        for tag, mode in (("wall", H.WALL_ONE_WAY), ("wall2", H.WALL_ONE_WAY2)):
            H.set_wall(d, mode)
            d.run(5, src, amp[:5], graph=True)
            wall_ms[tag] = d.run(steps, src, amp, graph=True) / steps
        H.set_wall(d, H.WALL_ZERO)
    params, meta = load_weights()
    mlp = H.Mlp(params, mode=int(meta["mode"]), sigma_max=float(meta["sigma_max"]))
    ncell = H.set_learned_layer(d, mlp, width)
    d.run(5, src, amp[:5], graph=True)
    learned = d.run(steps, src, amp, graph=True) / steps
    d.close()
    mlp.close()
    return (plain, learned, ncell, wall_ms) if walls else (plain, learned, ncell)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("cmd", choices=["train", "eval", "cost", "bound"])
    ap.add_argument("--models", type=int, default=64)
    ap.add_argument("--inner", type=int, default=64)
    ap.add_argument("--margin", type=int, default=48)
    ap.add_argument("--width", type=int, default=6)
    ap.add_argument("--steps", type=int, default=700)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[9001, 9002])
    ap.add_argument("--eval-steps", type=int, default=900)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--mode", choices=["free", "damping"], default="damping")
    ap.add_argument("--sigma-max", type=float, default=0.5)
    ap.add_argument("--loop-passes", type=int, default=0,
                    help="closed-loop collection passes after the first")
    ap.add_argument("--loop-iters", type=int, default=3000)
    ap.add_argument("--eval-margin", type=int, default=96)
    ap.add_argument("--taper", type=int, default=1)
    ap.add_argument("--src-depth", type=int, default=None,
                    help="evaluation source this many cells below the low-z layer")
    ap.add_argument("--wall", action="store_true",
                    help="also score the one-way wall, alone and around the "
                         "learned layer")
    ap.add_argument("--quick", action="store_true",
                    help="evaluate without the random-velocity shells")
    ap.add_argument("--out-bias", type=float, default=-3.0,
                    help="train --wall: initial output bias (damping near "
                         "zero at the start)")
    ap.add_argument("--eval-sigma", type=float, default=None,
                    help="eval: the damping cap to run the shipped network "
                         "with (default: the one it was trained with)")
    args = ap.parse_args(argv)
    kw = dict(inner=args.inner, width=args.width, margin=args.margin)
    ekw = dict(inner=args.inner, width=args.width, margin=args.eval_margin,
               src_depth=args.src_depth)
    mode = H.MLP_DAMPING if args.mode == "damping" else H.MLP_FREE
    sig = args.sigma_max if mode == H.MLP_DAMPING else 0.0

    def evaluate(mlp, tag):
        scores = {}
        for s in args.eval_seeds:
            c = make_case(s, **ekw)
            r = reflection(mlp, c, args.eval_steps, wall=args.wall,
                           sponge_widths=(6, 32) if not args.quick else ())
            log(f"[learned] {tag} eval seed {s} (source {c.fcut / 3.0:.1f} Hz): " +
                ", ".join(f"{k} {v:.3e}" for k, v in r.items()))
            del c
            for k, v in r.items():
                scores.setdefault(k, []).append(v)
        return scores

    if args.cmd == "bound":
        damping_bound(args.width)
        for order in (1, 2):
            for f in (10.0, 15.0, 20.0, 25.0):
                for c in (1500.0, 3000.0, 4500.0):
                    log(f"[learned] {order}-order one-way wall, no layer, "
                        f"{f:.0f} Hz, {c:.0f} m/s: "
                        f"{damping_layer_1d(f, c, 0, 0.0, 1, wall=order):.3e}")
        return
    if args.cmd == "cost":
        plain, learned, ncell, wall_ms = cost(args.inner, args.width, walls=True)
        for tag, ms in wall_ms.items():
            log(f"[learned] cost {args.inner}^3: plain {plain:.3f} ms/step, inside "
                f"the {'second' if tag == 'wall2' else 'first'}-order one-way wall "
                f"{ms:.3f} ms/step")
        log(f"[learned] cost {args.inner}^3: plain {plain:.3f} ms/step, with the "
            f"{args.width}-cell layer ({ncell} cells) {learned:.3f} ms/step; layer "
            f"{learned - plain:.3f} ms = {ncell / (learned - plain) / 1e6:.2f} Gcell/s")
        return
    if args.cmd == "train" and args.wall:
        # Inside the one-way wall a teacher-forced step has nothing to learn
        # (its frame is at rest, not the wall's), so every pass runs the
        # small domain on its own state with the current network and fits
        # the corrections that pull it back to the padded run.
        per_model = ((args.steps - 60 + 3) // 4) * args.cells
        samples = H.Samples(per_model * args.models * (1 + args.loop_passes) + 1)
        params = init_params(0)
        params[-2:] = args.out_bias
        mlp = H.Mlp(params, mode=mode, sigma_max=sig)
        for r in range(1 + args.loop_passes):
            collect(samples, [2000 * r + k for k in range(args.models)],
                    steps=args.steps, cells_per_step=args.cells, mlp=mlp,
                    rollout=True, taper=False, wall=True, **kw)
            prev = mlp.params()
            mlp.close()
            mlp, losses = train(samples, iters=args.iters if r == 0 else
                                args.loop_iters, batch=args.batch,
                                lr=args.lr if r == 0 else args.lr * 0.5,
                                seed=r, params=prev, mode=mode, sigma_max=sig)
            scores = evaluate(mlp, f"wall pass {r}")
        save_weights(args.out, mlp.params(), dict(
            width=args.width, inner=args.inner, margin=args.margin,
            models=args.models, steps=args.steps, iters=args.iters,
            loop_passes=args.loop_passes, loop_iters=args.loop_iters,
            batch=args.batch, lr=args.lr, samples=samples.count(),
            mode=mode, sigma_max=sig, eval_margin=args.eval_margin, taper=0,
            wall=1, out_bias=args.out_bias,
            eval_steps=args.eval_steps, eval_seeds=np.array(args.eval_seeds),
            final_loss=float(np.mean(losses[-100:])),
            **{f"eval_{k}": np.array(v) for k, v in scores.items()}))
        log(f"[learned] wrote {args.out}")
        return
    if args.cmd == "train":
        per_model = ((args.steps - 60 + 3) // 4) * args.cells
        samples = H.Samples(per_model * args.models * (1 + args.loop_passes) + 1)
        seeds = list(range(args.models))
        collect(samples, seeds, steps=args.steps, cells_per_step=args.cells,
                taper=bool(args.taper), **kw)
        mlp, losses = train(samples, iters=args.iters, batch=args.batch,
                            lr=args.lr, mode=mode, sigma_max=sig)
        scores = evaluate(mlp, "teacher-forced")
        for r in range(args.loop_passes):
            # Closed loop: the small domain runs with the current network.
            collect(samples, [1000 * (r + 1) + k for k in seeds],
                    steps=args.steps, cells_per_step=args.cells, mlp=mlp,
                    rollout=True, taper=bool(args.taper), **kw)
            mlp, losses = train(samples, iters=args.loop_iters,
                                batch=args.batch, lr=args.lr * 0.5,
                                seed=r + 1, params=mlp.params(), mode=mode,
                                sigma_max=sig)
            scores = evaluate(mlp, f"closed-loop pass {r + 1}")
        params = mlp.params()
        save_weights(args.out, params, dict(
            width=args.width, inner=args.inner, margin=args.margin,
            models=args.models, steps=args.steps, iters=args.iters,
            loop_passes=args.loop_passes, loop_iters=args.loop_iters,
            batch=args.batch, lr=args.lr, samples=samples.count(),
            mode=mode, sigma_max=sig, eval_margin=args.eval_margin,
            taper=args.taper,
            eval_steps=args.eval_steps, eval_seeds=np.array(args.eval_seeds),
            final_loss=float(np.mean(losses[-100:])),
            **{f"eval_{k}": np.array(v) for k, v in scores.items()}))
        log(f"[learned] wrote {args.out}")
    else:
        params, meta = load_weights(args.out)
        sig = float(meta["sigma_max"]) if args.eval_sigma is None else args.eval_sigma
        mlp = H.Mlp(params, mode=int(meta["mode"]), sigma_max=sig)
        evaluate(mlp, f"shipped (sigma_max {sig})")


if __name__ == "__main__":
    main()
