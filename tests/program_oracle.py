"""The reference program's multi-rank run restated on the CPU with the oracle.

Test infrastructure: emulates src/main.c + src/coordinator.c + src/worker.c for
a given rank topology in one process (every rank's arrays in numpy, halo
exchange by box copies), using only the oracle's restated kernels.  Its
output is the global .dc layout (pc, qc) of src/coordinator.c:116-219.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

import oracle as O

S = O.STENCIL


def iterations(time_max: str, dt: str) -> int:
    """ceil of the float division, src/main.c:132."""
    return int(math.ceil(float(np.float32(float(time_max)) /
                               np.float32(float(dt)))))


def _axis_box(d, s, recv):
    """src/worker.c:196-228 (send) and 679-711 (receive)."""
    if d == -1:
        return (0, S) if recv else (S, 2 * S)
    if d == 1:
        return (s - S, s) if recv else (s - 2 * S, s - S)
    return (S, s - S)


def slabs(sizes):
    """Region calls of src/worker.c:321-402 in order."""
    sx, sy, sz = sizes
    r = S
    if not (sx >= 4 * r and sy >= 4 * r and sz >= 4 * r):
        return [((r, r, r), (sx - r, sy - r, sz - r))]
    out = []
    for side in (0, 1):
        out.append(((r if side == 0 else sx - 2 * r, r, r),
                    (2 * r if side == 0 else sx - r, sy - r, sz - r)))
    for side in (0, 1):
        out.append(((2 * r, r if side == 0 else sy - 2 * r, r),
                    (sx - 2 * r, 2 * r if side == 0 else sy - r, sz - r)))
    for side in (0, 1):
        out.append(((2 * r, 2 * r, r if side == 0 else sz - 2 * r),
                    (sx - 2 * r, sy - 2 * r, 2 * r if side == 0 else sz - r)))
    out.append(((2 * r,) * 3, (sx - 2 * r, sy - 2 * r, sz - 2 * r)))
    return out


def run_program(size, absorption, d, dt, time_max, topology=(1, 1, 1)):
    """Returns (pc, qc) global arrays shaped (gz, gy, gx)."""
    g = [s + 2 * absorption + 2 * S for s in size]
    iters = iterations(time_max, dt)
    dtf = float(np.float32(float(dt)))
    df = float(np.float32(d))
    T = topology
    part = [(g[k] - 2 * S) // T[k] for k in range(3)]
    rem = [(g[k] - 2 * S) % T[k] for k in range(3)]
    src = [g[k] // 2 for k in range(3)]
    gn = g[0] * g[1] * g[2]
    # Default model (src/precomp.c:156-169) + random boundary over the global
    # grid; a rank's share is a slice (same draw per global cell).
    m = O.default_model(gn)
    vpz = m[0].reshape(g[2], g[1], g[0])
    vsv = m[1].reshape(g[2], g[1], g[0])
    O.random_boundary(g, g, (0, 0, 0), size, S, absorption, vpz, vsv, 0)
    coefg = O.precomp(vpz, vsv, *[a.reshape(vpz.shape) for a in m[2:]])
    ranks = {}
    for c in itertools.product(range(T[0]), range(T[1]), range(T[2])):
        ls = [part[k] + (rem[k] if c[k] == T[k] - 1 else 0) + 2 * S
              for k in range(3)]
        st = [c[k] * part[k] for k in range(3)]
        sl = (slice(st[2], st[2] + ls[2]), slice(st[1], st[1] + ls[1]),
              slice(st[0], st[0] + ls[0]))
        coef = [np.ascontiguousarray(a[sl]) for a in coefg]
        inside = all(st[k] <= src[k] < st[k] + ls[k] for k in range(3))
        sidx = ((src[0] - st[0]) + (src[1] - st[1]) * ls[0] +
                (src[2] - st[2]) * ls[0] * ls[1]) if inside else -1
        f = [np.zeros((ls[2], ls[1], ls[0]), np.float32) for _ in range(4)]
        ranks[c] = dict(sizes=ls, start=st, coef=coef, src=sidx, f=f)
    for it in range(iters):
        s = np.float32(O.source_f32(dtf, it))
        for r in ranks.values():
            if r["src"] >= 0:
                r["f"][1].reshape(-1)[r["src"]] += s
                r["f"][3].reshape(-1)[r["src"]] += s
        for r in ranks.values():
            f = r["f"]
            for a, b in slabs(r["sizes"]):
                O.propagate_f32(a, b, f[0], f[1], f[2], f[3], r["coef"], df, df,
                                df, dtf)
        # Halo exchange of the new pp/qp (src/worker.c:472-501).
        msgs = []
        for c, r in ranks.items():
            for dv in itertools.product((-1, 0, 1), repeat=3):
                if dv == (0, 0, 0):
                    continue
                nc = tuple(c[k] + dv[k] for k in range(3))
                if nc not in ranks:
                    continue
                send = [_axis_box(dv[k], r["sizes"][k], False) for k in range(3)]
                a = [b[0] for b in send]
                b = [b[1] for b in send]
                msgs.append((nc, dv, O.extract_box(r["f"][0], a, b),
                             O.extract_box(r["f"][2], a, b)))
        for nc, dv, bp, bq in msgs:
            r = ranks[nc]
            recv = [_axis_box(-dv[k], r["sizes"][k], True) for k in range(3)]
            a = [b[0] for b in recv]
            b = [b[1] for b in recv]
            O.insert_box(r["f"][0], bp, a, b)
            O.insert_box(r["f"][2], bq, a, b)
        for r in ranks.values():
            f = r["f"]
            r["f"] = [f[1], f[0], f[3], f[2]]
    pc = np.zeros((g[2], g[1], g[0]), np.float32)
    qc = np.zeros_like(pc)
    for c, r in ranks.items():
        ls = r["sizes"]
        gx0 = [S + c[k] * part[k] for k in range(3)]
        sl_g = tuple(slice(gx0[k], gx0[k] + ls[k] - 2 * S) for k in (2, 1, 0))
        sl_l = tuple(slice(S, ls[k] - S) for k in (2, 1, 0))
        pc[sl_g] = r["f"][1][sl_l]
        qc[sl_g] = r["f"][3][sl_l]
    return pc, qc
