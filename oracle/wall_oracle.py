"""CPU restatement of the one-way wall (test infrastructure only).

Used by tests/ as the checker for the library's frame fill (dcx_set_wall);
nothing in the library imports it.  The wall has no counterpart in the
reference, whose outer frame stays at zero (src/boundary.c:7-104), so this
file is pinned by mathematics (tests/test_wall_oracle.py):
  * the first-order condition is exact for an outgoing wave u = f(x - c t)
    with f linear: on every face, with c from the mixing of v2px and v2pz
    by the face's angle term, the filled frame equals the wave's closed form
    at the new time level;
  * a wave standing still (c = 0 is not reachable; instead fields constant
    in space and time) is left unchanged;
  * the fill order (x faces, then y over all x, then z over all x and y)
    reaches every frame cell exactly once per axis whose frame holds it.

Arrays are (z, y, x) float32, the library's host layout; the arithmetic is
float32 in the kernel's order.
"""
from __future__ import annotations

import numpy as np

STENCIL = 4
CXX, CYY, CZZ, CXY, CYZ, CXZ, V2PX, V2PZ, V2SZ, V2PN = range(10)


def _moved(a, axis):
    """View of a (z, y, x) array with the spatial axis `axis` (0 = x) first
    and the other two in array order."""
    return np.moveaxis(a, 2 - axis, 0)


def wall_fill(pn, qn, pc, qc, coef, dt, dx, dy, dz, order=1, po=None,
              qo=None):
    """Fill the 4-cell frame of the new level pn, qn (in place) from the
    outgoing wave, outward one layer at a time, r = (c dt - h) / (c dt + h),
    c^2 = v2px (1 - c_nn) + v2pz c_nn at the last compute cell along the
    normal.  With the first-order prediction
      pred(a, b, c) = a + r (b - c)
    of u_F^{n+1} from a = u_I^n, b = u_I^{n+1}, c = u_F^n:
    order 1: u_F^{n+1} = pred(u_I^n, u_I^{n+1}, u_F^n);
    order 2 (po, qo: level n-1 as it was before the step), B B u = 0 written
    as the first-order condition on the first-order residual
    v(G, m) = u_G^m - pred(u_{G-1}^{m-1}, u_{G-1}^m, u_G^{m-1}):
      u_F^{n+1} = pred(u_I^n, u_I^{n+1}, u_F^n)
                  + pred(v(I, n), v(I, n+1), v(F, n)),
    G-1 meaning one cell towards the compute region.  Each frame layer is
    one float32 array operation over the face (elementwise, so its results
    are those of a cell-by-cell loop)."""
    f32 = np.float32
    S = STENCIL
    n = pn.shape[::-1]                    # (nx, ny, nz)
    h = (f32(dx), f32(dy), f32(dz))
    cnn = (coef[CXX], coef[CYY], coef[CZZ])
    dt = f32(dt)
    for axis in range(3):
        # The two other axes in array order (z before y before x).
        others = [b for b in (2, 1, 0) if b != axis]
        rngs = [np.arange(0, n[b]) if b < axis else np.arange(S, n[b] - S)
                for b in others]
        clamp = [np.clip(r, S, n[b] - S - 1) for r, b in zip(rngs, others)]
        face = np.ix_(*rngs)
        cface = np.ix_(*clamp)
        for side in (0, 1):
            last = n[axis] - S - 1 if side else S
            dirn = 1 if side else -1
            cn = _moved(cnn[axis], axis)[last][cface].astype(f32)
            vx = _moved(coef[V2PX], axis)[last][cface].astype(f32)
            vz = _moved(coef[V2PZ], axis)[last][cface].astype(f32)
            c2 = (vx * (f32(1.0) - cn)).astype(f32) + (vz * cn).astype(f32)
            cdt = (np.sqrt(c2.astype(f32)) * dt).astype(f32)
            r = ((cdt - h[axis]).astype(f32) / (cdt + h[axis]).astype(f32)).astype(f32)

            def pred(a_, b_, c_):
                return (a_ + (r * (b_ - c_).astype(f32)).astype(f32)).astype(f32)
            for new, cur, old in ((pn, pc, po), (qn, qc, qo)):
                N, C = _moved(new, axis), _moved(cur, axis)
                for k in range(1, S + 1):
                    F = last + dirn * k
                    I = F - dirn
                    p1 = pred(C[I][face], N[I][face], C[F][face])
                    if order == 1:
                        N[F][face] = p1
                        continue
                    O = _moved(old, axis)
                    I2 = I - dirn
                    v_in1 = (N[I][face] - pred(C[I2][face], N[I2][face], C[I][face])).astype(f32)
                    v_in = (C[I][face] - pred(O[I2][face], C[I2][face], O[I][face])).astype(f32)
                    v_fn = (C[F][face] - pred(O[I][face], C[I][face], O[F][face])).astype(f32)
                    N[F][face] = (p1 + pred(v_in, v_in1, v_fn)).astype(f32)
    return pn, qn
