"""The one-way wall's CPU restatement (oracle/wall_oracle.py) against what
the mathematics fixes: the first-order condition reproduces an outgoing
linear wave exactly on every face, leaves a field at rest in space and time
unchanged, and its fill order reaches every frame cell."""
import numpy as np
import pytest

import wall_oracle as WO

S = WO.STENCIL


def const_model(shape, cnn=(0.3, 0.5, 0.2), v2pz=4.0e6, rx=1.4):
    """Ten constant coefficient arrays; the three angle terms and the two
    velocities differ, so that a wrong axis or a wrong mixing moves c."""
    coef = [np.zeros(shape, np.float32) for _ in range(10)]
    for a in range(3):
        coef[a][:] = cnn[a]
    coef[WO.V2PZ][:] = v2pz
    coef[WO.V2PX][:] = v2pz * rx
    return coef


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [0, 1])
def test_outgoing_linear_wave_is_exact(axis, side):
    nx, ny, nz = 19, 17, 15
    h = (20.0, 15.0, 25.0)
    dt = 1.5e-3
    coef = const_model((nz, ny, nx))
    cn = (0.3, 0.5, 0.2)[axis]
    c = float(np.sqrt(4.0e6 * 1.4 * (1 - cn) + 4.0e6 * cn))
    sgn = 1.0 if side else -1.0
    n = (nx, ny, nz)
    X = np.arange(n[axis], dtype=np.float64) * h[axis]
    shape = [1, 1, 1]
    shape[2 - axis] = n[axis]
    X = X.reshape(shape)
    A, B, t = 3.0, 1.0e-3, 0.37

    def wave(tt):
        return np.broadcast_to(A + B * (sgn * X - c * tt), (nz, ny, nx))

    pc = np.array(wave(t), np.float32)
    qc = np.array(-2.0 * wave(t), np.float32)
    pn = np.array(wave(t + dt), np.float32)
    qn = np.array(-2.0 * wave(t + dt), np.float32)
    # The frame of the new level holds garbage the fill must overwrite.
    frame = np.ones((nz, ny, nx), bool)
    frame[S:-S, S:-S, S:-S] = False
    pn[frame] = 1e3
    qn[frame] = -1e3
    WO.wall_fill(pn, qn, pc, qc, coef, dt, *h)
    want = wave(t + dt)
    # The filled frame of this axis's face, the other axes in compute range.
    sl = [slice(S, -S)] * 3
    sl[2 - axis] = slice(n[axis] - S, n[axis]) if side else slice(0, S)
    sl = tuple(sl)
    scale = B * h[axis]
    err = np.max(np.abs(pn[sl] - want[sl]))
    assert err < 2e-3 * scale
    assert np.max(np.abs(qn[sl] + 2.0 * want[sl])) < 4e-3 * scale
    # A mistake in c (another axis's angle term) misses by far more than the
    # rounding of the exact fill.
    others = [cc for k, cc in enumerate((0.3, 0.5, 0.2)) if k != axis]
    wrong = WO.wall_fill(np.array(pn), np.array(qn), pc, qc,
                         const_model((nz, ny, nx), cnn=(others[0],) * 3), dt,
                         *h)[0]
    assert np.max(np.abs(wrong[sl] - want[sl])) > 20 * max(err, 1e-7)


def test_rest_is_kept_and_every_frame_cell_is_filled():
    nx, ny, nz = 13, 14, 12
    coef = const_model((nz, ny, nx))
    pc = np.full((nz, ny, nx), 0.25, np.float32)
    qc = np.full((nz, ny, nx), -0.5, np.float32)
    pn = pc.copy()
    qn = qc.copy()
    frame = np.ones((nz, ny, nx), bool)
    frame[S:-S, S:-S, S:-S] = False
    pn[frame] = np.nan
    qn[frame] = np.nan
    WO.wall_fill(pn, qn, pc, qc, coef, 1.5e-3, 20.0, 20.0, 20.0)
    assert np.all(pn == np.float32(0.25))
    assert np.all(qn == np.float32(-0.5))
