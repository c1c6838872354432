"""Device-side model setup (dcx_device_data_init_model) against the reference.

The reference's worker builds its rank's model on the host: the default TTI
model, the random-velocity shell by a walk over the global grid
(src/boundary.c:25-103) and the ten coefficient arrays (src/precomp.c:85-105).
The device path computes every local cell's coefficients in one kernel.  Its
ten arrays must equal, bit for bit:
  * those of the reference's own boundary output (golden_boundary.npz, every
    rank of a 2x2x2 split) run through the precompute, and
  * those of the oracle's global walk and precompute on uneven splits,
and its generator state must end where the walk leaves it.  The whole program
(tests/test_cli.py, test_gpu_parity.py::test_program_*) now builds every
rank this way and stays bit-identical to the reference's mpirun output.
"""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

STENCIL = 4


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def default_inputs(O, shape):
    n = int(np.prod(shape))
    vpz, vsv, eps, dlt, phi, th = [a.reshape(shape) for a in O.default_model(n)]
    return vpz, vsv, eps, dlt, phi, th


def check_coef(d, want):
    for k in range(10):
        got = d.download_coef(k)
        assert np.array_equal(bits(got), bits(want[k])), f"coefficient {k}"


def test_device_model_matches_reference_boundary_fixture(H, O):
    g = golden("golden_boundary.npz")
    gl = tuple(int(v) for v in g["global_sizes"])
    prob = tuple(int(v) for v in g["problem"])
    ab = int(g["absorb"])
    for key in ("000", "100", "010", "110", "001", "101", "011", "111"):
        vpz = g[f"vpz_{key}"]
        shape = vpz.shape
        loc = (shape[2], shape[1], shape[0])
        start = tuple(int(v) for v in g[f"start_{key}"])
        print("device model rank", key, loc, start, file=sys.stderr, flush=True)
        _, _, eps, dlt, phi, th = default_inputs(O, shape)
        want = O.precomp(vpz, g[f"vsv_{key}"], eps, dlt, phi, th)
        d, seed = H.DeviceData.from_default_model(loc, gl, start, prob, ab, 0)
        check_coef(d, want)
        assert seed == int(g[f"seed_{key}"])
        d.close()


@pytest.mark.parametrize("case", range(6))
def test_device_model_matches_the_walk_on_uneven_splits(H, O, case):
    rng = np.random.default_rng(300 + case)
    absorb = int(rng.integers(1, 9))
    prob = tuple(int(v) for v in rng.integers(3, 20, 3))
    gl = tuple(p + 2 * (STENCIL + absorb) for p in prob)
    start = tuple(int(rng.integers(0, g - 9)) for g in gl)
    loc = tuple(int(rng.integers(9, g - s + 1)) for g, s in zip(gl, start))
    seed = int(rng.integers(0, 2**32 - 1))
    shape = (loc[2], loc[1], loc[0])
    print("device model case", case, loc, start, gl, absorb, file=sys.stderr,
          flush=True)
    vpz, vsv, eps, dlt, phi, th = default_inputs(O, shape)
    s_walk = O.random_boundary(loc, gl, start, prob, STENCIL, absorb, vpz, vsv,
                               seed)
    want = O.precomp(vpz, vsv, eps, dlt, phi, th)
    d, s_dev = H.DeviceData.from_default_model(loc, gl, start, prob, absorb, seed)
    check_coef(d, want)
    assert s_dev == s_walk
    d.close()


def test_device_model_runs_like_the_uploaded_model(H, O):
    """One rank's whole grid (1x1x1 split): 25 exact-kernel steps from the
    device-built model equal, bit for bit, 25 steps from the same model
    built on the host and uploaded through dc_device_data_init."""
    prob, ab = (14, 12, 10), 4
    gl = tuple(p + 2 * (STENCIL + ab) for p in prob)
    shape = gl[::-1]
    vpz, vsv, eps, dlt, phi, th = default_inputs(O, shape)
    O.random_boundary(gl, gl, (0, 0, 0), prob, STENCIL, ab, vpz, vsv, 0)
    coef = O.precomp(vpz, vsv, eps, dlt, phi, th)
    dt = 0.002
    src = gl[0] // 2 + gl[0] * (gl[1] // 2 + gl[1] * (gl[2] // 2))
    amp = H.ricker_table(dt, 25)
    a = H.DeviceData(coef, None, 20.0, 20.0, 20.0, dt, kernel=H.KERNEL_EXACT)
    b, _ = H.DeviceData.from_default_model(gl, gl, (0, 0, 0), prob, ab, 0,
                                           20.0, 20.0, 20.0, dt,
                                           kernel=H.KERNEL_EXACT)
    for d in (a, b):
        d.run(25, src, amp)
    for x, y in zip(a.get_results(), b.get_results()):
        assert np.array_equal(bits(x), bits(y))
    a.close()
    b.close()
