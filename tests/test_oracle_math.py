"""The oracle pinned by mathematics, independently of the golden fixtures.

* Weights: the float constants are the 8th-order central weights.
* Operators: the update isolates one derivative when the other coefficients
  are zero; on polynomials of degree <= 8 per axis the 8th-order operators
  are exact, so D_xx x^8 = 56 x^6, D_xy x^a y^b = a b x^(a-1) y^(b-1), ...
  on every axis pair -- a swapped stride, sign or weight fails one of them.
* Isotropic reduction: theta = 0, epsilon = delta = 0, vsv = 0 and p = q turn
  the Fletcher system into the textbook acoustic leapfrog with the 8th-order
  Laplacian, written here independently with numpy slicing.
"""
import numpy as np
import pytest

S = 4


def test_weights_are_the_8th_order_central_weights(O):
    w = O.weights().astype(np.float64)
    L = np.array([4 / 5, -1 / 5, 4 / 105, -1 / 280])
    K = np.array([-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    cross = [L[a] * L[b] for a, b in [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1),
                                      (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]]
    np.testing.assert_allclose(w[:4], L, rtol=1e-7)
    np.testing.assert_allclose(w[4:14], cross, rtol=1e-7)
    np.testing.assert_allclose(w[14:], K, rtol=1e-7)
    # consistency: second derivative of a constant is 0, of x^2 is 2
    k = np.arange(5)
    assert abs(K[0] + 2 * K[1:].sum()) < 1e-14
    assert abs((2 * K * k**2).sum() - 2) < 1e-13
    assert abs((2 * L * k[1:]).sum() - 1) < 1e-14  # D1 x = 1


def grid(n=17):
    c = np.arange(n, dtype=np.float64) - n // 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z


def isolate(O, which, field, d=(1.0, 1.0, 1.0), dt=1.0):
    """(p_new - 2p + p_old)/dt^2 with only coefficient `which` = 1."""
    shape = field.shape
    coef = [np.zeros(shape, np.float32) for _ in range(10)]
    coef[O.COEF_NAMES.index(which)][:] = 1.0
    coef[O.COEF_NAMES.index("v2sz")][:] = 1.0  # rhsp = H1(p) - H1(q), q = 0
    pc = field.astype(np.float64)
    pp = np.zeros_like(pc)
    zero = np.zeros_like(pc)
    n = shape[0]
    O.propagate_f64((S, S, S), (n - S,) * 3, pp, pc, zero.copy(), zero, coef,
                    *d, dt)
    inner = (slice(S, n - S),) * 3
    return ((pp - 2 * pc) / dt**2)[inner]


@pytest.mark.parametrize("which,f,want", [
    ("ch1dxx", lambda x, y, z: x**8, lambda x, y, z: 56 * x**6),
    ("ch1dyy", lambda x, y, z: y**7 + x, lambda x, y, z: 42 * y**5),
    ("ch1dzz", lambda x, y, z: z**6 * y, lambda x, y, z: 30 * z**4 * y),
    ("ch1dxy", lambda x, y, z: x**3 * y**2, lambda x, y, z: 6 * x**2 * y),
    ("ch1dyz", lambda x, y, z: y**2 * z**5, lambda x, y, z: 10 * y * z**4),
    ("ch1dxz", lambda x, y, z: x**4 * z**3 + y,
     lambda x, y, z: 12 * x**3 * z**2),
])
def test_operator_exact_on_polynomials(O, which, f, want):
    x, y, z = grid()
    got = isolate(O, which, f(x, y, z) / 1e4)
    n = x.shape[0]
    inner = (slice(S, n - S),) * 3
    ref = want(x, y, z)[inner] / 1e4
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_spacing_scales_each_axis(O):
    x, y, z = grid()
    d = (2.0, 3.0, 5.0)
    for which, f, scale in (("ch1dxx", x**2, 1 / 4), ("ch1dzz", z**2, 1 / 25),
                            ("ch1dxy", x * y, 1 / 6), ("ch1dyz", y * z, 1 / 15)):
        got = isolate(O, which, f, d)
        want = (2.0 if which in ("ch1dxx", "ch1dzz") else 1.0) * scale
        np.testing.assert_allclose(got, want, rtol=1e-9)


def acoustic_step(p_old, p, v2, dt, d):
    """8th-order acoustic leapfrog, written independently of the oracle."""
    K = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]
    n = p.shape[0]
    c = slice(S, n - S)
    lap = np.zeros_like(p[c, c, c])
    for axis in range(3):
        acc = K[0] * p[c, c, c]
        for a in range(1, 5):
            sl_p = [c, c, c]
            sl_m = [c, c, c]
            sl_p[axis] = slice(S + a, n - S + a)
            sl_m[axis] = slice(S - a, n - S - a)
            acc = acc + K[a] * (p[tuple(sl_p)] + p[tuple(sl_m)])
        lap += acc / d**2
    out = p_old.copy()
    out[c, c, c] = 2 * p[c, c, c] - p_old[c, c, c] + dt**2 * v2[c, c, c] * lap
    return out


def test_isotropic_reduction_is_acoustic_leapfrog(O, H):
    n, d = 24, 10.0
    rng = np.random.default_rng(4)
    shape = (n, n, n)
    vpz = rng.uniform(1500, 3000, shape).astype(np.float32)
    zero = np.zeros(shape, np.float32)
    coef = H.precompute(vpz, zero, zero, zero, zero, zero)  # theta=phi=0
    assert np.all(coef[2] == 1) and np.all(coef[0] == 0) and np.all(coef[8] == 0)
    p0 = rng.standard_normal(shape)
    p1 = rng.standard_normal(shape)
    dt = 1e-3
    f = O.run_f64([p0, p1, p0, p1], coef, d, d, d, dt, 5)
    want_old, want = p0.copy(), p1.copy()
    for _ in range(5):
        want_old, want = want, acoustic_step(want_old, want, coef[7].astype(
            np.float64), dt, d)
    np.testing.assert_allclose(f[1], want, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(f[1], f[3])  # p = q is preserved
    g = O.run_f32([p0.astype(np.float32), p1.astype(np.float32)] * 2, coef, d,
                  d, d, dt, 5)
    err = np.linalg.norm(g[1] - want) / np.linalg.norm(want)
    assert err < 1e-5


def test_fp32_oracle_tracks_fp64_on_a_wave(O, VM):
    w = VM.random_small((24, 26, 28), seed=1, steps=30)
    amp = O.source_table(w.dt, w.steps, w.fcut)
    z32 = [np.zeros(w.shape, np.float32)] * 4
    a = O.run_f32(z32, w.coef, w.dx, w.dy, w.dz, w.dt, w.steps, w.src_index, amp)
    b = O.run_f64(z32, w.coef, w.dx, w.dy, w.dz, w.dt, w.steps, w.src_index, amp)
    err = np.linalg.norm(a[1] - b[1]) / np.linalg.norm(b[1])
    assert 0 < err < 1e-5
