"""Pins the learned-layer oracle (oracle/layer_oracle.py) by mathematics: the
layer has no reference counterpart, so its checker is held to finite
differences, closed forms and counting arguments here."""
import numpy as np
import pytest

import layer_oracle as LO


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(LO.NPARAM) * 0.3


def test_parameter_count():
    assert LO.NPARAM == 64 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2 == 8450


def test_gradient_is_the_derivative_of_the_loss():
    """Backpropagation (float64, no rounding) against central differences,
    every parameter block."""
    rng = np.random.default_rng(0)
    p = _params(1)
    x = rng.uniform(-1, 1, (12, 64))
    t = rng.uniform(-1, 1, (12, 2))
    g = LO.gradient(p, x, t, fp16=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(4096, 40, replace=False),
                          4096 + rng.choice(64, 10, replace=False),
                          4160 + rng.choice(4096, 40, replace=False),
                          8256 + rng.choice(64, 10, replace=False),
                          8320 + np.arange(130)])
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (LO.loss(a, x, t, fp16=False) - LO.loss(b, x, t, fp16=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_forward_of_a_linear_network():
    """Hidden layers that pass a positive input through unchanged make the
    network the linear map W3 x[:64] + b3 (checks the parameter layout)."""
    W1 = np.eye(64)
    W2 = np.eye(64)
    rng = np.random.default_rng(2)
    W3 = rng.standard_normal((2, 64))
    b3 = np.array([0.5, -0.25])
    p = np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64),
                        W3.ravel(), b3])
    x = rng.uniform(0, 1, (5, 64))
    y, _ = LO.forward(p, x, fp16=False)
    assert np.allclose(y, x @ W3.T + b3, rtol=1e-12, atol=1e-12)


def test_fp16_emulation_rounds_only_operands():
    """With fp16 the result equals the float64 network evaluated on
    half-rounded operands (inputs, weights, hidden activations)."""
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.uniform(-1, 1, (7, 64))
    y16, (xh, z1, h1, z2, h2) = LO.forward(p, x, fp16=True)
    assert np.array_equal(xh, x.astype(np.float16).astype(np.float64))
    assert np.array_equal(h1, np.maximum(z1, 0).astype(np.float16).astype(np.float64))
    y64, _ = LO.forward(p, x, fp16=False)
    assert np.max(np.abs(y16 - y64)) < 0.05 * np.max(np.abs(y64))


@pytest.mark.parametrize("n,w", [((10, 12, 14), 3), ((20, 20, 20), 6), ((17, 9, 11), 4)])
def test_layer_cells_count(n, w):
    cells = LO.layer_cells(n, w)
    inner = np.prod([max(0, a - 2 * w) for a in n])
    assert len(cells) == np.prod(n) - inner
    assert len(set(cells)) == len(cells)


def _linear_fields(shape, a, b):
    """p^n - p^{n-1} = a . (x, y, z) + 1 and q^n - q^{n-1} = b . (x, y, z) - 2."""
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    pc = (a[0] * x + a[1] * y + a[2] * z + 1.0).astype(np.float32)
    qc = (b[0] * x + b[1] * y + b[2] * z - 2.0).astype(np.float32)
    zero = np.zeros(shape, np.float32)
    return pc, zero, qc, zero


@pytest.mark.parametrize("cell,axis,sign", [
    ((0, 7, 9), 0, -1), ((11, 7, 9), 0, 1), ((5, 0, 9), 1, -1),
    ((5, 13, 9), 1, 1), ((5, 7, 0), 2, -1), ((5, 7, 15), 2, 1)])
def test_features_orientation(cell, axis, sign):
    """On linear time differences the patch is a linear function of (i, j,
    k) whose slope along i is sign * (the slope along the face's axis) and
    along j, k the slopes of the two other axes in increasing order."""
    shape = (16 + 8, 14 + 8, 12 + 8)   # compute extents (12, 14, 16)
    a, b = np.array([0.3, -0.2, 0.1]), np.array([-0.05, 0.07, 0.11])
    pc, pp, qc, qp = _linear_fields(shape, a, b)
    v2 = np.full(shape, 4.0e6, np.float32)
    x, s = LO.features(pc, pp, qc, qp, cell, v2, v2 * 2, np.float32(1e-4))
    t1, t2 = [k for k in range(3) if k != axis]
    c = np.array(cell, np.float64) + 4
    for q, (g, off) in enumerate(((a, 1.0), (b, -2.0))):
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    pos = c.copy()
                    pos[axis] += i * sign
                    pos[t1] += j
                    pos[t2] += k
                    want = (g @ pos + off)
                    got = x[27 * q + (i + 1) * 9 + (j + 1) * 3 + (k + 1)] * s
                    assert abs(got - want) <= 1e-5 * max(1.0, abs(want))
    assert np.max(np.abs(x[:54])) == 1.0
    depth = min(min(cell[d], (12, 14, 16)[d] - 1 - cell[d]) for d in range(3))
    assert depth == 0
    assert x[54 + depth] == 1.0 and np.sum(x[54:62]) == 1.0
    assert x[62] == np.float32(2000.0) * np.float32(1e-4)
    assert x[63] == np.sqrt(np.float32(8.0e6)) * np.float32(1e-4)


def test_features_depth_and_ties():
    """Depth is the distance to the nearest face; ties go to the lowest axis,
    low side first; zero fields give zero inputs and s = 0."""
    shape = (12 + 8, 12 + 8, 12 + 8)
    z = np.zeros(shape, np.float32)
    v2 = np.ones(shape, np.float32)
    for cell, depth in (((2, 5, 5), 2), ((3, 3, 3), 3), ((9, 5, 5), 2), ((5, 5, 5), 5)):
        x, s = LO.features(z, z, z, z, cell, v2, v2, np.float32(1.0))
        assert s == 0 and np.all(x[:54] == 0)
        assert x[54 + depth] == 1.0
    assert LO.cell_frame((3, 3, 3), (12, 12, 12))[:2] == (0, -1)
    assert LO.cell_frame((8, 8, 8), (12, 12, 12))[:2] == (0, 1)
    assert LO.cell_frame((8, 3, 8), (12, 12, 12))[:2] == (0, 1)   # x-high ties y-low
    assert LO.cell_frame((8, 2, 8), (12, 12, 12))[:2] == (1, -1)


def test_courant_factor():
    assert LO.courant_factor(np.float32(1e-3), np.float32(20.0)) == \
        np.float32(1e-3) * np.sqrt(np.float32(1.0) / np.float32(400.0), dtype=np.float32)


def test_damping_gradient_is_the_derivative_of_the_loss():
    """The damping form's chain rule (sigmoid and the centre input) against
    central differences, float64."""
    rng = np.random.default_rng(5)
    p = _params(6)
    x = rng.uniform(-1, 1, (9, 64))
    t = rng.uniform(-0.3, 0.3, (9, 2))
    kw = dict(fp16=False, mode=LO.DAMPING, sigma_max=0.5)
    g = LO.gradient(p, x, t, **kw)
    h = 1e-6
    for i in np.concatenate([rng.choice(4096, 20, replace=False),
                             4160 + rng.choice(4096, 20, replace=False),
                             8320 + np.arange(130)]):
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (LO.loss(a, x, t, **kw) - LO.loss(b, x, t, **kw)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_damping_prediction_is_bounded():
    """|u| <= sigma_max |x_c|: the damping form never adds energy."""
    rng = np.random.default_rng(7)
    p = _params(8) * 10
    x = rng.uniform(-1, 1, (50, 64))
    u = LO.predict(p, x, fp16=False, mode=LO.DAMPING, sigma_max=0.5)
    assert np.all(np.abs(u) <= 0.5 * np.abs(x[:, LO.CENTRE]) + 1e-15)
    assert np.all(u * x[:, LO.CENTRE] <= 0)
