"""The library's host-side model setup (precompute, random boundary, Ricker,
default model) against the reference's outputs and the oracle.  CPU only."""
import numpy as np
import pytest

from conftest import golden


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_precompute_matches_reference(H):
    g = golden("golden_precomp.npz")
    out = H.precompute(g["vpz"], g["vsv"], g["epsilon"], g["delta"], g["phi"],
                       g["theta"])
    for name, a in zip(H.COEF_NAMES, out):
        assert np.array_equal(bits(a), bits(g[name])), name


def test_ricker_matches_reference(H):
    g = golden("golden_source.npz")
    for dt, row in zip(g["dt"], g["amp"]):
        ours = H.ricker_table(float(dt), row.size)
        assert np.array_equal(bits(ours), bits(row))


def test_ricker_other_cutoff_matches_oracle(H, O):
    for fcut in (45.0, 60.0, 75.0):
        a = H.ricker_table(6.4e-4, 300, fcut)
        b = np.array([O.source_f32(6.4e-4, i, fcut) for i in range(300)],
                     np.float32)
        assert np.array_equal(bits(a), bits(b))


def test_random_boundary_matches_reference(H):
    g = golden("golden_boundary.npz")
    gl = tuple(int(v) for v in g["global_sizes"])
    prob = tuple(int(v) for v in g["problem"])
    for key in ("000", "111", "010"):
        ref = g[f"vpz_{key}"]
        loc = (ref.shape[2], ref.shape[1], ref.shape[0])
        vpz = np.full(ref.shape, 3000.0, np.float32)
        vsv = np.full(ref.shape, 1296.1481, np.float32)
        H.random_boundary(loc, gl, tuple(g[f"start_{key}"]), prob, 4,
                          int(g["absorb"]), vpz, vsv, 0)
        assert np.array_equal(bits(vpz), bits(ref))
        assert np.array_equal(bits(vsv), bits(g[f"vsv_{key}"]))


def test_default_model_matches_oracle(H, O):
    a = H.default_model(100)
    b = O.default_model(100)
    for x, y in zip(a, b):
        assert np.array_equal(bits(x), bits(y))


def test_smooth_random_workload_is_sane(VM):
    w = VM.smooth_random(seed=3, problem=16, absorption=4, steps=10)
    assert w.shape == (32, 32, 32)
    v = np.sqrt(w.coef[7])  # vpz
    inner = v[8:-8, 8:-8, 8:-8]
    assert inner.min() >= 1500 - 1e-3 and inner.max() <= 4500 + 1e-3
    assert np.all(v[:4] == 0)  # the reference's zero frame
    assert 0 < w.dt < VM.cfl_dt(w.coef[6], w.dx, w.dy, w.dz)
    assert len(w.receivers) == 1024


@pytest.mark.parametrize("case", range(6))
def test_random_boundary_cell_by_cell_matches_the_walk(H, O, case):
    """The library evaluates each cell's draw in closed form (its index among
    the shell cells, the generator jumped ahead); the oracle walks the whole
    global grid as the reference does.  Same bits on uneven splits, several
    widths and seeds, and the same final generator state."""
    rng = np.random.default_rng(100 + case)
    bord = int(rng.integers(2, 6))
    absorb = int(rng.integers(1, 9))
    prob = tuple(int(v) for v in rng.integers(3, 14, 3))
    gl = tuple(p + 2 * (bord + absorb) for p in prob)
    start = tuple(int(rng.integers(0, g - 3)) for g in gl)
    loc = tuple(int(rng.integers(1, g - s + 1)) for g, s in zip(gl, start))
    seed = int(rng.integers(0, 2**32 - 1))
    shape = (loc[2], loc[1], loc[0])
    init = rng.uniform(1000, 4000, shape).astype(np.float32)
    a, b = init.copy(), (init * 0.5).astype(np.float32)
    c, d = a.copy(), b.copy()
    s1 = H.random_boundary(loc, gl, start, prob, bord, absorb, a, b, seed)
    s2 = O.random_boundary(loc, gl, start, prob, bord, absorb, c, d, seed)
    assert s1 == s2
    assert np.array_equal(bits(a), bits(c)) and np.array_equal(bits(b), bits(d))
