"""The library's host-side model setup (precompute, random boundary, Ricker,
default model) against the reference's outputs and the oracle.  CPU only."""
import numpy as np

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
