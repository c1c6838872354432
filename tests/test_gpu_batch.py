"""Batched shots (dcx_batch_run): one persistent launch per step over the
tiles of several volumes of different (ragged) sizes.

Bar: every shot's four fields after the batched run are bit-identical to the
same shot stepped alone by dcx_run with the same kernel, for both step
parities, with and without a source, in FP32 and in the half-precision
time-difference form, eagerly and from a HIP graph.  Shot 0 is also held to
the FP32 oracle at the fast kernel's 1e-5 relative L2.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5
# (sz, sy, sx) of the shots: ragged, none a multiple of the 32x16 tile.
SHAPES = [(40, 36, 44), (48, 40, 56), (33, 45, 37), (24, 24, 24), (70, 20, 30)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def shots(VM):
    out = []
    for k, shape in enumerate(SHAPES):
        w = VM.random_small(shape, seed=10 + k, steps=32, peak_hz=25.0 + 5 * k)
        src = -1 if k == 3 else w.src_index
        out.append((w, src, w.source_table()))
    return out


def _make(H, w, form):
    d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
    if form != (0, 0):
        assert d.set_storage(*form)
    return d


def _alone(H, w, src, amp, steps, form):
    d = _make(H, w, form)
    d.run(steps, src, amp[:steps])
    f = d.fields()
    d.close()
    return f


@pytest.mark.parametrize("form", [(0, 0), (0, 2), (0, 3), (0, 4), (1, 1)])
@pytest.mark.parametrize("steps,graph,tiling", [
    (24, False, (128, 1, 1)), (23, True, (16, 2, 2)), (7, False, (5, 1, 3)),
    (25, True, "auto"), (12, False, (64, 2, 0))])
def test_batch_equals_shots_alone(H, shots, form, steps, graph, tiling):
    say("batch form", form, "steps", steps, "graph", graph, "tiling", tiling)
    ds = [_make(H, w, form) for w, _, _ in shots]
    b = H.Batch(ds)
    if tiling == "auto":
        b.set_autotune(True)
    else:
        assert b.set_tiling(*tiling)
    idx = [src for _, src, _ in shots]
    amp = np.stack([a[:steps] for _, _, a in shots])
    ms = b.run(steps, idx, amp, graph=graph)
    assert ms > 0
    say("  tiling used", b.tiling(), f"{b.last_step_ms():.4f} ms/step")
    b.close()
    for k, ((w, src, a), d) in enumerate(zip(shots, ds)):
        got = d.fields()
        want = _alone(H, w, src, a, steps, form)
        for name, g, e in zip(("pp", "pc", "qp", "qc"), got, want):
            assert np.array_equal(bits(g), bits(e)), (k, name)
        # Shots stay usable after the batch: one more step alone.
        d.run(1, src, a[steps:steps + 1])
        d.close()


def test_batch_shot_against_oracle(H, O, shots):
    w, src, amp = shots[0]
    steps = 24
    d = _make(H, w, (0, 0))
    b = H.Batch([d])
    b.run(steps, [src], amp[None, :steps])
    b.close()
    pc, qc = d.get_results()
    d.close()
    z = [np.zeros(w.shape, np.float32) for _ in range(4)]
    ref = O.run_f32(z, w.coef, w.dx, w.dy, w.dz, w.dt, steps, src, amp[:steps])
    e = rel_l2(pc, ref[1]), rel_l2(qc, ref[3])
    say(f"batched shot vs FP32 oracle: p={e[0]:.3e} q={e[1]:.3e}")
    assert max(e) < FAST_TOL


def test_batch_rejects_mixed_forms(H, shots):
    """Shots of different storage forms cannot share a launch (host-side
    check; the library exits, so run it in a child process)."""
    import subprocess
    code = (
        "import importlib, numpy as np\n"
        "H = importlib.import_module('distributed-cube-average_amd').dc_hip\n"
        "c = [np.ones((24, 24, 24), np.float32) for _ in range(10)]\n"
        "a = H.DeviceData(c); b = H.DeviceData(c); b.set_storage(1, 1)\n"
        "H.Batch([a, b])\n")
    import os
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=repo,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "storage form" in r.stderr


def test_full_size_survey_shots(H, VM):
    """Six shots of the benchmark survey at their real apertures (up to
    320x320x256 compute cells): batched with the autotuner and a graph, each
    bit-identical to the shot run alone."""
    S = VM.layered_survey(seed=0, n_shots=6, steps=40)
    steps = 24
    ws = [S.shot(k) for k in range(len(S))]
    say("survey shots", [w.sizes for w in ws])
    ds = []
    for w in ws:
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        d.release_host()
        ds.append(d)
    amp = np.stack([w.source_table()[:steps] for w in ws])
    b = H.Batch(ds)
    b.set_autotune(True)
    b.run(steps, [w.src_index for w in ws], amp, graph=True)
    say("  tiling used", b.tiling())
    b.close()
    for k, (w, d) in enumerate(zip(ws, ds)):
        got = d.fields()
        d.close()
        a = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        a.run(steps, w.src_index, amp[k], graph=True)
        want = a.fields()
        a.close()
        for name, g, e in zip(("pp", "pc", "qp", "qc"), got, want):
            assert np.array_equal(bits(g), bits(e)), (k, name)


@pytest.mark.parametrize("form", [(0, 0), (0, 4), (1, 1)])
def test_whole_survey_batched_equals_shots_alone(H, VM, form):
    """The batching configuration at its workload: all 64 shots of the
    benchmark survey (compute extents U[128,320] x U[128,320] x U[128,256],
    589 M cells per step), each with its own Ricker, in one persistent
    launch per step with the autotuned tiling and a graph, in FP32, the
    axis coefficient form and FP16 storage (fields and coefficients).  A few
    steps (the autotuner's trials included), and every shot's four fields
    bit-identical to the shot stepped alone."""
    S = VM.layered_survey(seed=0, n_shots=64, steps=40)
    steps = 16
    ws, ds = [], []
    for k in range(len(S)):
        w = S.shot(k)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        assert d.set_storage(*form)
        d.release_host()
        w.coef = None  # the shot is rebuilt below for its run alone
        ws.append(w)
        ds.append(d)
    say("survey", len(ds), "shots on the device, form", form)
    idx = [w.src_index for w in ws]
    amp = np.stack([H.ricker_table(w.dt, steps, w.fcut) for w in ws])
    b = H.Batch(ds)
    b.set_autotune(True)
    b.run(steps, idx, amp, graph=True)
    say("  tiling used", b.tiling())
    b.close()
    for k, d in enumerate(ds):
        got = d.fields()
        d.close()
        w = S.shot(k)
        a = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        assert a.set_storage(*form)
        a.run(steps, idx[k], amp[k], graph=True)
        want = a.fields()
        a.close()
        for name, g, e in zip(("pp", "pc", "qp", "qc"), got, want):
            assert np.array_equal(bits(g), bits(e)), (k, name)


def test_freeing_a_shot_inside_a_batch_fails_loudly(H):
    """A shot freed while its batch is open would leave the batch with a
    dangling struct: the library refuses (message and exit), and a network
    still attached to a struct cannot be freed either."""
    import os
    import subprocess
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prelude = ("import importlib, numpy as np\n"
               "H = importlib.import_module('distributed-cube-average_amd').dc_hip\n"
               "c = [np.ones((24, 24, 24), np.float32) for _ in range(10)]\n")
    cases = [
        ("a = H.DeviceData(c); b = H.DeviceData(c); B = H.Batch([a, b])\n"
         "a.close()\n", "free the batch"),
        ("a = H.DeviceData(c); m = H.Mlp(np.zeros(8450, np.float32))\n"
         "H.set_learned_layer(a, m, 3)\nm.close()\n", "still attached"),
    ]
    for code, msg in cases:
        r = subprocess.run([sys.executable, "-c", prelude + code], cwd=repo,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode != 0 and msg in r.stderr, (r.returncode, r.stderr[-300:])


@pytest.mark.parametrize("form", [(0, 0), (0, 4)])
def test_batch_shots_with_their_own_spacings(H, VM, form):
    """Shots whose grid spacings differ from each other and between axes
    (10/20/35, 20/20/20 and 35/10/20 m), batched, against the same shots
    stepped alone, bit for bit.  Each shot's launch arguments carry its own
    1/(d_a d_b) factors and, in the axis form, its own spacing-scaled axis
    decode."""
    spacings = [(10.0, 20.0, 35.0), (20.0, 20.0, 20.0), (35.0, 10.0, 20.0)]
    steps = 20
    ws = []
    for k, (dx, dy, dz) in enumerate(spacings):
        w = VM.random_small(SHAPES[k], seed=40 + k)
        dt = float(np.float32(0.4 * VM.cfl_dt(w.coef[6], dx, dy, dz)))
        amp = H.ricker_table(dt, steps, w.fcut)
        ws.append((w, (dx, dy, dz, dt), amp))

    def make(w, s):
        d = H.DeviceData(w.coef, None, *s)
        if form != (0, 0):
            assert d.set_storage(*form)
        return d

    ds = [make(w, s) for w, s, _ in ws]
    b = H.Batch(ds)
    assert b.set_tiling(64, 2, 0)
    b.run(steps, [w.src_index for w, _, _ in ws],
          np.stack([a for _, _, a in ws]), graph=True)
    b.close()
    for k, ((w, s, amp), d) in enumerate(zip(ws, ds)):
        got = d.fields()
        d.close()
        alone = make(w, s)
        alone.run(steps, w.src_index, amp)
        want = alone.fields()
        alone.close()
        assert np.abs(want[1]).max() > 0
        for name, g, e in zip(("pp", "pc", "qp", "qc"), got, want):
            assert np.array_equal(bits(g), bits(e)), (k, name)
