"""dcx_halo_face_noalloc against the reference ABI's extract and insert.

The entry point stages the packed box in a fixed device array instead of a
buffer it allocates; its results must equal dc_device_extract_halo_face's
and dc_device_insert_halo_face's bit for bit, on two calls in a row (the
second with a smaller box, which must not see the first call's leftovers),
and it refuses a box beyond the array with a message and exit status 1.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_CELLS = 1 << 22


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _struct(H, VM, shape, seed):
    w = VM.random_small(shape, seed=seed)
    rng = np.random.default_rng(seed + 1)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    return w, f0


def test_noalloc_extract_equals_reference_entry(H, VM):
    shape = (150, 140, 212)  # (sz, sy, sx)
    w, f0 = _struct(H, VM, shape, 3)
    d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt)
    sx, sy, sz = d.sizes
    # 100 whole planes (2,968,000 cells, 71% of the array), then an x slab
    # of the compute region (74,976 cells), then a 24-cell box.
    boxes = [((0, 0, 0), (sx, sy, 100)), ((4, 4, 4), (8, sy - 4, sz - 4)),
             ((5, 9, 13), (9, 12, 15))]
    for field in ("pc", "qp"):
        for start, end in boxes:
            print("noalloc extract", field, start, end, file=sys.stderr, flush=True)
            want = d.extract_box(start, end, field)
            got = d.extract_box_noalloc(start, end, field)
            assert got.size == want.size
            assert np.array_equal(bits(got), bits(want))
    d.close()


def test_noalloc_insert_equals_reference_entry(H, VM):
    shape = (60, 70, 90)
    w, f0 = _struct(H, VM, shape, 5)
    a = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt)
    b = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt)
    sx, sy, sz = a.sizes
    rng = np.random.default_rng(9)
    boxes = [((0, 0, 0), (sx, sy, 4)), ((7, 3, 11), (20, 9, 13))]
    for field in ("pp", "qc"):
        for start, end in boxes:
            n = int(np.prod([e - s for s, e in zip(start, end)]))
            buf = rng.standard_normal(n).astype(np.float32)
            print("noalloc insert", field, start, end, file=sys.stderr, flush=True)
            a.insert_box(buf, start, end, field)
            b.insert_box_noalloc(buf, start, end, field)
            fa, fb = a.fields(), b.fields()
            for x, y in zip(fa, fb):
                assert np.array_equal(bits(x), bits(y))
    a.close()
    b.close()


def test_noalloc_refuses_boxes_beyond_its_array():
    """A 162^3 box (4,251,528 cells) exceeds the 4,194,304-cell array: the
    call prints the limit and exits 1 before any device work."""
    code = textwrap.dedent(f"""
        import importlib, sys
        import numpy as np
        sys.path.insert(0, {REPO!r})
        H = importlib.import_module("distributed-cube-average_amd").dc_hip
        VM = importlib.import_module("distributed-cube-average_amd.velocity_model")
        w = VM.random_small((162, 162, 162), seed=1)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        d.synchronize()
        print("ready", flush=True)
        d.extract_box_noalloc((0, 0, 0), (162, 162, 162), "pc")
        print("not reached", flush=True)
    """)
    assert 162 ** 3 > MAX_CELLS
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=120)
    assert "ready" in r.stdout, r.stderr
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "not reached" not in r.stdout
    assert "exceeds the 4194304-cell staging array" in r.stderr, r.stderr
