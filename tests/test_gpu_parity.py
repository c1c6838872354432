"""HIP paths through the C ABI against the oracle and the golden fixtures.

Bars: the exact kernel is bit-identical to the reference (golden fixtures and
oracle); the fast kernel is within 1e-5 relative L2 of the FP32 oracle on
wavefields (its error against the FP64 oracle is printed); halo packing and
all integer/index work are bit-exact.
"""
import sys

import numpy as np
import pytest

from conftest import golden

import program_oracle as PO

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(b), 1e-300))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _coef(H, g):
    return H.precompute(g["vpz"], g["vsv"], g["epsilon"], g["delta"], g["phi"],
                        g["theta"])


@pytest.mark.parametrize("kernel", [1, 0])
def test_one_propagate_call(H, kernel):
    g = golden("golden_cell.npz")
    say("golden_cell kernel", kernel)
    f0 = [g[k] for k in ("pp0", "pc0", "qp0", "qc0")]
    d = H.DeviceData(_coef(H, g), f0, float(g["dx"]), float(g["dy"]),
                     float(g["dz"]), float(g["dt"]), kernel=kernel)
    sz, sy, sx = f0[0].shape
    d.propagate((4, 4, 4), (sx - 4, sy - 4, sz - 4))
    pp, pc, qp, qc = d.fields()
    d.close()
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pp), bits(g["pp1"]))
        assert np.array_equal(bits(qp), bits(g["qp1"]))
    else:
        # one step from random fields: compare the update increments
        for new, old, ref in ((pp, g["pp0"], g["pp1"]), (qp, g["qp0"], g["qp1"])):
            assert rel_l2(new - old, ref - old) < FAST_TOL
    assert np.array_equal(bits(pc), bits(g["pc0"]))
    assert np.array_equal(bits(qc), bits(g["qc0"]))


def _reference_loop(H, d, steps, src_index, dt):
    """src/worker.c:466-511 for one rank: source, 7 region calls, swap."""
    for it in range(steps):
        d.add_source(src_index, H.ricker(dt, it))
        for a, b in PO.slabs(d.sizes):
            d.propagate(a, b)
        d.swap()


@pytest.mark.parametrize("kernel", [1, 0])
@pytest.mark.parametrize("mode", ["abi", "fused", "graph"])
def test_forty_step_loop(H, O, kernel, mode):
    g = golden("golden_run.npz")
    say("golden_run", kernel, mode)
    coef = _coef(H, g)
    dt, steps, src = float(g["dt"]), int(g["steps"]), int(g["src_index"])
    dx = float(g["dx"])
    d = H.DeviceData(coef, None, dx, dx, dx, dt, kernel=kernel)
    if mode == "abi":
        _reference_loop(H, d, steps, src, dt)
    else:
        d.run(steps, src, H.ricker_table(dt, steps), graph=(mode == "graph"))
    pc, qc = d.get_results()
    d.close()
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pc), bits(g["pc"]))
        assert np.array_equal(bits(qc), bits(g["qc"]))
    else:
        e = rel_l2(pc, g["pc"]), rel_l2(qc, g["qc"])
        f64 = O.run_f64([np.zeros(pc.shape)] * 4, coef, dx, dx, dx, dt, steps,
                        src, H.ricker_table(dt, steps))
        say(f"fast vs reference relL2 p={e[0]:.3e} q={e[1]:.3e}; vs FP64 "
            f"p={rel_l2(pc, f64[1]):.3e} q={rel_l2(qc, f64[3]):.3e}")
        assert max(e) < FAST_TOL


@pytest.mark.parametrize("tiling", [(32, 16, 64, 1), (32, 16, 5, 1),
                                    (32, 16, 64, 2), (32, 16, 3, 2),
                                    (32, 16, 64, 3), (32, 16, 2, 3),
                                    (32, 32, 7, 1), (64, 16, 64, 2),
                                    (64, 8, 64, 2), (64, 8, 5, 1)])
def test_fast_kernel_ragged_grid_and_boxes(H, O, VM, tiling):
    """Odd sizes (tile overhang in x, y and z) and sub-boxes of the grid."""
    shape = (29, 41, 70)  # (sz, sy, sx)
    w = VM.random_small(shape, seed=7)
    rng = np.random.default_rng(11)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    say("ragged", tiling)
    for box in [((4, 4, 4), (66, 37, 25)), ((9, 5, 4), (40, 36, 11)),
                ((4, 4, 20), (8, 37, 25))]:
        d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt)
        assert d.set_tiling(*tiling[:3])
        d.set_prefetch_depth(tiling[3])
        d.propagate(*box)
        pp, pc, qp, qc = d.fields()
        d.close()
        ref = [a.copy() for a in f0]
        O.propagate_f32(box[0], box[1], ref[0], ref[1], ref[2], ref[3], w.coef,
                        w.dx, w.dy, w.dz, w.dt)
        for got, want, old in ((pp, ref[0], f0[0]), (qp, ref[2], f0[2])):
            (x0, y0, z0), (x1, y1, z1) = box
            inside = np.zeros(shape, bool)
            inside[z0:z1, y0:y1, x0:x1] = True
            assert np.array_equal(bits(got[~inside]), bits(old[~inside]))
            assert rel_l2(got[inside] - old[inside],
                          want[inside] - old[inside]) < FAST_TOL


@pytest.mark.parametrize("kernel", [1, 0])
def test_unequal_spacings(H, O, VM, kernel):
    """dx, dy and dz all different (10, 20 and 35 m), so that every one of
    the six inverse spacings 1/(dx dx) ... 1/(dy dz) differs: a swapped axis
    or spacing in either kernel changes the result.  A 30-step fused loop
    with a source: exact kernel bit-identical to the oracle, fast kernel
    within the bar."""
    w = VM.random_small((30, 34, 40), seed=19)
    dx, dy, dz = 10.0, 20.0, 35.0
    dt = float(np.float32(0.4 * VM.cfl_dt(w.coef[6], dx, dy, dz)))
    steps = 30
    amp = H.ricker_table(dt, steps)
    d = H.DeviceData(w.coef, None, dx, dy, dz, dt, kernel=kernel)
    d.run(steps, w.src_index, amp, graph=True)
    pc, qc = d.get_results()
    d.close()
    zeros = [np.zeros(w.shape, np.float32)] * 4
    ref = O.run_f32(zeros, w.coef, dx, dy, dz, dt, steps, w.src_index, amp)
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pc), bits(ref[1]))
        assert np.array_equal(bits(qc), bits(ref[3]))
    else:
        e = rel_l2(pc, ref[1]), rel_l2(qc, ref[3])
        say(f"unequal spacings: fast vs oracle p={e[0]:.3e} q={e[1]:.3e}")
        assert max(e) < FAST_TOL


def test_empty_boxes_are_no_ops(H, VM):
    """A box with end <= start on any axis updates nothing, as the
    reference's loops over [start, end) do nothing (src/openmp_propagate.c)."""
    shape = (20, 22, 24)
    w = VM.random_small(shape, seed=15)
    rng = np.random.default_rng(16)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    for kernel in (H.KERNEL_EXACT, H.KERNEL_FAST):
        d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
        for box in [((4, 4, 4), (4, 18, 16)), ((4, 4, 4), (20, 4, 16)),
                    ((4, 4, 9), (20, 18, 9)), ((10, 10, 10), (5, 5, 5))]:
            d.propagate(*box)
        got = d.fields()
        d.close()
        for a, b in zip(got, f0):
            assert np.array_equal(bits(a), bits(b)), kernel


@pytest.mark.parametrize("kernel", [1, 0])
def test_smallest_grids(H, O, VM, kernel):
    """9 x 10 x 11 arrays (compute regions of 1 x 2 x 3 cells, far smaller
    than one tile): the exact kernel bit-identical to the oracle, the fast
    kernel within the bar."""
    shape = (11, 10, 9)  # (sz, sy, sx)
    w = VM.random_small(shape, seed=17)
    rng = np.random.default_rng(18)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
    box = ((4, 4, 4), (5, 6, 7))
    d.propagate(*box)
    pp, pc, qp, qc = d.fields()
    d.close()
    ref = [a.copy() for a in f0]
    O.propagate_f32(box[0], box[1], ref[0], ref[1], ref[2], ref[3], w.coef,
                    w.dx, w.dy, w.dz, w.dt)
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pp), bits(ref[0]))
        assert np.array_equal(bits(qp), bits(ref[2]))
    else:
        inside = np.zeros(shape, bool)
        inside[4:7, 4:6, 4:5] = True
        for got, want, old in ((pp, ref[0], f0[0]), (qp, ref[2], f0[2])):
            assert np.array_equal(bits(got[~inside]), bits(old[~inside]))
            assert rel_l2(got[inside] - old[inside],
                          want[inside] - old[inside]) < FAST_TOL


def test_prefetch_depths_are_bitwise_equal(H, VM):
    """The prefetch depth changes only when loads are issued: 1, 2 and 3
    give the same bits over a fused 30-step loop with a source."""
    w = VM.random_small((36, 40, 52), seed=12, steps=30)
    amp = w.source_table()
    out = []
    for depth in (1, 2, 3):
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        d.set_tiling(32, 16, 16)
        assert d.set_prefetch_depth(depth) == 3   # the default
        d.run(30, w.src_index, amp, graph=True)
        out.append(d.fields())
        d.close()
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(bits(a), bits(b)), k


def test_tile_shapes_are_bitwise_equal(H, VM):
    """Every supported tile shape (32x16, 32x32, 64x16, 64x8) computes each
    cell with the same expressions: a fused 30-step loop with a source gives
    the same bits with all four.  The 52 x 42 x 36 compute cells leave
    ragged tiles in x and y for every shape.  Other shapes and a negative z
    chunk are refused and leave the tiling as it was."""
    w = VM.random_small((44, 50, 60), seed=15, steps=30)
    amp = w.source_table()
    out = []
    for tx, ty in ((32, 16), (32, 32), (64, 16), (64, 8)):
        say("tile", tx, ty)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        assert d.set_tiling(tx, ty, 0)
        for bad in ((16, 16, 0), (32, 8, 0), (128, 4, 0), (tx, ty, -1)):
            assert not d.set_tiling(*bad)
        d.run(30, w.src_index, amp, graph=True)
        out.append(d.fields())
        d.close()
    for k in (1, 2, 3):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(bits(a), bits(b)), k


@pytest.mark.parametrize("form", [(0, 0), (0, 4)])
@pytest.mark.parametrize("shape", [(24, 72, 136), (44, 50, 60)])
def test_xcd_tile_orders_are_bitwise_equal(H, VM, shape, form):
    """The workgroup-to-tile orders of dcx_set_xcd_remap (0 round robin, 1
    contiguous runs per XCD, 2 whole tile rows per XCD, -1 the default per
    storage form) only change which XCD runs a tile: a fused 20-step loop
    with a source gives the same bits, in plain FP32 and in the axis form
    (whose default order is 2).  The first grid (128 x 64 x 16 compute
    cells: 4 x 4 tiles of 32 x 16 in two 8-plane z chunks, 32 tiles in 8
    tile rows) is remapped by both modes; the second (52 x 42 x 36, ragged
    tiles, 30 tiles in 15 rows) suits neither, which then keep the default
    order."""
    w = VM.random_small(shape, seed=16, steps=20)
    amp = w.source_table()
    out = []
    for mode in (0, 1, 2, -1):
        say("xcd remap", mode, "shape", shape, "form", form)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        assert d.set_tiling(32, 16, 8)
        assert d.set_storage(*form)
        assert d.set_xcd_remap(mode) == -1
        assert d.set_xcd_remap(mode) == mode
        d.run(20, w.src_index, amp, graph=True)
        out.append(d.fields())
        d.close()
    assert np.abs(out[0][1]).max() > 0
    for k in (1, 2, 3):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(bits(a), bits(b)), k


@pytest.mark.parametrize("kernel", [1, 0])
def test_reference_worker_seven_calls_equal_one_call(H, VM, kernel):
    """The reference's worker updates a step as six 4-cell boundary slabs,
    then the interior (src/worker.c:321-402).  The fast kernel starts each
    box's tile grid on the last 128-byte line boundary at or before its x0
    (the interior and the y and z slabs start at x = 8) and masks the cells
    before x0.  Each cell's update is the same, so 20 steps with a source
    give the bits of one whole-region call per step."""
    w = VM.random_small((44, 50, 60), seed=16, steps=20)
    amp = w.source_table()
    sx, sy, sz = w.sizes
    S = 4
    slabs = [((S, S, S), (2 * S, sy - S, sz - S)),
             ((sx - 2 * S, S, S), (sx - S, sy - S, sz - S)),
             ((2 * S, S, S), (sx - 2 * S, 2 * S, sz - S)),
             ((2 * S, sy - 2 * S, S), (sx - 2 * S, sy - S, sz - S)),
             ((2 * S, 2 * S, S), (sx - 2 * S, sy - 2 * S, 2 * S)),
             ((2 * S, 2 * S, sz - 2 * S), (sx - 2 * S, sy - 2 * S, sz - S))]
    interior = ((2 * S,) * 3, (sx - 2 * S, sy - 2 * S, sz - 2 * S))
    whole = ((S,) * 3, (sx - S, sy - S, sz - S))
    out = []
    for boxes in (slabs + [interior], [whole]):
        say("worker calls", len(boxes), "kernel", kernel)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
        for it in range(20):
            d.add_source(w.src_index, float(amp[it]))
            for b in boxes:
                d.propagate(*b)
            d.swap()
        out.append(d.fields())
        d.close()
    assert np.abs(out[1][1]).max() > 0
    for a, b in zip(*out):
        assert np.array_equal(bits(a), bits(b))


def test_default_z_chunk_is_bitwise_equal_to_whole_columns(H, VM):
    """The default z chunk is chosen per launch (two tiles per CU): the
    192 x 192 x 112 compute cells of these arrays make 72 columns of 32 x 16
    cells, so the launch splits them into 32-plane chunks.  Every cell's
    arithmetic is the same in any chunk: the fields match whole 512-plane
    columns bit for bit."""
    w = VM.random_small((120, 200, 200), seed=14, steps=12)
    amp = w.source_table()
    out = []
    for zc in (0, 512):
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        assert d.set_tiling(32, 16, zc)
        d.run(12, w.src_index, amp, graph=True)
        out.append(d.fields())
        d.close()
    for a, b in zip(*out):
        assert np.array_equal(bits(a), bits(b))


def test_stream_probe_touches_the_compute_region_only(H, VM):
    """The stencil-free probe (one untimed pass + reps timed passes) writes
    pp <- 2 pc - pp + dt * sum(coef) with dt = 0 at every compute cell and
    nothing else: three passes leave 2 pc - pp there, the rest bit for bit."""
    shape = (21, 26, 300)  # (sz, sy, sx): two 256-cell rows along x
    w = VM.random_small(shape, seed=3)
    rng = np.random.default_rng(4)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt)
    assert d.stream_probe(2) > 0
    pp, pc, qp, qc = d.fields()
    d.close()
    inside = np.zeros(shape, bool)
    inside[4:-4, 4:-4, 4:-4] = True
    for got, old, cur in ((pp, f0[0], f0[1]), (qp, f0[2], f0[3])):
        assert np.array_equal(bits(got[~inside]), bits(old[~inside]))
        want = 2.0 * cur[inside].astype(np.float64) - old[inside]
        assert np.max(np.abs(got[inside] - want)) < 1e-5
    assert np.array_equal(bits(pc), bits(f0[1]))
    assert np.array_equal(bits(qc), bits(f0[3]))


def test_halo_extract_insert_bitwise(H, O, VM):
    shape = (20, 22, 24)
    w = VM.random_small(shape, seed=5)
    rng = np.random.default_rng(2)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    d = H.DeviceData(w.coef, f0)
    for a, b in [((4, 4, 4), (8, 18, 16)), ((0, 0, 0), (24, 22, 20)),
                 ((20, 18, 16), (24, 22, 20)), ((5, 7, 3), (6, 8, 4))]:
        for name, k in (("pp", 0), ("qc", 3)):
            got = d.extract_box(a, b, name)
            want = O.extract_box(f0[k], a, b)
            assert np.array_equal(bits(got), bits(want))
    buf = rng.standard_normal(4 * 14 * 12).astype(np.float32)
    d.insert_box(buf, (0, 4, 4), (4, 18, 16), "qp")
    want = f0[2].copy()
    O.insert_box(want, buf, (0, 4, 4), (4, 18, 16))
    assert np.array_equal(bits(d.download(2)), bits(want))
    d.close()


@pytest.mark.parametrize("kernel", [1, 0])
def test_arrays_past_2_gib(H, O, VM, kernel):
    """Arrays of 1000 x 40 x 13,200 cells: with rows pitched to 1024 floats
    each array spans 2.16 GB, so its last planes sit past 2^31 bytes, where a
    signed 32-bit byte offset would wrap (the fast kernel's limit is 4 GiB).
    A box in those planes matches the oracle (exact kernel bit for bit, fast
    kernel within 1e-5) and nothing below the box changes.  Only the last 16
    planes of the host arrays are written; the rest stay untouched zero
    pages, so host memory stays small."""
    sz, sy, sx = 13_200, 40, 1000
    tail = 16
    assert (sz - tail) * 1024 * sy * 4 > 2 ** 31
    small = VM.random_small((tail, sy, sx), seed=21)
    rng = np.random.default_rng(22)
    coef, f0 = [], []
    for c in small.coef:
        a = np.zeros((sz, sy, sx), np.float32)
        a[-tail:] = c
        coef.append(a)
    for _ in range(4):
        a = np.zeros((sz, sy, sx), np.float32)
        a[-tail:] = rng.standard_normal((tail, sy, sx)).astype(np.float32)
        f0.append(a)
    ref = [a[-tail:].copy() for a in f0]
    O.propagate_f32((4, 4, 4), (sx - 4, sy - 4, 12), ref[0], ref[1], ref[2],
                    ref[3], small.coef, small.dx, small.dy, small.dz, small.dt)
    say("past 2 GiB, kernel", kernel)
    d = H.DeviceData(coef, f0, small.dx, small.dy, small.dz, small.dt,
                     kernel=kernel)
    d.propagate((4, 4, sz - 12), (sx - 4, sy - 4, sz - 4))
    inside = np.zeros((tail, sy, sx), bool)
    inside[4:12, 4:sy - 4, 4:sx - 4] = True
    for name, want, old in (("pp", ref[0], f0[0][-tail:]),
                            ("qp", ref[2], f0[2][-tail:])):
        got = d.extract_box((0, 0, sz - tail), (sx, sy, sz), name).reshape(
            tail, sy, sx)
        assert np.array_equal(bits(got[~inside]), bits(old[~inside]))
        if kernel == 1:
            assert np.array_equal(bits(got[inside]), bits(want[inside]))
        else:
            assert rel_l2(got[inside] - old[inside],
                          want[inside] - old[inside]) < FAST_TOL
        below = d.extract_box((0, 0, sz - tail - 4), (sx, sy, sz - tail), name)
        assert not below.any()
    d.close()


def test_multi_box_pack_unpack_bitwise(H, VM):
    """Many boxes per launch (dcx_pack_boxes / dcx_unpack_boxes): faces one
    to four cells thin, an edge, a corner, one cell, an empty box, a ragged
    interior box and the array's last row.  The packed stream is unpacked as
    consecutive x-rows of another field, so that field's leading cells hold
    the stream itself: box after box, each x fastest, then y, then z (the
    order of src/device_data.c:61-92).  Then the inverse: rows packed and
    unpacked into the boxes.  Hundreds of row boxes also exercise the box
    search of one launch."""
    shape = (23, 37, 45)  # (sz, sy, sx)
    sz, sy, sx = shape
    w = VM.random_small(shape, seed=9)
    rng = np.random.default_rng(4)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    boxes = [((4, 4, 4), (8, 33, 19)), ((8, 4, 4), (41, 8, 19)),
             ((8, 8, 15), (41, 33, 19)), ((0, 0, 0), (4, 4, 23)),
             ((41, 33, 19), (45, 37, 23)), ((2, 30, 20), (3, 31, 21)),
             ((5, 5, 5), (5, 9, 9)), ((8, 8, 2), (40, 29, 12)),
             ((0, 36, 22), (41, 37, 23)), ((44, 0, 0), (45, 33, 19)),
             ((4, 3, 0), (43, 4, 23))]
    # Disjoint, as halo boxes are: one launch writes its boxes in no set
    # order, so overlapping destinations would have no defined result.
    cover = np.zeros(shape, int)
    for (x0, y0, z0), (x1, y1, z1) in boxes:
        cover[z0:z1, y0:y1, x0:x1] += 1
    assert cover.max() == 1
    stream = np.concatenate([f0[1][z0:z1, y0:y1, x0:x1].ravel()
                             for (x0, y0, z0), (x1, y1, z1) in boxes])
    total = stream.size
    rows = []
    for t in range(-(-total // sx)):
        z, y = divmod(t, sy)
        rows.append(((0, y, z), (min(sx, total - t * sx), y + 1, z + 1)))
    assert 256 < len(rows) <= 512
    say("multi-box", len(boxes), "boxes,", len(rows), "rows,", total, "cells")
    d = H.DeviceData(w.coef, f0)
    d.pack_unpack(boxes, "pc", rows, "qp")
    want = f0[2].copy()
    want.reshape(-1)[:total] = stream
    assert np.array_equal(bits(d.download(2)), bits(want))
    d.pack_unpack(rows, "qc", boxes, "pp")
    want = f0[0].copy()
    src = f0[3].reshape(-1)
    pos = 0
    for (x0, y0, z0), (x1, y1, z1) in boxes:
        n = (x1 - x0) * (y1 - y0) * (z1 - z0)
        want[z0:z1, y0:y1, x0:x1] = src[pos:pos + n].reshape(
            z1 - z0, y1 - y0, x1 - x0)
        pos += n
    assert np.array_equal(bits(d.download(0)), bits(want))
    d.close()


@pytest.mark.parametrize("case", ["a", "b"])
@pytest.mark.parametrize("ranks,topo", [(1, (1, 1, 1)), (8, (2, 2, 2))])
@pytest.mark.parametrize("kernel", [1, 0])
def test_program_matches_reference_output(H, case, ranks, topo, kernel):
    g = golden(f"golden_dc_{case}_np{ranks}.npz")
    say("program", case, ranks, kernel)
    size = tuple(int(v) for v in g["size"])
    pc, qc, st = H.program_run(size, int(g["absorption"]), float(g["d"]),
                               float(np.float32(float(g["dt"]))),
                               float(np.float32(float(g["tmax"]))), topo,
                               kernel)
    assert st["iterations"] == PO.iterations(str(g["tmax"]), str(g["dt"]))
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pc), bits(g["pc"]))
        assert np.array_equal(bits(qc), bits(g["qc"]))
    else:
        assert rel_l2(pc, g["pc"]) < FAST_TOL and rel_l2(qc, g["qc"]) < FAST_TOL


def test_program_uneven_topology_matches_oracle(H):
    """3x2x1 ranks: remainders on the last rank of each axis."""
    size, ab = (13, 11, 9), 3
    pc, qc, _ = H.program_run(size, ab, 20.0, 0.002, 0.03, (3, 2, 1), 1)
    want = PO.run_program(size, ab, 20.0, "0.002", "0.03", (3, 2, 1))
    assert np.array_equal(bits(pc), bits(want[0]))
    assert np.array_equal(bits(qc), bits(want[1]))


def test_smooth_random_200_steps(H, O, VM):
    """The 104^3 smooth-random workload, 200 steps, fast kernel vs oracle."""
    w = VM.smooth_random(seed=0)
    say("smooth_random 104^3 x200")
    amp = w.source_table()
    d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
    d.run(w.steps, w.src_index, amp, graph=True)
    pc, qc = d.get_results()
    d.close()
    zeros = [np.zeros(w.shape, np.float32)] * 4
    ref = O.run_f32(zeros, w.coef, w.dx, w.dy, w.dz, w.dt, w.steps,
                    w.src_index, amp)
    f64 = O.run_f64(zeros, w.coef, w.dx, w.dy, w.dz, w.dt, w.steps,
                    w.src_index, amp)
    rec = w.receivers
    e = rel_l2(pc, ref[1]), rel_l2(qc, ref[3])
    tr = rel_l2(pc.reshape(-1)[rec], ref[1].reshape(-1)[rec])
    say(f"smooth_random relL2 vs reference FP32: p={e[0]:.3e} q={e[1]:.3e} "
        f"receivers={tr:.3e}")
    say(f"  vs FP64: fast p={rel_l2(pc, f64[1]):.3e} "
        f"receivers={rel_l2(pc.reshape(-1)[rec], f64[1].reshape(-1)[rec]):.3e}; "
        f"reference FP32 p={rel_l2(ref[1], f64[1]):.3e} receivers="
        f"{rel_l2(ref[1].reshape(-1)[rec], f64[1].reshape(-1)[rec]):.3e}")
    assert max(e) < FAST_TOL and tr < FAST_TOL


def test_full_size_octant_against_oracle(H, O, VM):
    """The benchmark's step itself: the 520^3 octant workload with the 7-box
    halo loopback registered as bench.py registers it, 30 steps with the
    source.  The fields start from small seeded noise everywhere, so that the
    halos the loopback copies are not zero.  The oracle runs the reference's
    per-step order with its own box extract and insert on the host
    (tests/test_gpu_loopback.py:oracle_loop); exact kernel bit-identical,
    fast kernel's change over the 30 steps within the 1e-5 bar."""
    import bench
    from test_gpu_loopback import check_boxes_copied, oracle_loop
    w = VM.layered_octant(seed=0, local=512)
    steps = 30
    amp = H.ricker_table(w.dt, steps, w.fcut)
    send, recv = bench.octant_halo_boxes(w.sizes)
    rng = np.random.default_rng(41)
    f0 = [np.float32(1e-3) * rng.standard_normal(w.shape, dtype=np.float32)
          for _ in range(4)]
    say("octant 520^3: oracle", steps, "steps with the loopback")
    ref = oracle_loop(O, w, f0, steps, w.src_index, amp, send, recv)
    for kernel in (H.KERNEL_EXACT, H.KERNEL_FAST):
        say("octant 520^3: kernel", kernel)
        d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
        d.release_host()
        d.set_halo_loopback(send, recv)
        d.run(steps, w.src_index, amp, graph=True)
        pc, qc = d.get_results()
        d.close()
        for f in (pc, qc):
            check_boxes_copied(f, send, recv)
        if kernel == H.KERNEL_EXACT:
            assert np.array_equal(bits(pc), bits(ref[1]))
            assert np.array_equal(bits(qc), bits(ref[3]))
        else:
            e = (rel_l2(pc - f0[1], ref[1] - f0[1]),
                 rel_l2(qc - f0[3], ref[3] - f0[3]))
            say(f"octant fast vs oracle (change over {steps} steps): "
                f"p={e[0]:.3e} q={e[1]:.3e}")
            assert max(e) < FAST_TOL
