"""The halo loopback every benchmark step runs (dcx_set_halo_loopback).

After each step's stencil launch the library packs the send boxes of the new
p and q into one device buffer and unpacks it into the receive boxes of the
same arrays: the on-device half of the reference's exchange
(src/worker.c:160-319 send and receive, 657-718 insert) with the transfer
taken out.  The oracle for one step is therefore its own update followed by
its own box extract and insert on the new fields, then the swap; the
reference's per-step order (source, update, exchange, swap;
src/worker.c:466-511) fixes where the source goes.

Bars: the exact kernel is bit-identical to that oracle loop; the fast kernel
is within 1e-5 relative L2 of it, and in every case each receive box holds,
bit for bit, the send box of the same field after the same step.
"""
import sys

import numpy as np
import pytest

import bench

S = 4
FAST_TOL = 1e-5


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(b), 1e-300))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def box_slice(box):
    (x0, y0, z0), (x1, y1, z1) = box
    return (slice(z0, z1), slice(y0, y1), slice(x0, x1))


def test_octant_boxes_geometry():
    """The 7 boxes the benchmark registers (bench.octant_halo_boxes) have the
    geometry of src/worker.c:196-228 and 679-711, checked by properties and
    not by restating the formula: each receive box is its send box moved 4
    cells along the neighbour's direction; send boxes lie in the compute
    region; receive boxes lie in the halo frame, are disjoint, and together
    cover exactly the frame cells on the high sides of the compute region's
    extension (every coordinate in [4, s), at least one in [s-4, s))."""
    sizes = (37, 30, 26)
    send, recv = bench.octant_halo_boxes(sizes)
    assert len(send) == len(recv) == 7
    cover = np.zeros(sizes[::-1], np.int32)
    dirs = set()
    for (s0, s1), (r0, r1) in zip(send, recv):
        d = tuple(int(r0[k] != s0[k]) for k in range(3))
        dirs.add(d)
        for k in range(3):
            assert r0[k] - s0[k] == S * d[k] and r1[k] - s1[k] == S * d[k]
            assert S <= s0[k] < s1[k] <= sizes[k] - S
            if d[k]:
                assert (r0[k], r1[k]) == (sizes[k] - S, sizes[k])
        cover[box_slice((r0, r1))] += 1
    assert dirs == {(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)} - {(0, 0, 0)}
    want = np.zeros(sizes[::-1], bool)
    want[S:, S:, S:] = True
    want[S:-S, S:-S, S:-S] = False
    assert cover.max() == 1
    assert np.array_equal(cover.astype(bool), want)


def oracle_loop(O, w, fields, steps, src, amp, send, recv):
    """src/worker.c:466-511 for one rank whose neighbours are itself: add the
    source to pc and qc, update the compute region, copy each send box of the
    new pp and qp into its receive box, swap."""
    pp, pc, qp, qc = [np.ascontiguousarray(a, np.float32).copy() for a in fields]
    sx, sy, sz = w.sizes
    for it in range(steps):
        if src >= 0:
            pc.reshape(-1)[src] += np.float32(amp[it])
            qc.reshape(-1)[src] += np.float32(amp[it])
        O.propagate_f32((S, S, S), (sx - S, sy - S, sz - S), pp, pc, qp, qc,
                        w.coef, w.dx, w.dy, w.dz, w.dt)
        for (s0, s1), (r0, r1) in zip(send, recv):
            for arr in (pp, qp):
                O.insert_box(arr, O.extract_box(arr, s0, s1), r0, r1)
        pp, pc, qp, qc = pc, pp, qc, qp
    return pp, pc, qp, qc


def check_boxes_copied(field, send, recv, skip=-1):
    """Each receive box equals its send box bit for bit; the cell of linear
    index `skip` (a source cell that already holds the next step's
    amplitude) and its image in the receive box are left out."""
    keep = np.ones(field.shape, bool)
    if skip >= 0:
        keep.reshape(-1)[skip] = False
    for s, r in zip(send, recv):
        m = keep[box_slice(s)]
        assert np.array_equal(bits(field[box_slice(r)])[m],
                              bits(field[box_slice(s)])[m])


def _source_index(w, where):
    sx, sy, sz = w.sizes
    if where == "interior":
        x, y, z = sx // 2 - 3, sy // 2, sz // 2 - 2
    else:  # inside the +x+y+z corner's send box (and the +x face's)
        x, y, z = sx - 6, sy - 7, sz - 5
    return x + sx * (y + sy * z)


@pytest.mark.gpu
@pytest.mark.parametrize("where", ["interior", "send_box"])
@pytest.mark.parametrize("mode", ["step", "run", "graph"])
@pytest.mark.parametrize("kernel", [1, 0])
def test_loopback_against_oracle(H, O, VM, kernel, mode, where):
    shape = (26, 30, 37)  # (sz, sy, sx): tiles overhang in x, y and z
    w = VM.random_small(shape, seed=23)
    rng = np.random.default_rng(5)
    f0 = [0.1 * rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    steps = 10
    src = _source_index(w, where)
    amp = (50.0 * H.ricker_table(w.dt, steps, w.fcut)).astype(np.float32)
    send, recv = bench.octant_halo_boxes(w.sizes)
    say("loopback kernel", kernel, "mode", mode, "source", where)
    d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
    d.set_halo_loopback(send, recv)
    if mode == "step":
        d.add_source(src, float(amp[0]))
        for it in range(steps):
            more = it + 1 < steps
            d.step(src, float(amp[it + 1]) if more else 0.0, more)
            # pc and qc already hold the next step's source at its cell.
            for which in (1, 3):
                check_boxes_copied(d.download(which), send, recv,
                                   src if more else -1)
    else:
        d.run(steps, src, amp, graph=(mode == "graph"))
    pc, qc = d.get_results()
    d.close()
    for f in (pc, qc):
        check_boxes_copied(f, send, recv)
    ref = oracle_loop(O, w, f0, steps, src, amp, send, recv)
    if kernel == H.KERNEL_EXACT:
        assert np.array_equal(bits(pc), bits(ref[1]))
        assert np.array_equal(bits(qc), bits(ref[3]))
    else:
        e = rel_l2(pc - f0[1], ref[1] - f0[1]), rel_l2(qc - f0[3], ref[3] - f0[3])
        say(f"loopback fast vs oracle (change over {steps} steps): "
            f"p={e[0]:.3e} q={e[1]:.3e}")
        assert max(e) < FAST_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["run", "graph"])
def test_loopback_fp16_storage(H, O, VM, mode):
    """FP16 field storage: the loopback copies the new P level's 16-bit
    values, so the decoded receive boxes equal the decoded send boxes bit for
    bit; the fields stay within the half form's error of the oracle."""
    shape = (26, 30, 37)
    w = VM.random_small(shape, seed=29)
    steps = 30
    src = _source_index(w, "interior")
    amp = H.ricker_table(w.dt, steps, w.fcut)
    send, recv = bench.octant_halo_boxes(w.sizes)
    zeros = [np.zeros(shape, np.float32) for _ in range(4)]
    d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
    assert d.set_storage(H.STORE_F16, H.COEF_F32)
    d.set_halo_loopback(send, recv)
    d.run(steps, src, amp, graph=(mode == "graph"))
    pc, qc = d.get_results()
    d.close()
    for f in (pc, qc):
        check_boxes_copied(f, send, recv)
    ref = oracle_loop(O, w, zeros, steps, src, amp, send, recv)
    e = rel_l2(pc, ref[1]), rel_l2(qc, ref[3])
    say(f"loopback FP16 storage vs oracle: p={e[0]:.3e} q={e[1]:.3e}")
    assert max(e) < 2e-2


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 0])
def test_reference_worker_exchange_pattern(H, O, VM, kernel):
    """The reference worker's step with neighbours, through the ABI
    (src/worker.c:466-511): add the source, update the six boundary slabs,
    extract the send boxes of pp and qp (which launches the queued slabs,
    here as one multi-box launch), update the interior, insert the halos
    (the send boxes, looped back), swap.  Exact kernel: bit-identical to the
    oracle loop; fast kernel: bit-identical to the fused step with the
    loopback registered (one launch over the compute region per step)."""
    import program_oracle as PO
    shape = (40, 44, 52)
    w = VM.random_small(shape, seed=31)
    steps = 12
    src = _source_index(w, "interior")
    amp = H.ricker_table(w.dt, steps, w.fcut)
    send, recv = bench.octant_halo_boxes(w.sizes)
    regions = PO.slabs(w.sizes)
    assert len(regions) == 7
    d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
    for it in range(steps):
        d.add_source(src, float(amp[it]))
        for b in regions[:6]:
            d.propagate(*b)
        halos = [(d.extract_box(s0, s1, "pp"), d.extract_box(s0, s1, "qp"))
                 for s0, s1 in send]
        d.propagate(*regions[6])
        for (r0, r1), (hp, hq) in zip(recv, halos):
            d.insert_box(hp, r0, r1, "pp")
            d.insert_box(hq, r0, r1, "qp")
        d.swap()
    got = d.get_results()
    d.close()
    if kernel == H.KERNEL_EXACT:
        zeros = [np.zeros(shape, np.float32) for _ in range(4)]
        ref = oracle_loop(O, w, zeros, steps, src, amp, send, recv)
        want = (ref[1], ref[3])
    else:
        f = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        f.set_halo_loopback(send, recv)
        f.run(steps, src, amp, graph=True)
        want = f.get_results()
        f.close()
    assert np.abs(want[0]).max() > 0
    for a, b in zip(got, want):
        assert np.array_equal(bits(a), bits(b))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 0])
def test_overlapping_propagate_calls_run_in_order(H, O, VM, kernel):
    """dc_propagate calls are queued and launched together; a box that
    overlaps a queued one must still see the queued update first, as two
    separate calls on the reference do (the second reads pp as the first
    left it).  Exact kernel bit-identical to the oracle applying the calls in
    order; fast kernel within the bar."""
    shape = (30, 34, 40)
    w = VM.random_small(shape, seed=33)
    rng = np.random.default_rng(34)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    calls = [((4, 4, 4), (20, 30, 26)), ((10, 12, 4), (36, 30, 20)),
             ((30, 4, 20), (36, 10, 26)), ((4, 4, 4), (36, 30, 26))]
    d = H.DeviceData(w.coef, f0, w.dx, w.dy, w.dz, w.dt, kernel=kernel)
    for b in calls:
        d.propagate(*b)
    got = d.fields()
    d.close()
    ref = [a.copy() for a in f0]
    for b in calls:
        O.propagate_f32(b[0], b[1], ref[0], ref[1], ref[2], ref[3], w.coef,
                        w.dx, w.dy, w.dz, w.dt)
    for k in (0, 2):
        if kernel == H.KERNEL_EXACT:
            assert np.array_equal(bits(got[k]), bits(ref[k]))
        else:
            assert rel_l2(got[k] - f0[k], ref[k] - f0[k]) < FAST_TOL
    for k in (1, 3):
        assert np.array_equal(bits(got[k]), bits(f0[k]))
