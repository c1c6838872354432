"""Benchmark: Fletcher TTI time steps of one rank's 512^3 share on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the plain-FP32 configuration the metric is quoted on): the (0,0,0)
octant of a 1024^3 domain split 2x2x2 -- 520^3 local arrays (512^3 compute
cells + 4-cell halo), a 32-cell random-velocity absorbing shell on its three
outer faces, a 10-layer 1500->4500 m/s model with 5% seeded noise and random
TTI anisotropy, a 15 Hz Ricker source; every step is one fused stencil launch
plus on-device packing of the 7 halo boxes (3 faces, 3 edges, 1 corner) of p
and q and unpacking them into the receive boxes (loopback: the transfer to
neighbour GPUs is the part one GPU cannot show).

Timing: W untimed steps, then exactly K steps between two stream
synchronisations (wall clock); value = 512^3 * K / seconds / 1e6.  The
stencil kernel's average launch time comes from HIP events around each
launch on the library's stream (a second pass), and gives the roofline
fraction against MI355X's 8 TB/s with 64 algorithmic bytes per cell-update
(read pc, qc, pp, qp and ten coefficient arrays, write pp, qp), and against
a stencil-free pass that streams the same 64 B per cell over the same arrays
(dcx_stream_probe: what the memory system gives this byte pattern).  HBM traffic
per launch is read from two rocprofv3 --pmc child runs (FETCH_SIZE,
WRITE_SIZE; --no-traffic skips them) scaled by the same counters on a copy
kernel of known size.  The CPU baseline runs the reference's own OpenMP dc_propagate
(oracle/_ref, when built) or the oracle's C restatement on the host cores,
on a bounded sample of the same workload: a few steps of the 520^3 arrays
from the GPU's field state after the timed steps.
"""
from __future__ import annotations

import argparse
import glob
import importlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("Mcell-updates/s on 1 GPU, mixed/batched vs plain FP32, "
          "at <=1e-3 trace err")
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_FP32 = 64
# Published single-GPU Fletcher step throughput at 512^3 (MSamples/s =
# Mcell-updates/s): the reference's single-node baseline code on an RTX 4070.
PUBLISHED_512 = 3870.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def octant_halo_boxes(sizes):
    """Send and receive boxes of the (0,0,0) rank's 7 neighbours (+x,+y,+z
    and their combinations), src/worker.c:196-228 and 679-711."""
    S = 4
    send, recv = [], []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                if dx == dy == dz == 0:
                    continue
                a0, a1, b0, b1 = [], [], [], []
                for d, s in zip((dx, dy, dz), sizes):
                    if d == 1:
                        a0.append(s - 2 * S); a1.append(s - S)
                        b0.append(s - S); b1.append(s)
                    else:
                        a0.append(S); a1.append(s - S)
                        b0.append(S); b1.append(s - S)
                send.append((tuple(a0), tuple(a1)))
                recv.append((tuple(b0), tuple(b1)))
    return send, recv


def reference_call_pattern(d, sizes, steps=50):
    """The step as the reference's worker drives a backend
    (src/worker.c:321-402): dc_propagate on the six 4-cell boundary slabs
    (x, then y over x in [8, sx-8), then z over x, y in [8, s-8); low side
    first), then on the interior [8, s-8)^3, then dc_device_swap_arrays.
    Beside it, one dc_propagate over the whole compute region.  Wall-clock
    ms per step between stream synchronisations, after one untimed step.
    The struct runs with the library's default tiling, as a reference
    worker's would (32 x 16 tiles, z chunk chosen per launch); the caller
    restores its own tiling afterwards."""
    assert d.set_tiling(32, 16, 0)
    sx, sy, sz = sizes
    S = 4
    slabs = [((S, S, S), (2 * S, sy - S, sz - S)),
             ((sx - 2 * S, S, S), (sx - S, sy - S, sz - S)),
             ((2 * S, S, S), (sx - 2 * S, 2 * S, sz - S)),
             ((2 * S, sy - 2 * S, S), (sx - 2 * S, sy - S, sz - S)),
             ((2 * S, 2 * S, S), (sx - 2 * S, sy - 2 * S, 2 * S)),
             ((2 * S, 2 * S, sz - 2 * S), (sx - 2 * S, sy - 2 * S, sz - S))]
    interior = ((2 * S, 2 * S, 2 * S), (sx - 2 * S, sy - 2 * S, sz - 2 * S))
    whole = ((S, S, S), (sx - S, sy - S, sz - S))
    out = {}
    # seven_calls: one rank with no neighbours (the queued boxes tile the
    # compute region: one launch).  seven_calls_exchange: the slabs launched
    # before the interior, as a worker with neighbours must (its halo
    # extract needs them; dcx_synchronize stands in for the extract here):
    # the six slabs as one multi-box launch, then the interior.
    for name, boxes, sync_after in (
            ("seven_calls", slabs + [interior], None),
            ("seven_calls_exchange", slabs + [interior], 6),
            ("one_call", [whole], None)):
        for _ in range(2):
            for b in boxes:
                d.propagate(*b)
            d.swap()
        d.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for k, b in enumerate(boxes):
                if k == sync_after:
                    d.synchronize()
                d.propagate(*b)
            d.swap()
        d.synchronize()
        out[name + "_ms"] = round((time.perf_counter() - t0) * 1e3 / steps, 4)
    out["steps"] = steps
    return out


def make_workload(VM, args):
    t = time.time()
    w = VM.layered_octant(seed=args.seed, local=args.local)
    log(f"[bench] workload {w.name} {w.sizes} built in {time.time() - t:.1f}s")
    return w


def open_device(H, w, args):
    d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt, kernel=H.KERNEL_FAST)
    tx, ty, zc = args.tiling
    if not d.set_tiling(tx, ty, zc):
        raise SystemExit(f"unsupported tiling {args.tiling}")
    d.set_xcd_remap(args.xcd_remap)
    d.set_prefetch_depth(args.prefetch)
    send, recv = octant_halo_boxes(w.sizes)
    d.set_halo_loopback(send, recv)
    return d


def cpu_baseline(w, fields, steps, log_prefix="[bench]"):
    """Time `steps` one-rank steps of the reference path on the host."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    S = 4
    sx, sy, sz = w.sizes
    box = ((S, S, S), (sx - S, sy - S, sz - S))
    f = [np.ascontiguousarray(a).copy() for a in fields]
    kind = "port"
    prop = O.propagate_f32
    try:
        if O.ref_available():
            O.ref()
            kind, prop = "reference", O.ref_propagate
    except OSError as e:
        log(f"{log_prefix} reference backend not loadable ({e}); using the port")
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {}
    for mode in ("default", "ftz"):
        O.set_ftz(mode == "ftz")
        g = [a.copy() for a in f]
        t = time.perf_counter()
        for _ in range(steps):
            prop(box[0], box[1], g[0], g[1], g[2], g[3], w.coef, w.dx, w.dy,
                 w.dz, w.dt)
            g = [g[1], g[0], g[3], g[2]]
        sec = time.perf_counter() - t
        out[mode] = w.compute_cells() * steps / sec / 1e6
        log(f"{log_prefix} cpu {kind} {mode}: {out[mode]:.2f} Mcell/s "
            f"({steps} steps, {sec:.1f}s, {cores} threads)")
    O.set_ftz(False)
    return {"value": round(out["default"], 3), "unit": "Mcell-updates/s",
            "cores": cores, "kind": kind,
            "sample": f"{steps} steps of the full {sx}x{sy}x{sz} arrays from the "
                      f"GPU's field state after the timed steps, default FP "
                      f"environment (denormals live, as the reference runs)",
            "value_ftz_daz": round(out["ftz"], 3)}


# Storage forms measured beside the plain FP32 step: (name, field, coef).
FORMS = [("fp32", 0, 0), ("fp32_n16coef", 0, 4), ("fp32_u16coef", 0, 2),
         ("fp32_a16coef", 0, 3), ("fp16_f32coef", 1, 0), ("fp16", 1, 1),
         ("fp16_u16coef", 1, 2), ("bf16_u16coef", 2, 2)]
# Half field storage scale 2^8: the field peaks near 40 (true units) at the
# source cell and half's largest finite value is 65504.
FIELD_SCALE_LOG2 = 8


def time_forms(H, w, d, args, amp, cells):
    """Mcell/s and stencil launch time of each storage form on the same rank
    (same model, halo loopback on); the struct is converted in place."""
    out = {}
    base = None
    for name, fs, cs in FORMS:
        if not d.set_storage(fs, cs, FIELD_SCALE_LOG2):
            continue
        d.run(5, w.src_index, amp[:5])
        d.synchronize()
        t0 = time.perf_counter()
        d.run(args.form_steps, w.src_index, amp[:args.form_steps])
        d.synchronize()
        sec = time.perf_counter() - t0
        d.set_timing_detail(True)
        d.run(20, -1, None)
        kms = d.last_kernel_ms()
        d.set_timing_detail(False)
        bpc = d.bytes_per_cell()
        v = cells * args.form_steps / sec / 1e6
        base = base or v
        out[name] = {"mcell_per_s": round(v, 1), "kernel_ms": round(kms, 4),
                     "bytes_per_cell": bpc,
                     "hbm_frac": round(bpc * cells / (kms * 1e-3) / 1e9 /
                                       HBM_PEAK_GBS, 4),
                     "speedup_vs_fp32": round(v / base, 3)}
        log(f"[bench] form {name}: {v:.0f} Mcell/s, launch {kms:.3f} ms, "
            f"{bpc} B/cell")
    d.set_storage(0, 0, FIELD_SCALE_LOG2)
    return out


def precision_vs_fp32(H, VM, w, steps=1000):
    """The compact forms' traces against the plain FP32 kernel's, GPU only:
    the 136^3 crop of the benchmark's own model around its source, `steps`
    steps, p at the 128x128 compute cells of the first plane below the
    absorbing shell.  Relative L2 of each form's traces (and final p) from
    the FP32 run's.  With the FP32 kernel's own trace error against FP64
    (tests/test_gpu_precision.py asserts it), this bounds each form's error
    against FP64 from above."""
    c = VM.crop(w, (208, 208, 0), 136, receiver_z=36)
    amp = H.ricker_table(c.dt, steps, c.fcut)
    out = {"workload": f"136^3 crop of {w.name}, {steps} steps, 128x128 "
                       f"receivers", "vs": "fp32 kernel"}
    ref = None
    for name, fs, cs in (("fp32", 0, 0), ("fp32_n16coef", 0, 4),
                         ("fp32_a16coef", 0, 3), ("fp16", 1, 1)):
        d = H.DeviceData(c.coef, None, c.dx, c.dy, c.dz, c.dt)
        d.set_storage(fs, cs, FIELD_SCALE_LOG2)
        tr, _ = d.run_traces(steps, c.src_index, amp, c.receivers)
        pc, _ = d.get_results()
        d.close()
        if ref is None:
            ref = (tr.astype(np.float64), pc.astype(np.float64))
            continue

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out[name] = {"trace_rel_l2": float(f"{rel(tr, ref[0]):.4g}"),
                     "wavefield_rel_l2": float(f"{rel(pc, ref[1]):.4g}"),
                     "finite": bool(np.isfinite(tr).all() and np.isfinite(pc).all())}
        log(f"[bench] precision vs FP32 kernel, {steps} steps, {name}: "
            f"traces {out[name]['trace_rel_l2']:.3e}")
    return out


def precision_study(H, VM, steps_list, log_prefix="[bench]"):
    """Relative L2 against the FP64 oracle (CPU) of the storage forms on a
    136^3 crop of the octant model around the source: every step's p at the
    128x128 compute cells of the first plane below the absorbing shell
    (traces), and the final p wavefield."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    full = VM.layered_octant(seed=0, local=512)
    w = VM.crop(full, (208, 208, 0), 136, receiver_z=36)
    del full
    res = {}
    for steps in steps_list:
        amp = H.ricker_table(w.dt, steps, w.fcut)
        zeros = [np.zeros(w.shape, np.float32)] * 4
        t = time.time()
        f64, tr64 = O.run_traces(zeros, w.coef, w.dx, w.dy, w.dz, w.dt, steps,
                                 w.src_index, amp, w.receivers, double=True)
        log(f"{log_prefix} precision: FP64 oracle {steps} steps "
            f"in {time.time() - t:.0f}s")
        f32, tr32 = O.run_traces(zeros, w.coef, w.dx, w.dy, w.dz, w.dt, steps,
                                 w.src_index, amp, w.receivers)
        log(f"{log_prefix} precision: FP64/FP32 oracle {steps} steps "
            f"in {time.time() - t:.0f}s")

        def rel(a, b):
            b = np.asarray(b, np.float64)
            return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                         np.linalg.norm(b))
        row = {"reference_fp32_cpu": {"trace_rel_l2": rel(tr32, tr64),
                                      "wavefield_rel_l2": rel(f32[1], f64[1])}}
        for name, fs, cs in FORMS:
            d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
            d.set_storage(fs, cs, FIELD_SCALE_LOG2)
            tr, _ = d.run_traces(steps, w.src_index, amp, w.receivers)
            pc, _ = d.get_results()
            flagged = d.storage_overflow()
            d.close()
            row[name] = {"trace_rel_l2": rel(tr, tr64),
                         "wavefield_rel_l2": rel(pc, f64[1]),
                         "left_half_range": float(flagged)}
            log(f"{log_prefix} precision {steps} steps {name}: traces "
                f"{row[name]['trace_rel_l2']:.3e} wavefield "
                f"{row[name]['wavefield_rel_l2']:.3e}")
        res[str(steps)] = {k: {kk: float(f"{vv:.4g}") for kk, vv in v.items()}
                           for k, v in row.items()}
    return res


BATCH_FORMS = [("fp32", 0, 0), ("fp32_n16coef", 0, 4), ("fp16", 1, 1)]
# Half field scale of the survey's shots, which start from rest.  In FP32
# the lowest-frequency shot (10 Hz) reaches 130 field units at its source
# cell, half the 256 that 2^8 leaves, and the half form's rounding grows the
# field further (one shot saturated at 2^8); 2^4 leaves 16x more room.
BATCH_FIELD_SCALE_LOG2 = 4


def batch_study(H, VM, args):
    """The batched form: a survey of independent shots of uneven apertures
    on one shared layered model (velocity_model.layered_survey), every shot
    stepped args.batch_steps times with its own Ricker.

    Compared, for each storage form, on device time of the same steps:
    every shot run alone (dcx_run with a HIP graph, one shot after another)
    with the library's default z chunk (chosen per launch from the shot's
    size) and with 64-plane chunks, and
    all shots in one persistent launch per step (dcx_batch_run, autotuned on
    its own first steps, HIP graph).  The batched shots' fields are bitwise
    those of the shots run alone (tests/test_gpu_batch.py)."""
    t = time.time()
    S = VM.layered_survey(seed=args.seed, n_shots=args.batch_shots,
                          steps=args.batch_steps)
    log(f"[bench] batch: survey of {len(S)} shots, "
        f"{S.compute_cells() / 1e6:.1f} M compute cells per step")
    ds, idx, amps = [], [], []
    for k in range(len(S)):
        w = S.shot(k)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        d.release_host()
        ds.append(d)
        idx.append(w.src_index)
        amps.append(w.source_table())
        if k % 8 == 7 or k == len(S) - 1:
            log(f"[bench] batch: {k + 1} shots on the device "
                f"({time.time() - t:.0f}s)")
    cells = S.compute_cells()
    steps = args.batch_steps
    amp = np.stack(amps)
    out = {"shots": len(S), "steps": steps, "compute_cells_per_step": cells,
           "apertures": "compute extents U[128,320]x U[128,320]x U[128,256], "
                        "16-cell absorbing shell, 10-25 Hz Ricker at a random "
                        "surface position", "seed": args.seed}
    # Every leg runs its steps from rest (dcx_zero_fields), as a survey's
    # shots do: half storage then keeps the scales it takes from zero fields
    # and stays in range over the legs.
    for name, fs, cs in BATCH_FORMS:
        for d in ds:
            d.zero_fields()
            d.set_storage(fs, cs, BATCH_FIELD_SCALE_LOG2)
        row = {}
        out_of_range = 0
        for zc in (0, 64):
            tot = 0.0
            for d, i, a in zip(ds, idx, amps):
                d.zero_fields()
                d.set_tiling(32, 16, zc)
                tot += d.run(steps, i, a, graph=True)
            key = "alone_default" if zc == 0 else f"alone_zchunk{zc}"
            row[key] = round(cells * steps / tot / 1e3, 1)
            out_of_range = max(out_of_range, sum(d.storage_overflow() for d in ds))
            log(f"[bench] batch {name}: shots alone ({key}) {row[key]:.0f} Mcell/s")
        for d in ds:
            d.zero_fields()
        b = H.Batch(ds)
        b.set_autotune(True)
        t0 = time.perf_counter()
        ms = b.run(steps, idx, amp, graph=True)
        wall = time.perf_counter() - t0
        row["batched"] = round(cells * steps / ms / 1e3, 1)
        row["batched_wall"] = round(cells * steps / wall / 1e6, 1)
        row["batched_steady"] = round(cells / b.last_step_ms() / 1e3, 1)
        # Shots whose half storage saturated in any leg (always 0 for FP32
        # fields).
        row["shots_out_of_range"] = max(out_of_range,
                                        sum(d.storage_overflow() for d in ds))
        row["tiling"] = dict(zip(("z_chunk", "prefetch_planes",
                                  "workgroups_per_cu"), b.tiling()))
        row["speedup_vs_alone_default"] = round(
            row["batched"] / row["alone_default"], 3)
        row["speedup_vs_alone_zchunk64"] = round(
            row["batched"] / row["alone_zchunk64"], 3)
        b.set_autotune(False)
        b.close()
        out[name] = row
        log(f"[bench] batch {name}: batched {row['batched']:.0f} Mcell/s "
            f"(steady {row['batched_steady']:.0f}, tiling {row['tiling']}), "
            f"x{row['speedup_vs_alone_default']:.2f} vs shots alone")
    for d in ds:
        d.close()
    return out


def survey_shots(H, VM, args, fs, cs):
    """The batch study's survey on the device in one storage form, fields at
    rest: (survey, shots, source cells, source tables)."""
    S = VM.layered_survey(seed=args.seed, n_shots=args.batch_shots,
                          steps=args.batch_steps)
    ds, idx, amps = [], [], []
    for k in range(len(S)):
        w = S.shot(k)
        d = H.DeviceData(w.coef, None, w.dx, w.dy, w.dz, w.dt)
        d.release_host()
        d.set_storage(fs, cs, BATCH_FIELD_SCALE_LOG2)
        ds.append(d)
        idx.append(w.src_index)
        amps.append(w.source_table())
    return S, ds, idx, np.stack(amps)


def batch_sweep(H, VM, args):
    """Every fixed batch tiling (z chunk, prefetch depth, workgroups per CU;
    0 = one per tile) on the survey, FP32 and the axis form, 40 steps each
    from rest: the space the autotuner searches."""
    out = {}
    for name, fs, cs in (("fp32", 0, 0), ("fp32_n16coef", 0, 4)):
        S, ds, idx, amp = survey_shots(H, VM, args, fs, cs)
        cells = S.compute_cells()
        b = H.Batch(ds)
        b.set_autotune(False)
        rows = []
        for zc in (128, 256, 512):
            for pf in (1, 2, 3):
                if pf == 3 and cs != 0:
                    continue  # the three-plane prefetch is built for FP32 only
                for bpc in (0, 1, 2, 4, 8):
                    for d in ds:
                        d.zero_fields()
                    b.set_tiling(zc, pf, bpc)
                    b.run(40, idx, amp, graph=True)
                    v = round(cells / b.last_step_ms() / 1e3, 1)
                    rows.append({"z_chunk": zc, "prefetch_planes": pf,
                                 "workgroups_per_cu": bpc, "mcell_per_s": v})
                    log(f"[bench] sweep {name} z {zc} prefetch {pf} "
                        f"wg/cu {bpc}: {v:.0f} Mcell/s")
        b.close()
        for d in ds:
            d.close()
        out[name] = sorted(rows, key=lambda r: -r["mcell_per_s"])
    return out


def batch_child(args):
    """Run under rocprofv3: the survey batched in one storage form with one
    fixed tiling, --steps steps (no graph, so each step is a dispatch)."""
    pkg = importlib.import_module("distributed-cube-average_amd")
    H = pkg.dc_hip
    VM = importlib.import_module("distributed-cube-average_amd.velocity_model")
    fs, cs, zc, pf, bpc = args.child_batch
    S, ds, idx, amp = survey_shots(H, VM, args, fs, cs)
    b = H.Batch(ds)
    b.set_autotune(False)
    b.set_tiling(zc, pf, bpc)
    ms = b.run(args.steps, idx, amp, graph=False)
    log(f"[bench] batch child: {ms / args.steps:.3f} ms per step")
    b.close()


def learned_study(H, args):
    """The learned absorbing layer at 384^3 (6-cell layer, shipped weights):
    reflection ratios against a 640^3 padded reference run for a source 40 cells
    below one face over 1000 steps, beside the zero frame, the reference's
    random-velocity shells and the one-way walls (alone and around the
    layer), and the device time per step of the thin domain with and without
    the layer and inside each wall."""
    LB = importlib.import_module("distributed-cube-average_amd.learned_boundary")
    params, meta = LB.load_weights()
    mlp = H.Mlp(params, mode=int(meta["mode"]), sigma_max=float(meta["sigma_max"]))
    case = LB.make_case(7001, inner=384, width=int(meta["width"]), margin=128,
                        src_depth=40)
    r = LB.reflection(mlp, case, 1000, wall=True)
    mlp.close()
    del case
    plain, learned, ncell, wall_ms = LB.cost(384, int(meta["width"]), walls=True)
    out = {"reflection": {k: float(f"{v:.4g}") for k, v in r.items()},
           "ms_per_step_plain": round(plain, 4),
           "ms_per_step_learned": round(learned, 4), "layer_cells": ncell,
           "ms_per_step_wall": {k: round(v, 4) for k, v in wall_ms.items()},
           "width": int(meta["width"]), "reference_run": "640^3 padded domain"}
    log(f"[bench] learned layer: {out}")
    return out


def pmc_child(args):
    """Run under rocprofv3 --pmc: one calibration copy, then a few steps."""
    pkg = importlib.import_module("distributed-cube-average_amd")
    H = pkg.dc_hip
    VM = importlib.import_module("distributed-cube-average_amd.velocity_model")
    w = make_workload(VM, args)
    d = open_device(H, w, args)
    if args.child_form:
        if not d.set_storage(args.child_form[0], args.child_form[1],
                             FIELD_SCALE_LOG2):
            raise SystemExit(f"storage form {args.child_form} refused")
    nbytes = d.copy_field(0, 1)  # known size: nbytes read, nbytes written
    d.synchronize()
    with open(os.environ["DC_BENCH_COPY_BYTES"], "w") as fh:
        fh.write(str(nbytes))
    amp = w.source_table()
    d.run(args.steps, w.src_index, amp[:args.steps])
    d.synchronize()
    d.close()


def read_counters(path_glob):
    """{kernel_name: [per-dispatch values]} from rocprofv3 counter CSVs."""
    import csv
    out = {}
    for fn in glob.glob(path_glob, recursive=True):
        with open(fn) as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name") or row.get("KernelName") or ""
                val = row.get("Counter_Value") or row.get("Value")
                if val is None:
                    continue
                out.setdefault(name, []).append(float(val))
    return out


def measure_traffic(args):
    """HBM bytes per stencil launch from FETCH_SIZE/WRITE_SIZE (KiB)."""
    res = {}
    with tempfile.TemporaryDirectory(prefix="dc_bench_pmc_") as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(td, counter)
            cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc", counter,
                   "-d", out, "-o", "pmc", "--output-format", "csv", "--",
                   sys.executable, os.path.abspath(__file__), "--child-pmc",
                   "--steps", "3", "--local", str(args.local), "--seed",
                   str(args.seed), "--tiling", *map(str, args.tiling),
                   "--xcd-remap", str(args.xcd_remap),
                   "--prefetch", str(args.prefetch)]
            log(f"[bench] traffic pass {counter}")
            nfile = os.path.join(td, counter + ".bytes")
            r = subprocess.run(cmd, capture_output=True, text=True,
                               env=dict(os.environ, DC_BENCH_COPY_BYTES=nfile))
            if r.returncode != 0:
                log(f"[bench] rocprofv3 {counter} failed rc={r.returncode}: "
                    f"{r.stderr[-400:]}")
                return None
            vals = read_counters(os.path.join(out, "**", "*counter_collection.csv"))
            sten = [v for k, vs in vals.items() if "k_fast" in k for v in vs]
            copy = [v for k, vs in vals.items() if "k_box_copy1" in k for v in vs]
            if not sten or not copy:
                log(f"[bench] no {counter} rows for the kernels: {list(vals)}")
                return None
            # The calibration copy is the first box-copy dispatch.
            n_copy_bytes = float(open(nfile).read())
            scale = n_copy_bytes / (copy[0] * 1024.0)
            res[counter] = float(np.mean(sten)) * 1024.0 * scale
            res[counter + "_scale"] = scale
    return res


def copy_bandwidth(gib=1.0, reps=20):
    """HIP runtime device-to-device copy rate (GB/s, read + write) of a
    buffer four times the 256 MB last-level cache, timed with HIP events:
    what a plain streaming copy gets from this GPU's HBM, reported beside the
    published peak."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    ms = C.c_float()
    try:
        for call in (hip.hipMalloc(C.byref(a), C.c_size_t(n)),
                     hip.hipMalloc(C.byref(b), C.c_size_t(n)),
                     hip.hipEventCreate(C.byref(e0)),
                     hip.hipEventCreate(C.byref(e1))):
            if call != 0:
                raise RuntimeError(f"HIP error {call}")
        hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None)  # device to device
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    finally:
        for p in (a, b):
            if p.value:
                hip.hipFree(p)
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
    return 2.0 * n * reps / (ms.value / 1e3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--local", type=int, default=512,
                    help="compute cells per axis of the local block")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tiling", type=int, nargs=3, default=(32, 16, 512))
    ap.add_argument("--xcd-remap", type=int, default=-1,
                    help="tile order over the XCDs (-1: the library's "
                         "default per storage form)")
    ap.add_argument("--prefetch", type=int, default=3, choices=(1, 2, 3))
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form-steps", type=int, default=100)
    ap.add_argument("--no-forms", action="store_true")
    ap.add_argument("--no-calls", action="store_true",
                    help="skip the reference worker's seven-call step "
                         "(reference_call_pattern is then null)")
    ap.add_argument("--batch-shots", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=800)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--learned", action="store_true",
                    help="add the 384^3 learned absorbing layer study")
    ap.add_argument("--precision-study", type=int, nargs="*", default=None,
                    help="step counts of the storage-form precision study")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc runs that measure HBM bytes "
                         "per stencil launch (traffic is then null)")
    ap.add_argument("--batch-sweep", action="store_true",
                    help="only time every fixed batch tiling on the survey "
                         "(FP32 and the axis form) and print them as JSON")
    ap.add_argument("--child-pmc", action="store_true", help=argparse.SUPPRESS)
    # Counter child only: storage form, z chunk, prefetch, workgroups per CU
    # of a batched survey run.
    ap.add_argument("--child-batch", type=int, nargs=5, default=None,
                    help=argparse.SUPPRESS)
    # Counter child only: storage form (field, coefficient) of the stepped
    # rank, for counter runs of the compact kernels.
    ap.add_argument("--child-form", type=int, nargs=2, default=None,
                    help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library steps one rank's "
              f"subdomain per GPU, and the exchange between GPUs is not part "
              f"of it (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.child_pmc:
        pmc_child(args)
        return
    if args.child_batch:
        batch_child(args)
        return
    if args.batch_sweep:
        pkg = importlib.import_module("distributed-cube-average_amd")
        VM = importlib.import_module("distributed-cube-average_amd.velocity_model")
        print(json.dumps({"batch_sweep": batch_sweep(pkg.dc_hip, VM, args)}),
              flush=True)
        return

    traffic = None
    if not args.no_traffic:
        traffic = measure_traffic(args)

    pkg = importlib.import_module("distributed-cube-average_amd")
    H = pkg.dc_hip
    VM = importlib.import_module("distributed-cube-average_amd.velocity_model")
    w = make_workload(VM, args)
    d = open_device(H, w, args)
    amp = w.source_table() if w.steps >= args.warmup + args.steps else \
        H.ricker_table(w.dt, args.warmup + args.steps, w.fcut)

    log(f"[bench] warmup {args.warmup} steps")
    d.run(args.warmup, w.src_index, amp[:args.warmup])
    d.synchronize()
    log(f"[bench] timing {args.steps} steps")
    d.synchronize()
    t0 = time.perf_counter()
    dev_ms = d.run(args.steps, w.src_index, amp[args.warmup:args.warmup + args.steps])
    d.synchronize()
    sec = time.perf_counter() - t0
    cells = w.compute_cells()
    value = cells * args.steps / sec / 1e6
    log(f"[bench] {value:.1f} Mcell/s wall ({sec * 1e3 / args.steps:.3f} ms/step, "
        f"device {dev_ms / args.steps:.3f} ms/step)")

    # Per-launch stencil time (events around each launch, same stream).
    n_detail = min(50, args.steps)
    d.set_timing_detail(True)
    d.run(n_detail, -1, None)
    kern_ms = d.last_kernel_ms()
    d.set_timing_detail(False)
    alg_bytes = BYTES_PER_CELL_FP32 * cells
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    log(f"[bench] stencil launch {kern_ms:.3f} ms -> {achieved:.0f} GB/s "
        f"algorithmic ({achieved / HBM_PEAK_GBS:.1%} of 8 TB/s)")

    # The reference worker's seven dc_propagate calls per step, and one.
    calls = None
    if not args.no_calls:
        calls = reference_call_pattern(d, w.sizes)
        d.set_tiling(*args.tiling)
        log(f"[bench] dc_propagate as the reference's worker calls it: "
            f"{calls['seven_calls_ms']:.3f} ms/step (seven boxes), "
            f"{calls['seven_calls_exchange_ms']:.3f} ms/step (slabs launched "
            f"before the interior) vs {calls['one_call_ms']:.3f} ms/step "
            f"(one box)")

    forms = None
    if not args.no_forms:
        forms = time_forms(H, w, d, args, amp, cells)
    fields = d.fields()
    copy_gbs = copy_bandwidth()
    log(f"[bench] device-to-device copy of 1 GiB: {copy_gbs:.0f} GB/s (read + write)")
    # The step's 64 B per cell streamed pointwise with no stencil, over the
    # same arrays (it overwrites pp and qp, so it runs after the readback).
    probe_ms = d.stream_probe(20)
    stream_gbs = alg_bytes / (probe_ms * 1e-3) / 1e9
    log(f"[bench] stencil-free pass of the same 64 B/cell: {probe_ms:.3f} ms "
        f"-> {stream_gbs:.0f} GB/s; the stencil reaches "
        f"{achieved / stream_gbs:.1%} of it")
    d.close()
    batch = None
    if not args.no_batch:
        batch = batch_study(H, VM, args)
    learned = None
    if args.learned:
        learned = learned_study(H, args)
    precision = None
    if not args.no_forms:
        precision = {"vs_fp32_kernel": precision_vs_fp32(H, VM, w)}
    if args.precision_study:
        precision = precision or {}
        precision["vs_fp64_oracle"] = precision_study(H, VM, args.precision_study)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, fields, args.cpu_steps)

    roof = {"bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else
            round(traffic["FETCH_SIZE"] + traffic["WRITE_SIZE"]),
            "kernel": "k_fast<32,16>" if tuple(args.tiling[:2]) == (32, 16)
            else f"k_fast<{args.tiling[0]},{args.tiling[1]}>",
            "kernel_ms": round(kern_ms, 4),
            "algorithmic_bytes": alg_bytes}
    if traffic is not None:
        roof["traffic_detail"] = {k: round(v, 4) for k, v in traffic.items()}
    if copy_gbs is not None:
        roof["copy_gbs"] = round(copy_gbs, 1)
        roof["frac_of_copy"] = round(achieved / copy_gbs, 4)
    roof["stream_ms"] = round(probe_ms, 4)
    roof["stream_gbs"] = round(stream_gbs, 1)
    roof["frac_of_stream"] = round(achieved / stream_gbs, 4)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED_512, 3),
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"layered_octant_{args.local}",
                   "local_arrays": list(w.sizes),
                   "compute_cells_per_step": cells,
                   "form": "plain FP32, fused source, 7-box halo pack/unpack "
                           "loopback per step",
                   "tiling": list(args.tiling), "xcd_remap": args.xcd_remap,
                   "prefetch_planes": args.prefetch,
                   "dt": w.dt,
                   "device_ms_per_step": round(dev_ms / args.steps, 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_call_pattern": calls,
        "forms": forms,
        "batch": batch,
        "learned_layer": learned,
        "precision": precision,
    }
    # The metric's comparison: the fastest compact form whose traces stay
    # within 1e-3 of the FP32 kernel's (FP32 is ~1e-5 from FP64), on the
    # block and batched, against plain FP32.
    if forms and precision and precision.get("vs_fp32_kernel"):
        pv = precision["vs_fp32_kernel"]
        ok = [k for k in forms if k != "fp32" and k in pv and pv[k]["finite"]
              and pv[k]["trace_rel_l2"] <= 1e-3]
        if ok:
            best = max(ok, key=lambda k: forms[k]["mcell_per_s"])
            mixed = {"form": best, "mcell_per_s": forms[best]["mcell_per_s"],
                     "speedup_vs_fp32": forms[best]["speedup_vs_fp32"],
                     "launch_speedup_vs_fp32": round(
                         forms["fp32"]["kernel_ms"] / forms[best]["kernel_ms"], 3),
                     "trace_rel_l2_vs_fp32_kernel": pv[best]["trace_rel_l2"]}
            if batch and best in batch:
                mixed["batched_mcell_per_s"] = batch[best]["batched"]
                mixed["batched_speedup_vs_fp32_block"] = round(
                    batch[best]["batched"] / forms["fp32"]["mcell_per_s"], 3)
                # The same survey in plain FP32, every shot run alone (the
                # library default tiling, a graph per shot).
                if "fp32" in batch:
                    mixed["batched_speedup_vs_fp32_shots_alone"] = round(
                        batch[best]["batched"] / batch["fp32"]["alone_default"], 3)
                    # The same survey batched in plain FP32.
                    mixed["batched_speedup_vs_fp32_batched"] = round(
                        batch[best]["batched"] / batch["fp32"]["batched"], 3)
            line["mixed_vs_fp32"] = mixed
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
