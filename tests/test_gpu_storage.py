"""Compact storage forms (dcx_set_storage) against the FP64 oracle.

FP32 fields with 16-bit coefficient codes (all ten, or the six angle terms
with the v^2 terms in FP32) must stay within 1e-4 of FP64 on the 40-step
fixture; half/bfloat16 field storage is a precision trade-off whose error is
printed and bounded loosely here (the study at 1000-4000 steps is in
bench.py --precision-study).
"""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def run40(H, O):
    g = golden("golden_run.npz")
    coef = H.precompute(g["vpz"], g["vsv"], g["epsilon"], g["delta"], g["phi"],
                        g["theta"])
    dt, steps, src, d = float(g["dt"]), int(g["steps"]), int(g["src_index"]), float(g["dx"])
    amp = H.ricker_table(dt, steps)
    f64 = O.run_f64([np.zeros(g["vpz"].shape)] * 4, coef, d, d, d, dt, steps, src, amp)
    return coef, dt, steps, src, d, amp, f64


@pytest.mark.parametrize("form,bpc,tol", [
    ((0, 0), 64, 1e-5), ((0, 1), 44, 1e-3), ((0, 2), 44, 1e-4), ((0, 3), 52, 1e-4),
    ((0, 4), 40, 1e-4),
    ((1, 1), 36, 2e-2), ((1, 2), 36, 2e-2), ((2, 2), 36, 1e-1)])
def test_storage_form_accuracy(H, run40, form, bpc, tol):
    coef, dt, steps, src, d, amp, f64 = run40
    dd = H.DeviceData(coef, None, d, d, d, dt)
    assert dd.set_storage(*form)
    assert dd.bytes_per_cell() == bpc
    dd.run(steps, src, amp, graph=True)
    pc, qc = dd.get_results()
    dd.close()
    e = rel_l2(pc, f64[1]), rel_l2(qc, f64[3])
    say(f"storage {form} ({bpc} B/cell): relL2 vs FP64 p={e[0]:.3e} q={e[1]:.3e}")
    assert max(e) < tol


def test_storage_round_trip(H, run40):
    """F32 -> half fields -> F32 reproduces pc exactly up to half rounding and
    keeps stepping consistently."""
    coef, dt, steps, src, d, amp, f64 = run40
    dd = H.DeviceData(coef, None, d, d, d, dt)
    dd.run(10, src, amp[:10])
    before = dd.fields()
    assert dd.set_storage(1, 2, 8)
    mid = dd.fields()
    for a, b in zip(before, mid):
        assert np.allclose(a, b, rtol=2e-3, atol=2e-3 * np.abs(a).max())
    assert dd.set_storage(0, 0)
    after = dd.fields()
    for a, b in zip(mid, after):
        np.testing.assert_array_equal(a, b)
    dd.close()


@pytest.mark.parametrize("form", [(0, 2), (0, 3), (0, 4), (1, 1), (1, 2), (2, 2)])
def test_storage_forms_tile_shapes_bitwise_equal(H, run40, form):
    """A compact form computes each cell with the same expressions whatever
    the tile: 64x8 tiles (two apron cells per thread) give the bits of 32x16
    tiles over the 40-step fixture loop."""
    coef, dt, steps, src, d, amp, f64 = run40
    out = []
    for tx, ty in ((32, 16), (64, 8)):
        say("storage", form, "tile", tx, ty)
        dd = H.DeviceData(coef, None, d, d, d, dt)
        assert dd.set_tiling(tx, ty, 0)
        assert dd.set_storage(*form)
        dd.run(steps, src, amp, graph=True)
        out.append(dd.fields())
        dd.close()
    for a, b in zip(*out):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                              np.ascontiguousarray(b).view(np.uint32))


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_axis_form_refuses_arrays_of_another_form(H, run40):
    """DCX_COEF_N16 assumes the reference's TTI structure: angle terms
    n_i n_j of a unit axis, v^2 terms proportional to v2pz.  Arrays without
    it (the ch1dxy term replaced by noise) are refused: set_storage returns
    -1, the struct keeps its FP32 coefficients and runs as before, bit for
    bit."""
    coef, dt, steps, src, d, amp, f64 = run40
    bad = [c.copy() for c in coef]
    rng = np.random.default_rng(3)
    bad[3] = rng.uniform(-0.5, 0.5, bad[3].shape).astype(np.float32)
    out = []
    for try_axis in (True, False):
        dd = H.DeviceData(bad, None, d, d, d, dt)
        if try_axis:
            assert not dd.set_storage(0, 4)
            assert dd.bytes_per_cell() == 64
        dd.run(steps, src, amp)
        out.append(dd.get_results())
        dd.close()
    for a, b in zip(*out):
        assert np.array_equal(_bits(a), _bits(b))


def test_constant_coefficient_arrays_code_exactly(H, O):
    """The reference's default model (src/precomp.c:156-169) has the same
    six angle terms in every cell.  Their 16-bit code ranges come from the
    array's own cells (not its row padding), so each decodes to its FP32
    value exactly, and 30 steps with the angle terms coded (DCX_COEF_A16)
    give the bits of the FP32 arrays."""
    n = 24
    shape = (n, n, n)
    vpz, vsv, eps, dlt, phi, th = [a.reshape(shape) for a in O.default_model(n ** 3)]
    coef = O.precomp(vpz, vsv, eps, dlt, phi, th)
    for k in range(6):
        assert np.all(coef[k] == coef[k].flat[0])
    dt = 0.002
    src = 12 + n * (11 + n * 10)
    amp = H.ricker_table(dt, 30)
    out = []
    for form in (0, 3):
        dd = H.DeviceData(coef, None, 20.0, 20.0, 20.0, dt)
        assert dd.set_storage(0, form)
        dd.run(30, src, amp)
        out.append(dd.get_results())
        dd.close()
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(*out):
        assert np.array_equal(_bits(a), _bits(b))


def test_half_storage_saturates_and_reports(H, run40, capfd):
    """A source far beyond half's range at the caller's scale: the stored
    values saturate at +-65504 field units instead of becoming inf or NaN,
    dcx_storage_overflow reports it, and the next synchronizing call prints
    one warning."""
    coef, dt, steps, src, d, amp, f64 = run40
    dd = H.DeviceData(coef, None, d, d, d, dt)
    assert dd.set_storage(1, 1, 8)
    assert not dd.storage_overflow()
    dd.run(steps, src, (1e5 * amp).astype(np.float32))
    pc, qc = dd.get_results()
    flagged = dd.storage_overflow()
    dd.close()
    err = capfd.readouterr().err
    assert flagged
    assert np.isfinite(pc).all() and np.isfinite(qc).all()
    assert np.abs(pc).max() <= 65504.0 / 256.0 * (1 + 1e-6)
    assert err.count("field storage left its range") == 1


def test_half_storage_scale_follows_the_fields(H, run40):
    """Converting fields whose values are far beyond 65504 / 2^8 (the
    caller's scale) picks the scale from the fields: no saturation, and the
    converted fields equal the FP32 ones to half precision."""
    coef, dt, steps, src, d, amp, f64 = run40
    dd = H.DeviceData(coef, None, d, d, d, dt)
    dd.run(steps, src, (1e4 * amp).astype(np.float32))
    before = dd.fields()
    big = np.abs(before[1]).max()
    assert big * 256.0 > 65504.0
    assert dd.set_storage(1, 1, 8)
    after = dd.fields()
    flagged = dd.storage_overflow()
    dd.close()
    assert not flagged
    for k in (1, 3):
        assert np.isfinite(after[k]).all()
        assert np.abs(after[k] - before[k]).max() <= 2e-3 * big


@pytest.mark.parametrize("form", [(0, 0), (0, 4), (1, 1)])
def test_zero_fields_is_a_run_from_rest(H, run40, form):
    """dcx_zero_fields after 20 steps, then 15 steps with the source: the
    bits of a fresh struct of the same form stepped 15 times (the scales a
    half form picked at conversion from zero fields are kept)."""
    coef, dt, steps, src, d, amp, f64 = run40
    out = []
    for warm in (20, 0):
        dd = H.DeviceData(coef, None, d, d, d, dt)
        assert dd.set_storage(*form, 8)
        if warm:
            dd.run(warm, src, amp[:warm])
            dd.zero_fields()
            for f in dd.fields():
                assert not f.any()
        dd.run(15, src, amp[:15], graph=True)
        out.append(dd.fields())
        dd.close()
    assert np.abs(out[1][1]).max() > 0
    for a, b in zip(*out):
        assert np.array_equal(_bits(a), _bits(b))


def test_zero_fields_clears_the_overflow_flag(H, run40, capfd):
    """A half form that saturated (source far beyond its range) is in range
    again after dcx_zero_fields: the flag is clear, and a later saturation is
    reported again."""
    coef, dt, steps, src, d, amp, f64 = run40
    dd = H.DeviceData(coef, None, d, d, d, dt)
    assert dd.set_storage(1, 1, 8)
    big = (1e5 * amp).astype(np.float32)
    dd.run(steps, src, big)
    assert dd.storage_overflow()
    dd.zero_fields()
    assert not dd.storage_overflow()
    dd.run(steps, src, big)
    dd.synchronize()
    flagged = dd.storage_overflow()
    dd.close()
    assert flagged
    assert capfd.readouterr().err.count("field storage left its range") == 2


@pytest.mark.parametrize("form", [(0, 4), (0, 3)])
def test_coefficient_forms_with_unequal_spacings(H, O, VM, form):
    """dx, dy and dz all different (10, 20 and 35 m).  The axis form carries
    the spacings in its decoded axis (sqrt(2) n_x/dx, ...), one factor per
    axis, so a swapped axis or spacing there changes the result at O(1):
    30 fused steps with a source on a random TTI model, within 1e-4 of the
    FP64 oracle."""
    w = VM.random_small((30, 34, 40), seed=19)
    dx, dy, dz = 10.0, 20.0, 35.0
    dt = float(np.float32(0.4 * VM.cfl_dt(w.coef[6], dx, dy, dz)))
    steps = 30
    amp = H.ricker_table(dt, steps)
    d = H.DeviceData(w.coef, None, dx, dy, dz, dt)
    assert d.set_storage(*form)
    d.run(steps, w.src_index, amp, graph=True)
    pc, qc = d.get_results()
    d.close()
    zeros = [np.zeros(w.shape)] * 4
    ref = O.run_f64(zeros, w.coef, dx, dy, dz, dt, steps, w.src_index, amp)
    e = rel_l2(pc, ref[1]), rel_l2(qc, ref[3])
    say(f"form {form}, spacings 10/20/35 m: vs FP64 p={e[0]:.3e} q={e[1]:.3e}")
    assert max(e) < 1e-4
