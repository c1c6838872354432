"""The oracle against the reference's own outputs (tests/golden, written by
tests/golden/make_golden.py from the reference's compiled sources).

Every comparison is bitwise: the oracle restates the reference's FP32
arithmetic in the same order and operand types.
"""
import numpy as np
import pytest

from conftest import golden

import program_oracle as PO


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_precomp_matches_reference(O):
    g = golden("golden_precomp.npz")
    out = O.precomp(g["vpz"], g["vsv"], g["epsilon"], g["delta"], g["phi"],
                    g["theta"])
    for name, a in zip(O.COEF_NAMES, out):
        assert np.array_equal(bits(a), bits(g[name])), name


def test_source_matches_reference(O):
    g = golden("golden_source.npz")
    for dt, row in zip(g["dt"], g["amp"]):
        ours = np.array([O.source_f32(float(dt), i) for i in range(row.size)],
                        np.float32)
        assert np.array_equal(bits(ours), bits(row))


def test_random_boundary_matches_reference_on_every_rank(O):
    g = golden("golden_boundary.npz")
    gl = tuple(int(v) for v in g["global_sizes"])
    prob = tuple(int(v) for v in g["problem"])
    for key in ("000", "100", "010", "110", "001", "101", "011", "111"):
        ref_vpz = g[f"vpz_{key}"]
        loc = (ref_vpz.shape[2], ref_vpz.shape[1], ref_vpz.shape[0])
        vpz = np.full(ref_vpz.shape, 3000.0, np.float32)
        vsv = np.full(ref_vpz.shape, 1296.1481, np.float32)
        seed = O.random_boundary(loc, gl, tuple(g[f"start_{key}"]), prob, 4,
                                 int(g["absorb"]), vpz, vsv, 0)
        assert np.array_equal(bits(vpz), bits(ref_vpz)), key
        assert np.array_equal(bits(vsv), bits(g[f"vsv_{key}"])), key
        assert seed == int(g[f"seed_{key}"])


def _coef(O, g):
    return O.precomp(g["vpz"], g["vsv"], g["epsilon"], g["delta"], g["phi"],
                     g["theta"])


def test_one_propagate_call_matches_reference(O):
    g = golden("golden_cell.npz")
    coef = _coef(O, g)
    f = [g[k].copy() for k in ("pp0", "pc0", "qp0", "qc0")]
    sz, sy, sx = f[0].shape
    O.propagate_f32((4, 4, 4), (sx - 4, sy - 4, sz - 4), f[0], f[1], f[2], f[3],
                    coef, float(g["dx"]), float(g["dy"]), float(g["dz"]),
                    float(g["dt"]))
    assert np.array_equal(bits(f[0]), bits(g["pp1"]))
    assert np.array_equal(bits(f[2]), bits(g["qp1"]))
    # Cells outside the box are untouched.
    assert np.array_equal(bits(f[1]), bits(g["pc0"]))


def test_forty_step_loop_matches_reference(O):
    g = golden("golden_run.npz")
    coef = _coef(O, g)
    shape = g["vpz"].shape
    dt = float(g["dt"])
    steps = int(g["steps"])
    amp = O.source_table(dt, steps)
    f = [np.zeros(shape, np.float32) for _ in range(4)]
    d = float(g["dx"])
    out = O.run_f32(f, coef, d, d, d, dt, steps, int(g["src_index"]), amp)
    assert np.array_equal(bits(out[1]), bits(g["pc"]))
    assert np.array_equal(bits(out[3]), bits(g["qc"]))
    assert np.abs(out[1]).max() > 0  # the wave reached the field


@pytest.mark.parametrize("case,ranks,topo", [
    ("a", 1, (1, 1, 1)), ("a", 8, (2, 2, 2)),
    ("b", 1, (1, 1, 1)), ("b", 8, (2, 2, 2))])
def test_reference_program_output(O, case, ranks, topo):
    g = golden(f"golden_dc_{case}_np{ranks}.npz")
    pc, qc = PO.run_program(tuple(int(v) for v in g["size"]),
                            int(g["absorption"]), float(g["d"]), str(g["dt"]),
                            str(g["tmax"]), topo)
    assert np.array_equal(bits(pc), bits(g["pc"]))
    assert np.array_equal(bits(qc), bits(g["qc"]))


def test_reference_decomposition_invariance():
    """The reference's 1-rank and 8-rank outputs are the same bits."""
    for case in ("a", "b"):
        a = golden(f"golden_dc_{case}_np1.npz")
        b = golden(f"golden_dc_{case}_np8.npz")
        assert np.array_equal(bits(a["pc"]), bits(b["pc"]))
        assert np.array_equal(bits(a["qc"]), bits(b["qc"]))


def test_float_ceil_iteration_count():
    """0.075f/0.0025f rounds above 30: the reference runs 31 steps."""
    assert PO.iterations("0.075", "0.0025") == 31
    assert PO.iterations("0.08", "0.002") == 40
