"""ctypes front end of the oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the repo's C restatement of the reference's
one-rank Fletcher TTI step, see fletcher_oracle.c) and, where it has been
built, oracle/_ref/libdcref.so (the reference's own OpenMP backend compiled
from its sources).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.

Arrays are numpy float32 (float64 for the FP64 oracle), C-contiguous, shaped
(sz, sy, sx) so that x is the fastest index as in include/indexing.h:15-20.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libdcref.so")
REF_BIN_PATH = os.path.join(HERE, "_ref", "dc")

STENCIL = 4
COEF_NAMES = ("ch1dxx", "ch1dyy", "ch1dzz", "ch1dxy", "ch1dyz", "ch1dxz",
              "v2px", "v2pz", "v2sz", "v2pn")

_SZ3 = C.c_size_t * 3
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_propagate_f32.argtypes = [_SZ3, _SZ3, _SZ3, _FP, _FP, _FP, _FP,
                                        _FP * 10, C.c_float, C.c_float,
                                        C.c_float, C.c_float]
        L.orc_propagate_f64.argtypes = [_SZ3, _SZ3, _SZ3, _DP, _DP, _DP, _DP,
                                        _FP * 10, C.c_double, C.c_double,
                                        C.c_double, C.c_double]
        L.orc_run_f32.argtypes = [_SZ3, _FP * 4, _FP * 10, C.c_float, C.c_float,
                                  C.c_float, C.c_float, C.c_uint, C.c_longlong,
                                  _FP]
        L.orc_run_f64.argtypes = [_SZ3, _DP * 4, _FP * 10, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_uint,
                                  C.c_longlong, _FP]
        _LLP = C.POINTER(C.c_longlong)
        L.orc_run_f32_traces.argtypes = L.orc_run_f32.argtypes + [C.c_int, _LLP, _FP]
        L.orc_run_f64_traces.argtypes = L.orc_run_f64.argtypes + [C.c_int, _LLP, _DP]
        L.orc_source_f32.argtypes = [C.c_float, C.c_int, C.c_double]
        L.orc_source_f32.restype = C.c_float
        L.orc_default_model.argtypes = [C.c_size_t] + [_FP] * 6
        L.orc_precomp.argtypes = [C.c_size_t] + [_FP] * 6 + [_FP * 10]
        L.orc_random_boundary.argtypes = [C.c_int] * 14 + [
            _FP, _FP, C.POINTER(C.c_uint)]
        L.orc_extract_box.argtypes = [_FP, _SZ3, _SZ3, _SZ3, _FP]
        L.orc_insert_box.argtypes = [_FP, _SZ3, _SZ3, _SZ3, _FP]
        L.orc_weights.argtypes = [_FP]
        L.orc_set_ftz.argtypes = [C.c_int]
        L.orc_set_ftz.restype = C.c_int
        _lib = L
    return _lib


def fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _sz3(v):
    return _SZ3(*[int(x) for x in v])


def _coef_arr(coef):
    return (_FP * 10)(*[fp(c) for c in coef])


def sizes_of(a: np.ndarray):
    """(sx, sy, sz) of an array shaped (sz, sy, sx)."""
    return (a.shape[2], a.shape[1], a.shape[0])


def set_ftz(on: bool) -> int:
    """FTZ/DAZ on or off in every OpenMP thread (CPU timing modes)."""
    return lib().orc_set_ftz(int(on))


def weights() -> np.ndarray:
    out = np.zeros(19, np.float32)
    lib().orc_weights(fp(out))
    return out


def propagate_f32(start, end, pp, pc, qp, qc, coef, dx, dy, dz, dt):
    """Region update in place on pp/qp; src/openmp_propagate.c:6-32."""
    lib().orc_propagate_f32(_sz3(start), _sz3(end), _sz3(sizes_of(pc)), fp(pp),
                            fp(pc), fp(qp), fp(qc), _coef_arr(coef), dx, dy, dz,
                            dt)


def propagate_f64(start, end, pp, pc, qp, qc, coef, dx, dy, dz, dt):
    lib().orc_propagate_f64(_sz3(start), _sz3(end), _sz3(sizes_of(pc)), dp(pp),
                            dp(pc), dp(qp), dp(qc), _coef_arr(coef), dx, dy, dz,
                            dt)


def source_f32(dt: float, it: int, fcut: float = 40.0) -> float:
    return lib().orc_source_f32(dt, it, fcut)


def source_table(dt: float, n: int, fcut: float = 40.0) -> np.ndarray:
    return np.array([source_f32(dt, i, fcut) for i in range(n)], np.float32)


def run_f32(fields, coef, dx, dy, dz, dt, n_steps, src_index=-1, src_amp=None):
    """One-rank time loop (src/worker.c:466-511 with no neighbours).

    fields = [pp, pc, qp, qc] (copied, not modified).  Returns the final
    [pp, pc, qp, qc] after the last swap; pc/qc are the reference's results.
    """
    f = [np.ascontiguousarray(a, np.float32).copy() for a in fields]
    ptrs = (_FP * 4)(*[fp(a) for a in f])
    amp = None if src_amp is None else np.ascontiguousarray(src_amp, np.float32)
    lib().orc_run_f32(_sz3(sizes_of(f[0])), ptrs, _coef_arr(coef), dx, dy, dz,
                      dt, n_steps, src_index, None if amp is None else fp(amp))
    addr = {a.ctypes.data: a for a in f}
    return [addr[C.cast(ptrs[k], C.c_void_p).value] for k in range(4)]


def run_f64(fields, coef, dx, dy, dz, dt, n_steps, src_index=-1, src_amp=None):
    """FP64 twin of run_f32 (float coefficients and source table, promoted)."""
    f = [np.ascontiguousarray(a, np.float64).copy() for a in fields]
    ptrs = (_DP * 4)(*[dp(a) for a in f])
    amp = None if src_amp is None else np.ascontiguousarray(src_amp, np.float32)
    lib().orc_run_f64(_sz3(sizes_of(f[0])), ptrs, _coef_arr(coef), dx, dy, dz,
                      dt, n_steps, src_index, None if amp is None else fp(amp))
    addr = {a.ctypes.data: a for a in f}
    return [addr[C.cast(ptrs[k], C.c_void_p).value] for k in range(4)]


def run_traces(fields, coef, dx, dy, dz, dt, n_steps, src_index, src_amp,
               receivers, double=False):
    """run_f32/run_f64 that also returns pc at the receivers after each
    step, shaped (n_steps, n_receivers)."""
    dt_ = np.float64 if double else np.float32
    P = _DP if double else _FP
    f = [np.ascontiguousarray(a, dt_).copy() for a in fields]
    ptrs = (P * 4)(*[a.ctypes.data_as(P) for a in f])
    rec = np.ascontiguousarray(receivers, np.int64)
    tr = np.zeros((n_steps, rec.size), dt_)
    amp = np.ascontiguousarray(src_amp, np.float32)
    fn = lib().orc_run_f64_traces if double else lib().orc_run_f32_traces
    fn(_sz3(sizes_of(f[0])), ptrs, _coef_arr(coef), dx, dy, dz, dt, n_steps,
       src_index, fp(amp), rec.size, rec.ctypes.data_as(C.POINTER(C.c_longlong)),
       tr.ctypes.data_as(P))
    addr = {a.ctypes.data: a for a in f}
    return [addr[C.cast(ptrs[k], C.c_void_p).value] for k in range(4)], tr


def default_model(n: int):
    """vpz, vsv, epsilon, delta, phi, theta of src/precomp.c:156-169."""
    arrs = [np.empty(n, np.float32) for _ in range(6)]
    lib().orc_default_model(n, *[fp(a) for a in arrs])
    return arrs


def precomp(vpz, vsv, eps, delta, phi, theta):
    """The 10 coefficient arrays of src/precomp.c:85-105 (same shape as vpz)."""
    ins = [np.ascontiguousarray(a, np.float32) for a in
           (vpz, vsv, eps, delta, phi, theta)]
    outs = [np.empty_like(ins[0]) for _ in range(10)]
    lib().orc_precomp(ins[0].size, *[fp(a) for a in ins],
                      (_FP * 10)(*[fp(o) for o in outs]))
    return outs


def random_boundary(local_sizes, global_sizes, start, problem, bord, absorb,
                    vpz, vsv, seed=0):
    """In place on vpz/vsv; src/boundary.c:7-104.  Returns the final seed."""
    s = C.c_uint(seed)
    args = list(local_sizes) + list(global_sizes) + list(start) + list(problem)
    lib().orc_random_boundary(*[int(a) for a in args], int(bord), int(absorb),
                              fp(vpz), fp(vsv), C.byref(s))
    return s.value


def extract_box(arr, start, end):
    n = int(np.prod([e - s for s, e in zip(start, end)]))
    buf = np.empty(n, np.float32)
    lib().orc_extract_box(fp(buf), _sz3(start), _sz3(end), _sz3(sizes_of(arr)),
                          fp(arr))
    return buf


def insert_box(arr, buf, start, end):
    lib().orc_insert_box(fp(np.ascontiguousarray(buf, np.float32)), _sz3(start),
                         _sz3(end), _sz3(sizes_of(arr)), fp(arr))


# --------------------------------------------------------------- reference --
class _PrecompVars(C.Structure):
    _fields_ = [(n, _FP) for n in COEF_NAMES]


class _DeviceData(C.Structure):
    _fields_ = [("pp", _FP), ("pc", _FP), ("qp", _FP), ("qc", _FP),
                ("vpz", _FP), ("vsv", _FP), ("precomp_vars", _PrecompVars)]


class _Anisotropy(C.Structure):
    _fields_ = [(n, _FP) for n in ("theta", "phi", "vsv", "vpz", "epsilon",
                                   "delta")]


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref() -> C.CDLL:
    """The reference's own OpenMP backend (oracle/_ref/libdcref.so)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB_PATH)
        R.dc_propagate.argtypes = [_SZ3, _SZ3, _SZ3, C.c_int * 3, C.c_int * 3,
                                   C.POINTER(_DeviceData), C.c_float, C.c_float,
                                   C.c_float, C.c_float]
        R.dc_calculate_source.argtypes = [C.c_float, C.c_int]
        R.dc_calculate_source.restype = C.c_float
        R.randomVelocityBoundaryPartition.argtypes = [C.c_int] * 14 + [
            _FP, _FP, C.POINTER(C.c_uint)]
        R.dc_compute_precomp_vars.argtypes = [C.c_int] * 3 + [_Anisotropy]
        R.dc_compute_precomp_vars.restype = _PrecompVars
        R.dc_free_precomp_vars.argtypes = [C.POINTER(_PrecompVars)]
        R.dc_compute_anisotropy_vars.argtypes = [C.c_int] * 3
        R.dc_compute_anisotropy_vars.restype = _Anisotropy
        R.dc_free_anisotropy_vars.argtypes = [C.POINTER(_Anisotropy)]
        R.dc_device_extract_halo_face.argtypes = [C.POINTER(_DeviceData), _FP,
                                                  _SZ3, _SZ3, _SZ3, _FP]
        R.dc_device_insert_halo_face.argtypes = [C.POINTER(_DeviceData), _FP,
                                                 _SZ3, _SZ3, _SZ3, _FP]
        _ref = R
    return _ref


def ref_propagate(start, end, pp, pc, qp, qc, coef, dx, dy, dz, dt):
    """The reference's dc_propagate (src/openmp_propagate.c) on these arrays."""
    d = _DeviceData(fp(pp), fp(pc), fp(qp), fp(qc), None, None,
                    _PrecompVars(*[fp(c) for c in coef]))
    ref().dc_propagate(_sz3(start), _sz3(end), _sz3(sizes_of(pc)),
                       (C.c_int * 3)(0, 0, 0), (C.c_int * 3)(1, 1, 1),
                       C.byref(d), dx, dy, dz, dt)


def ref_precomp(vpz, vsv, eps, delta, phi, theta):
    """src/precomp.c:9-108 through the reference's own code."""
    sz, sy, sx = vpz.shape
    ins = {n: np.ascontiguousarray(a, np.float32) for n, a in
           zip(("vpz", "vsv", "epsilon", "delta", "phi", "theta"),
               (vpz, vsv, eps, delta, phi, theta))}
    an = _Anisotropy(**{k: fp(v) for k, v in ins.items()})
    pv = ref().dc_compute_precomp_vars(sx, sy, sz, an)
    n = vpz.size
    outs = [np.ctypeslib.as_array(getattr(pv, nm), shape=(n,)).copy().reshape(
        vpz.shape) for nm in COEF_NAMES]
    ref().dc_free_precomp_vars(C.byref(pv))
    return outs


def ref_default_model(sx, sy, sz):
    an = ref().dc_compute_anisotropy_vars(sx, sy, sz)
    n = sx * sy * sz
    out = [np.ctypeslib.as_array(getattr(an, nm), shape=(n,)).copy() for nm in
           ("vpz", "vsv", "epsilon", "delta", "phi", "theta")]
    ref().dc_free_anisotropy_vars(C.byref(an))
    return out


def ref_random_boundary(local_sizes, global_sizes, start, problem, bord, absorb,
                        vpz, vsv, seed=0):
    s = C.c_uint(seed)
    args = list(local_sizes) + list(global_sizes) + list(start) + list(problem)
    ref().randomVelocityBoundaryPartition(*[int(a) for a in args], int(bord),
                                          int(absorb), fp(vpz), fp(vsv),
                                          C.byref(s))
    return s.value


def ref_source(dt, it):
    return ref().dc_calculate_source(dt, it)
