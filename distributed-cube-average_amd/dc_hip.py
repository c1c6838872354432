"""ctypes bindings of libdc_hip.so (the MI355X Fletcher TTI backend).

Mirrors the reference's backend ABI (include/propagate.h, include/device_data.h)
with the same names and argument meaning, plus the library's dcx_* API.  The
structs below are field-for-field copies of include/dc_process.h and
include/precomp.h so a dc_process_t built here is what the reference's C
worker would hand to dc_device_data_init.

Arrays are numpy float32 shaped (sz, sy, sx) -- x fastest, as in
include/indexing.h:15-20.  Every call goes to the HIP library; there is no
CPU fallback, and loading fails loudly if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import sys
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdc_hip.so")

STENCIL = 4
NEIGHBOURHOOD = 27
COEF_NAMES = ("ch1dxx", "ch1dyy", "ch1dzz", "ch1dxy", "ch1dyz", "ch1dxz",
              "v2px", "v2pz", "v2sz", "v2pn")
KERNEL_FAST = 0
KERNEL_EXACT = 1
STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
COEF_F32, COEF_F16, COEF_U16, COEF_A16 = 0, 1, 2, 3

_FP = C.POINTER(C.c_float)
_SZ3 = C.c_size_t * 3
_I3 = C.c_int * 3


class PrecompVars(C.Structure):
    """dc_precomp_vars (include/precomp.h:6-17)."""
    _fields_ = [(n, _FP) for n in COEF_NAMES]


class Anisotropy(C.Structure):
    """dc_anisotropy_t (include/precomp.h:19-26)."""
    _fields_ = [(n, _FP) for n in ("theta", "phi", "vsv", "vpz", "epsilon",
                                   "delta")]


class Process(C.Structure):
    """dc_process_t (include/dc_process.h:7-21)."""
    _fields_ = [("rank", C.c_int), ("coordinates", C.c_int * 3),
                ("neighbours", C.c_int * NEIGHBOURHOOD),
                ("topology", C.c_int * 3), ("iterations", C.c_uint),
                ("source_index", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("sizes", C.c_size_t * 3), ("anisotropy_vars", Anisotropy),
                ("precomp_vars", PrecompVars), ("pp", _FP), ("pc", _FP),
                ("qp", _FP), ("qc", _FP), ("hostnames", C.c_char_p),
                ("num_workers", C.c_size_t)]


class DeviceDataStruct(C.Structure):
    """dc_device_data (include/device_data.h:11-15)."""
    _fields_ = [("pp", _FP), ("pc", _FP), ("qp", _FP), ("qc", _FP),
                ("vpz", _FP), ("vsv", _FP), ("precomp_vars", PrecompVars)]


class ProgramArgs(C.Structure):
    """dcx_program_args (include/dc_hip.h)."""
    _fields_ = [("size_x", C.c_size_t), ("size_y", C.c_size_t),
                ("size_z", C.c_size_t), ("dx", C.c_float), ("dy", C.c_float),
                ("dz", C.c_float), ("dt", C.c_float), ("time_max", C.c_float),
                ("absorption", C.c_size_t), ("topology", C.c_int * 3),
                ("kernel", C.c_int)]


_DD = C.POINTER(DeviceDataStruct)

# (name, restype, argtypes) of every symbol the headers in include/ declare.
SIGNATURES = {
    # include/propagate.h
    "dc_propagate": (None, [_SZ3, _SZ3, _SZ3, _I3, _I3, _DD, C.c_float,
                            C.c_float, C.c_float, C.c_float]),
    # include/device_data.h
    "dc_device_data_init": (_DD, [C.POINTER(Process)]),
    "dc_device_data_free": (None, [_DD]),
    "dc_device_data_get_results": (None, [C.POINTER(Process), _DD]),
    "dc_device_swap_arrays": (None, [_DD]),
    "dc_device_add_source": (None, [_DD, C.c_size_t, C.c_float]),
    "dc_device_extract_halo_face": (None, [_DD, _FP, _SZ3, _SZ3, _SZ3, _FP]),
    "dc_device_insert_halo_face": (None, [_DD, _FP, _SZ3, _SZ3, _SZ3, _FP]),
    # include/dc_hip.h
    "dcx_set_kernel": (C.c_int, [_DD, C.c_int]),
    "dcx_set_tiling": (C.c_int, [_DD, C.c_int, C.c_int, C.c_int]),
    "dcx_set_xcd_remap": (C.c_int, [_DD, C.c_int]),
    "dcx_set_prefetch_depth": (C.c_int, [_DD, C.c_int]),
    "dcx_synchronize": (None, [_DD]),
    "dcx_set_storage": (C.c_int, [_DD, C.c_int, C.c_int, C.c_int]),
    "dcx_bytes_per_cell": (C.c_int, [_DD]),
    "dcx_storage_overflow": (C.c_int, [_DD]),
    "dcx_zero_fields": (None, [_DD]),
    "dcx_download": (None, [_DD, C.c_int, _FP]),
    "dcx_upload": (None, [_DD, C.c_int, _FP]),
    "dcx_step": (None, [_DD, C.c_float, C.c_float, C.c_float, C.c_float,
                        C.c_longlong, C.c_float, C.c_int]),
    "dcx_run": (C.c_double, [_DD, C.c_float, C.c_float, C.c_float, C.c_float,
                             C.c_uint, C.c_longlong, _FP, C.c_int]),
    "dcx_run_traces": (C.c_double, [_DD, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_uint, C.c_longlong, _FP,
                                    C.c_int, C.POINTER(C.c_longlong), _FP,
                                    C.c_int]),
    "dcx_set_timing_detail": (None, [_DD, C.c_int]),
    "dcx_last_kernel_ms": (C.c_double, [_DD]),
    "dcx_stream_probe": (C.c_double, [_DD, C.c_int]),
    "dcx_pack_boxes": (None, [_DD, _FP, C.c_int, C.POINTER(C.c_size_t),
                              C.POINTER(C.c_size_t), _FP]),
    "dcx_unpack_boxes": (None, [_DD, _FP, C.c_int, C.POINTER(C.c_size_t),
                                C.POINTER(C.c_size_t), _FP]),
    "dcx_scratch": (_FP, [_DD, C.c_size_t]),
    "dcx_set_halo_loopback": (None, [_DD, C.c_int] + [C.POINTER(C.c_size_t)] * 4),
    "dcx_copy_field": (C.c_size_t, [_DD, C.c_int, C.c_int]),
    "dcx_halo_face_noalloc": (None, [_DD, _FP, _SZ3, _SZ3, _SZ3, _FP, C.c_int]),
    "dcx_batch_create": (C.c_void_p, [C.POINTER(_DD), C.c_int]),
    "dcx_batch_free": (None, [C.c_void_p]),
    "dcx_batch_set_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "dcx_batch_set_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "dcx_batch_get_tiling": (None, [C.c_void_p, _I3]),
    "dcx_batch_last_step_ms": (C.c_double, [C.c_void_p]),
    "dcx_batch_run": (C.c_double, [C.c_void_p, C.c_uint, _FP,
                                   C.POINTER(C.c_longlong), _FP, C.c_int]),
    "dcx_mlp_nparam": (C.c_int, []),
    "dcx_mlp_create": (C.c_void_p, [C.c_int]),
    "dcx_mlp_free": (None, [C.c_void_p]),
    "dcx_mlp_set_output": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    "dcx_mlp_set_params": (None, [C.c_void_p, _FP]),
    "dcx_mlp_get_params": (None, [C.c_void_p, _FP]),
    "dcx_samples_create": (C.c_void_p, [C.c_int, C.c_longlong]),
    "dcx_samples_free": (None, [C.c_void_p]),
    "dcx_samples_count": (C.c_longlong, [C.c_void_p]),
    "dcx_samples_clear": (None, [C.c_void_p]),
    "dcx_samples_upload": (None, [C.c_void_p, C.c_longlong, _FP, _FP]),
    "dcx_samples_download": (None, [C.c_void_p, C.c_longlong, C.c_longlong,
                                    _FP, _FP, _FP, C.POINTER(C.c_uint)]),
    "dcx_mlp_forward": (None, [C.c_void_p, C.c_void_p, _FP]),
    "dcx_mlp_gradient": (C.c_double, [C.c_void_p, C.c_void_p, _FP]),
    "dcx_mlp_train": (C.c_double, [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                   C.c_float, C.c_longlong, _FP]),
    "dcx_set_learned_layer": (C.c_int, [_DD, C.c_void_p, C.c_int]),
    "dcx_set_wall": (C.c_int, [_DD, C.c_int]),
    "dcx_learned_sample_step": (C.c_longlong, [
        _DD, _DD, C.c_longlong * 3, C.c_int, C.c_int, C.c_uint, C.c_void_p,
        C.c_float, C.c_float, C.c_float, C.c_float, C.c_longlong, C.c_float,
        C.c_int, C.c_int]),
    "dcx_program_run": (C.c_int, [C.POINTER(ProgramArgs), _FP, _FP,
                                  C.POINTER(C.c_double)]),
    "dcx_write_dc_file": (C.c_int, [C.c_char_p, C.c_size_t, C.c_size_t,
                                    C.c_size_t, _FP, _FP]),
    "dcx_dims_create": (C.c_int, [C.c_int, _I3]),
    "dcx_default_model": (None, [C.c_size_t] + [_FP] * 6),
    "dcx_precompute": (None, [C.c_size_t] + [_FP] * 6 + [_FP * 10]),
    "dcx_device_data_init_model": (_DD, [C.POINTER(Process), _I3, _I3, _I3,
                                         C.c_int, C.POINTER(C.c_uint)]),
    "dcx_download_coef": (None, [_DD, C.c_int, _FP]),
    "dcx_random_boundary": (None, [C.c_int] * 14 + [_FP, _FP,
                                                    C.POINTER(C.c_uint)]),
    "dcx_ricker": (C.c_float, [C.c_float, C.c_int, C.c_double]),
}

_lib = None


def load() -> C.CDLL:
    """Load libdc_hip.so (build it first with build.py or build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python distributed-cube-average_amd/build.py)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def fp(a: np.ndarray):
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_FP)


def sizes_of(a: np.ndarray):
    return (a.shape[2], a.shape[1], a.shape[0])


# ------------------------------------------------------------ model setup --
def default_model(n: int):
    """vpz, vsv, epsilon, delta, phi, theta of src/precomp.c:156-169."""
    arrs = [np.empty(n, np.float32) for _ in range(6)]
    load().dcx_default_model(n, *[fp(a) for a in arrs])
    return arrs


def precompute(vpz, vsv, epsilon, delta, phi, theta):
    """The 10 TTI coefficient arrays (src/precomp.c:85-105), same shape."""
    ins = [np.ascontiguousarray(a, np.float32) for a in
           (vpz, vsv, epsilon, delta, phi, theta)]
    outs = [np.empty_like(ins[0]) for _ in range(10)]
    load().dcx_precompute(ins[0].size, *[fp(a) for a in ins],
                          (_FP * 10)(*[fp(o) for o in outs]))
    return outs


def random_boundary(local_sizes, global_sizes, start, problem, bord, absorb,
                    vpz, vsv, seed=0):
    """src/boundary.c:7-104 in place on vpz/vsv; returns the final seed."""
    s = C.c_uint(seed)
    args = list(local_sizes) + list(global_sizes) + list(start) + list(problem)
    load().dcx_random_boundary(*[int(v) for v in args], int(bord), int(absorb),
                               fp(vpz), fp(vsv), C.byref(s))
    return s.value


def ricker(dt: float, it: int, fcut: float = 40.0) -> float:
    return load().dcx_ricker(dt, it, fcut)


def ricker_table(dt: float, n: int, fcut: float = 40.0) -> np.ndarray:
    return np.array([ricker(dt, i, fcut) for i in range(n)], np.float32)


# ------------------------------------------------------------ device data --
class DeviceData:
    """One rank's arrays on the GPU, created through dc_device_data_init.

    fields: [pp, pc, qp, qc] host arrays (or None for zeros); coef: the 10
    coefficient arrays in dc_precomp_vars order.  The host arrays are kept
    referenced for the lifetime of the object, as the reference's caller
    keeps its process arrays.
    """

    def __init__(self, coef: Sequence[np.ndarray], fields=None, dx=20.0,
                 dy=20.0, dz=20.0, dt=1e-3, kernel=KERNEL_FAST):
        lib = load()
        self._coef = [np.ascontiguousarray(c, np.float32) for c in coef]
        shape = self._coef[0].shape
        if fields is None:
            fields = [np.zeros(shape, np.float32) for _ in range(4)]
        self._fields = [np.ascontiguousarray(f, np.float32) for f in fields]
        for a in self._coef + self._fields:
            if a.shape != shape:
                raise ValueError("all arrays must have the same shape")
        self.shape = shape
        self.sizes = sizes_of(self._coef[0])
        p = Process()
        p.sizes = _SZ3(*self.sizes)
        p.dx, p.dy, p.dz, p.dt = dx, dy, dz, dt
        p.source_index = -1
        p.topology = _I3(1, 1, 1)
        p.precomp_vars = PrecompVars(*[fp(c) for c in self._coef])
        p.pp, p.pc, p.qp, p.qc = [fp(f) for f in self._fields]
        self._proc = p
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.ptr = lib.dc_device_data_init(C.byref(p))
        self.set_kernel(kernel)

    @classmethod
    def from_default_model(cls, local_sizes, global_sizes, start, problem,
                           absorption, seed=0, dx=20.0, dy=20.0, dz=20.0,
                           dt=1e-3, kernel=KERNEL_FAST):
        """dcx_device_data_init_model: the reference worker's model (default
        TTI model, random-velocity shell, coefficients) built on the device
        for the local box `local_sizes` (x, y, z) at global cell `start`.
        Returns (DeviceData, seed after the shell's draws); fields zero."""
        self = cls.__new__(cls)
        p = Process()
        p.sizes = _SZ3(*local_sizes)
        p.dx, p.dy, p.dz, p.dt = dx, dy, dz, dt
        p.source_index = -1
        p.topology = _I3(1, 1, 1)
        self._proc, self._coef, self._fields = p, None, None
        self.sizes = tuple(int(v) for v in local_sizes)
        self.shape = self.sizes[::-1]
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        s = C.c_uint(seed)
        self.ptr = load().dcx_device_data_init_model(
            C.byref(p), _I3(*global_sizes), _I3(*start), _I3(*problem),
            int(absorption), C.byref(s))
        self.set_kernel(kernel)
        return self, s.value

    def download_coef(self, k: int) -> np.ndarray:
        """Coefficient array k (dc_precomp_vars order) as the device holds it."""
        out = np.empty(self.shape, np.float32)
        load().dcx_download_coef(self.ptr, k, fp(out))
        return out

    # -- reference ABI ------------------------------------------------------
    def propagate(self, start, end, dx=None, dy=None, dz=None, dt=None):
        """dc_propagate over the box [start, end)."""
        load().dc_propagate(_SZ3(*start), _SZ3(*end), _SZ3(*self.sizes),
                            _I3(0, 0, 0), _I3(1, 1, 1), self.ptr,
                            self.dx if dx is None else dx,
                            self.dy if dy is None else dy,
                            self.dz if dz is None else dz,
                            self.dt if dt is None else dt)

    def swap(self):
        load().dc_device_swap_arrays(self.ptr)

    def add_source(self, index: int, amp: float):
        load().dc_device_add_source(self.ptr, index, amp)

    def field_token(self, name: str):
        """The device pointer the caller passes back (data->pp etc.)."""
        return getattr(self.ptr.contents, name)

    def extract_box(self, start, end, field: str) -> np.ndarray:
        n = int(np.prod([e - s for s, e in zip(start, end)]))
        buf = np.empty(n, np.float32)
        load().dc_device_extract_halo_face(self.ptr, fp(buf), _SZ3(*start),
                                           _SZ3(*end), _SZ3(*self.sizes),
                                           self.field_token(field))
        return buf

    def insert_box(self, buf: np.ndarray, start, end, field: str):
        load().dc_device_insert_halo_face(
            self.ptr, fp(np.ascontiguousarray(buf, np.float32)), _SZ3(*start),
            _SZ3(*end), _SZ3(*self.sizes), self.field_token(field))

    def extract_box_noalloc(self, start, end, field: str) -> np.ndarray:
        """extract_box through dcx_halo_face_noalloc (no device allocation)."""
        n = int(np.prod([e - s for s, e in zip(start, end)]))
        buf = np.empty(n, np.float32)
        load().dcx_halo_face_noalloc(self.ptr, fp(buf), _SZ3(*start),
                                     _SZ3(*end), _SZ3(*self.sizes),
                                     self.field_token(field), 0)
        return buf

    def insert_box_noalloc(self, buf: np.ndarray, start, end, field: str):
        """insert_box through dcx_halo_face_noalloc (no device allocation)."""
        load().dcx_halo_face_noalloc(
            self.ptr, fp(np.ascontiguousarray(buf, np.float32)), _SZ3(*start),
            _SZ3(*end), _SZ3(*self.sizes), self.field_token(field), 1)

    def get_results(self):
        """dc_device_data_get_results: returns copies of (pc, qc)."""
        pc = np.empty(self.shape, np.float32)
        qc = np.empty(self.shape, np.float32)
        self._proc.pc, self._proc.qc = fp(pc), fp(qc)
        load().dc_device_data_get_results(C.byref(self._proc), self.ptr)
        return pc, qc

    # -- dcx extensions ------------------------------------------------------
    def set_kernel(self, kernel: int) -> int:
        return load().dcx_set_kernel(self.ptr, kernel)

    def set_tiling(self, tile_x: int, tile_y: int, z_chunk: int) -> bool:
        return load().dcx_set_tiling(self.ptr, tile_x, tile_y, z_chunk) == 0

    def set_prefetch_depth(self, planes: int) -> int:
        return load().dcx_set_prefetch_depth(self.ptr, planes)

    def set_xcd_remap(self, mode) -> int:
        """dcx_set_xcd_remap: 0 off, 1 (or True) contiguous runs, 2 whole
        tile rows per XCD, -1 (the default) 0 for plain FP32 and 2 for the
        compact forms; returns the previous mode."""
        return int(load().dcx_set_xcd_remap(self.ptr, int(mode)))

    def set_storage(self, field_store: int, coef_store: int,
                    field_scale_log2: int = 8) -> bool:
        """Switch to a compact storage form (see include/dc_hip.h)."""
        return load().dcx_set_storage(self.ptr, field_store, coef_store,
                                      field_scale_log2) == 0

    def storage_overflow(self) -> bool:
        """dcx_storage_overflow: whether a 16-bit field store has saturated
        since the last conversion."""
        return bool(load().dcx_storage_overflow(self.ptr))

    def zero_fields(self):
        """dcx_zero_fields: both time levels of p and q back to rest, in the
        current storage form (half storage keeps its scales)."""
        load().dcx_zero_fields(self.ptr)

    def bytes_per_cell(self) -> int:
        return load().dcx_bytes_per_cell(self.ptr)

    def synchronize(self):
        load().dcx_synchronize(self.ptr)

    def download(self, which: int) -> np.ndarray:
        out = np.empty(self.shape, np.float32)
        load().dcx_download(self.ptr, which, fp(out))
        return out

    def upload(self, which: int, arr: np.ndarray):
        load().dcx_upload(self.ptr, which,
                          fp(np.ascontiguousarray(arr, np.float32)))

    def fields(self):
        """Current [pp, pc, qp, qc] copied to host."""
        return [self.download(k) for k in range(4)]

    def step(self, src_index=-1, src_next=0.0, add_next=False):
        load().dcx_step(self.ptr, self.dx, self.dy, self.dz, self.dt,
                        src_index, src_next, int(add_next))

    def run(self, n_steps: int, src_index=-1, src_amp=None, graph=False):
        """n_steps of the one-rank loop; returns elapsed device ms."""
        amp = None
        if src_amp is not None:
            amp = np.ascontiguousarray(src_amp, np.float32)
            if amp.size < n_steps:
                raise ValueError("source table shorter than n_steps")
        return load().dcx_run(self.ptr, self.dx, self.dy, self.dz, self.dt,
                              n_steps, src_index,
                              None if amp is None else fp(amp), int(graph))

    def set_halo_loopback(self, send_boxes, recv_boxes):
        """Boxes as ((x0,y0,z0),(x1,y1,z1)); see dcx_set_halo_loopback."""
        n = len(send_boxes)
        arr = [np.ascontiguousarray(np.array([b[i] for b in boxes], np.uint64)
                                    .reshape(-1)) for boxes in
               (send_boxes, recv_boxes) for i in (0, 1)]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_size_t)) for a in arr]
        load().dcx_set_halo_loopback(self.ptr, n, *ptrs)

    def pack_unpack(self, send_boxes, src: str, recv_boxes, dst: str):
        """dcx_pack_boxes: send_boxes of field src into the struct's device
        scratch buffer (one launch); then dcx_unpack_boxes: that buffer into
        recv_boxes of field dst (one launch).  Boxes as
        ((x0,y0,z0),(x1,y1,z1)); both lists must hold the same number of
        cells."""
        def vol(boxes):
            return sum(int(np.prod([e - s for s, e in zip(*b)])) for b in boxes)

        n = vol(send_boxes)
        if n != vol(recv_boxes):
            raise ValueError(f"send boxes hold {n} cells, receive boxes "
                             f"{vol(recv_boxes)}")
        arr = [np.ascontiguousarray(np.array([b[i] for b in boxes], np.uint64)
                                    .reshape(-1)) for boxes in
               (send_boxes, recv_boxes) for i in (0, 1)]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_size_t)) for a in arr]
        buf = load().dcx_scratch(self.ptr, max(n, 1))
        load().dcx_pack_boxes(self.ptr, buf, len(send_boxes), ptrs[0], ptrs[1],
                              self.field_token(src))
        load().dcx_unpack_boxes(self.ptr, buf, len(recv_boxes), ptrs[2],
                                ptrs[3], self.field_token(dst))

    def stream_probe(self, reps: int = 20) -> float:
        """ms per pointwise pass with the step's bytes and no stencil
        (overwrites pp and qp)."""
        return load().dcx_stream_probe(self.ptr, reps)

    def copy_field(self, dst: int, src: int) -> int:
        """Calibration copy of a whole allocation; returns bytes read."""
        return load().dcx_copy_field(self.ptr, dst, src)

    def run_traces(self, n_steps: int, src_index, src_amp, receivers,
                   graph=False):
        """run() that also returns p at the receivers after every step,
        shaped (n_steps, n_receivers); returns (traces, elapsed ms)."""
        amp = np.ascontiguousarray(src_amp, np.float32)
        rec = np.ascontiguousarray(receivers, np.int64)
        tr = np.zeros((n_steps, rec.size), np.float32)
        ms = load().dcx_run_traces(
            self.ptr, self.dx, self.dy, self.dz, self.dt, n_steps, src_index,
            fp(amp), rec.size, rec.ctypes.data_as(C.POINTER(C.c_longlong)),
            fp(tr), int(graph))
        return tr, ms

    def set_timing_detail(self, on: bool):
        load().dcx_set_timing_detail(self.ptr, int(on))

    def last_kernel_ms(self) -> float:
        return load().dcx_last_kernel_ms(self.ptr)

    def release_host(self):
        """Drop the host copies of the initial arrays (the device holds
        everything after dc_device_data_init)."""
        self._coef = self._fields = None
        self._proc.precomp_vars = PrecompVars()
        self._proc.pp = self._proc.pc = self._proc.qp = self._proc.qc = None

    def close(self):
        if getattr(self, "ptr", None):
            load().dc_device_data_free(self.ptr)   # also detaches its layer
            self.ptr = None
            self._layer = None

    def __del__(self):
        # At interpreter exit objects go in no particular order; the process
        # exit releases device memory, so nothing is freed then.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Shots (DeviceData objects, one storage form) advanced together by one
    persistent launch per step (dcx_batch_run).  The shots stay usable; they
    move back to streams of their own when the batch is closed."""

    def __init__(self, shots: Sequence[DeviceData]):
        self.shots = list(shots)
        arr = (_DD * len(self.shots))(*[s.ptr for s in self.shots])
        self.ptr = load().dcx_batch_create(arr, len(self.shots))

    def set_tiling(self, z_chunk: int, prefetch_planes: int = 1,
                   blocks_per_cu: int = 1) -> bool:
        return load().dcx_batch_set_tiling(self.ptr, z_chunk, prefetch_planes,
                                           blocks_per_cu) == 0

    def set_autotune(self, on: bool = True):
        load().dcx_batch_set_autotune(self.ptr, int(on))

    def tiling(self):
        """(z chunk, prefetch planes, workgroups per CU; 0 = one per tile)."""
        out = _I3()
        load().dcx_batch_get_tiling(self.ptr, out)
        return tuple(out)

    def last_step_ms(self) -> float:
        return load().dcx_batch_last_step_ms(self.ptr)

    def run(self, n_steps: int, src_index=None, src_amp=None, graph=False):
        """src_index: per-shot source cells (-1 for none); src_amp: per-shot
        amplitude tables shaped (n_shots, >= n_steps).  Returns device ms."""
        n = len(self.shots)
        d4 = np.array([[s.dx, s.dy, s.dz, s.dt] for s in self.shots],
                      np.float32).reshape(-1)
        idx = amp = None
        if src_index is not None:
            idx = np.ascontiguousarray(src_index, np.int64)
            if idx.size != n:
                raise ValueError("one source index per shot")
            amp = np.ascontiguousarray(np.asarray(src_amp, np.float32)[:, :n_steps])
            if amp.shape != (n, n_steps):
                raise ValueError("source tables shorter than n_steps")
        return load().dcx_batch_run(
            self.ptr, n_steps, fp(d4),
            None if idx is None else idx.ctypes.data_as(C.POINTER(C.c_longlong)),
            None if amp is None else fp(amp), int(graph))

    def close(self):
        if getattr(self, "ptr", None):
            load().dcx_batch_free(self.ptr)
            self.ptr = None

    def __del__(self):
        # At interpreter exit objects go in no particular order; the process
        # exit releases device memory, so nothing is freed then.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


MLP_FREE, MLP_DAMPING = 0, 1


class Mlp:
    """The learned layer's network on the device (dcx_mlp_*)."""

    def __init__(self, params=None, device=0, mode=MLP_FREE, sigma_max=0.0):
        self.nparam = load().dcx_mlp_nparam()
        self.ptr = load().dcx_mlp_create(device)
        if params is not None:
            self.set_params(params)
        self.set_output(mode, sigma_max)

    def set_output(self, mode: int, sigma_max: float = 0.0):
        if load().dcx_mlp_set_output(self.ptr, mode, sigma_max) != 0:
            raise ValueError(f"bad output form {mode}, sigma_max {sigma_max}")
        self.mode, self.sigma_max = mode, sigma_max

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters")
        load().dcx_mlp_set_params(self.ptr, fp(p))

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        load().dcx_mlp_get_params(self.ptr, fp(out))
        return out

    def forward(self, samples: "Samples") -> np.ndarray:
        y = np.empty((samples.count(), 2), np.float32)
        load().dcx_mlp_forward(self.ptr, samples.ptr, fp(y))
        return y

    def gradient(self, samples: "Samples"):
        g = np.empty(self.nparam, np.float32)
        l = load().dcx_mlp_gradient(self.ptr, samples.ptr, fp(g))
        return g, l

    def train(self, samples: "Samples", n_iter: int, batch: int, lr: float,
              seed: int = 0):
        losses = np.empty(n_iter, np.float32)
        load().dcx_mlp_train(self.ptr, samples.ptr, n_iter, batch, lr, seed,
                             fp(losses))
        return losses

    def close(self):
        if getattr(self, "ptr", None):
            load().dcx_mlp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        # At interpreter exit objects go in no particular order; the process
        # exit releases device memory, so nothing is freed then.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Samples:
    """Training samples on the device (dcx_samples_*)."""

    def __init__(self, capacity: int, device=0):
        self.ptr = load().dcx_samples_create(device, capacity)

    def count(self) -> int:
        return load().dcx_samples_count(self.ptr)

    def clear(self):
        load().dcx_samples_clear(self.ptr)

    def upload(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        load().dcx_samples_upload(self.ptr, x.shape[0], fp(x), fp(t))

    def download(self, first=0, n=None):
        n = self.count() - first if n is None else n
        x = np.empty((n, 64), np.float32)
        t = np.empty((n, 2), np.float32)
        s = np.empty(n, np.float32)
        c = np.empty(n, np.uint32)
        load().dcx_samples_download(self.ptr, first, n, fp(x), fp(t), fp(s),
                                    c.ctypes.data_as(C.POINTER(C.c_uint)))
        return x, t, s, c

    def close(self):
        if getattr(self, "ptr", None):
            load().dcx_samples_free(self.ptr)
            self.ptr = None

    def __del__(self):
        # At interpreter exit objects go in no particular order; the process
        # exit releases device memory, so nothing is freed then.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def set_learned_layer(d: DeviceData, mlp, width: int) -> int:
    """Attach (or, with mlp None, detach) the learned layer; returns the
    number of layer cells.  The struct keeps the network alive while
    attached."""
    n = load().dcx_set_learned_layer(d.ptr, None if mlp is None else mlp.ptr,
                                     width)
    d._layer = mlp if (mlp is not None and width) else None
    return n


WALL_ZERO, WALL_ONE_WAY, WALL_ONE_WAY2 = 0, 1, 2


def set_wall(d: DeviceData, mode: int) -> None:
    """dcx_set_wall: WALL_ZERO keeps the frame at rest (the reference's
    rigid outer frame), WALL_ONE_WAY / WALL_ONE_WAY2 fill it each step from
    the outgoing wave by the first- / second-order one-way condition."""
    load().dcx_set_wall(d.ptr, mode)


def learned_sample_step(padded: DeviceData, small: DeviceData, origin, width,
                        n_cells, seed, samples: Samples, src_index=-1,
                        src_next=0.0, add_next=False, rollout=False,
                        taper=False) -> int:
    """dcx_learned_sample_step with padded's spacing and dt."""
    flags = (1 if rollout else 0) | (2 if taper else 0)
    return load().dcx_learned_sample_step(
        padded.ptr, small.ptr, (C.c_longlong * 3)(*origin), width, n_cells,
        seed & 0xffffffff, samples.ptr, padded.dx, padded.dy, padded.dz,
        padded.dt, src_index, src_next, int(add_next), flags)


def dims_create(n):
    """The rank topology the reference program uses for n ranks
    (src/main.c:89-90, MPI_Dims_create with zeroed dims); host only."""
    d = _I3(0, 0, 0)
    if load().dcx_dims_create(int(n), d) != 0:
        raise ValueError(f"dims_create: n must be >= 1, got {n}")
    return tuple(d)


def program_run(size, absorption, d, dt, time_max, topology=(1, 1, 1),
                kernel=KERNEL_FAST, want_fields=True):
    """The reference program on one GPU; returns (pc, qc, stats).

    pc/qc are global arrays shaped (gz, gy, gx) in the output-file layout.
    """
    a = ProgramArgs()
    a.size_x, a.size_y, a.size_z = size
    a.dx = a.dy = a.dz = d
    a.dt, a.time_max = dt, time_max
    a.absorption = absorption
    a.topology = _I3(*topology)
    a.kernel = kernel
    g = [s + 2 * absorption + 2 * STENCIL for s in size]
    pc = qc = None
    if want_fields:
        pc = np.empty((g[2], g[1], g[0]), np.float32)
        qc = np.empty_like(pc)
    stats = (C.c_double * 3)()
    rc = load().dcx_program_run(C.byref(a), None if pc is None else fp(pc),
                                None if qc is None else fp(qc), stats)
    if rc != 0:
        raise RuntimeError(f"dcx_program_run failed ({rc})")
    return pc, qc, {"seconds": stats[0], "iterations": int(stats[1]),
                    "mcells_per_s": stats[2]}
