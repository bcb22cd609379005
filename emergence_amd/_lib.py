"""ctypes loader of the HIP library (emergence_amd/libemhip.so).

The library is the product: every call below runs gfx950 kernels.  There is
no CPU fallback; a missing or unloadable library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemhip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "em_hip.h")

# every entry point declared in include/em_hip.h
EXPORTS = (
    "em_last_error",
    "em_solver_create",
    "em_solver_create_f32",
    "em_solver_destroy",
    "em_solver_set_variant",
    "em_solver_set_tiles",
    "em_solver_autotune",
    "em_solver_upload",
    "em_solver_download",
    "em_solver_rhs",
    "em_solver_step",
    "em_solver_dt_cfl",
    "em_solver_run_to",
    "em_solver_run_steps",
    "em_solver_counters",
    "em_solver_time",
    "em_solver_synchronize",
    "em_solver_adjoint_rhs",
    "em_solver_set_adjoint_tile",
    "em_solver_forward",
    "em_solver_backward",
    "em_solver_backward_device",
    "em_case_init",
    "em_rhs_noalloc",
    "em_ensemble_create",
    "em_ensemble_destroy",
    "em_ensemble_set_batched",
    "em_ensemble_set_tiles",
    "em_ensemble_autotune",
    "em_ensemble_run_steps",
    "em_ensemble_upload",
    "em_ensemble_download",
    "em_ensemble_time",
    "em_ensemble_forward",
    "em_ensemble_dts",
    "em_ensemble_backward",
    "em_amr_last_error",
    "em_amr_create",
    "em_amr_create_case",
    "em_amr_create_f32",
    "em_amr_create_case_f32",
    "em_amr_destroy",
    "em_amr_set_variant",
    "em_amr_set_batched",
    "em_amr_set_graph",
    "em_amr_set_shapes",
    "em_amr_autotune",
    "em_amr_set_profile",
    "em_amr_profile",
    "em_amr_set_next_regrid",
    "em_amr_nblocks",
    "em_amr_download",
    "em_amr_upload",
    "em_amr_rhs",
    "em_amr_step",
    "em_amr_run_steps",
    "em_amr_run_to",
    "em_amr_dt_cfl",
    "em_amr_details",
    "em_amr_regrid",
    "em_amr_counters",
    "em_amr_time",
    "em_amr_synchronize",
    "em_amr_group_create",
    "em_amr_group_destroy",
    "em_amr_group_run_steps",
    "em_fiveeq_last_error",
    "em_fiveeq_create",
    "em_fiveeq_create_f32",
    "em_fiveeq_destroy",
    "em_fiveeq_upload",
    "em_fiveeq_download",
    "em_fiveeq_rhs",
    "em_fiveeq_step",
    "em_fiveeq_dt_cfl",
    "em_fiveeq_run_to",
    "em_fiveeq_set_tiles",
    "em_fiveeq_autotune",
    "em_fiveeq_counters",
    "em_fiveeq_time",
    "em_fiveeq_synchronize",
)

_lib = None


class EmError(RuntimeError):
    """Failure reported by the HIP library (code, message)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load():
    """Load libemhip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C emergence_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    I64 = C.POINTER(C.c_int64)
    I32 = C.POINTER(C.c_int32)
    U64 = C.POINTER(C.c_uint64)
    lib.em_last_error.restype = C.c_char_p
    lib.em_last_error.argtypes = []
    lib.em_solver_create.argtypes = [C.POINTER(P), I64, D, D, C.c_int32, D, I32, D, C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_double, D]
    lib.em_solver_create_f32.argtypes = [C.POINTER(P), I64, D, D, C.c_int32, D, I32, D, C.c_int32, C.c_int32,
                                         C.c_double, C.c_double, C.c_double, D]
    lib.em_solver_destroy.argtypes = [P]
    lib.em_solver_set_variant.argtypes = [P, C.c_int32]
    lib.em_solver_set_tiles.argtypes = [P, I32]
    lib.em_solver_autotune.argtypes = [P, C.c_int32, I32]
    lib.em_solver_upload.argtypes = [P, D]
    lib.em_solver_download.argtypes = [P, D]
    lib.em_solver_rhs.argtypes = [P, C.c_double, D, U64]
    lib.em_solver_step.argtypes = [P, C.c_double]
    lib.em_solver_dt_cfl.argtypes = [P, D]
    lib.em_solver_run_to.argtypes = [P, C.c_double, I64]
    lib.em_solver_run_steps.argtypes = [P, C.c_int64, D]
    lib.em_solver_counters.argtypes = [P, U64]
    lib.em_solver_time.argtypes = [P, D]
    lib.em_solver_synchronize.argtypes = [P]
    lib.em_solver_adjoint_rhs.argtypes = [P, C.c_double, D, D]
    lib.em_solver_set_adjoint_tile.argtypes = [P, C.c_int32]
    lib.em_solver_forward.argtypes = [P, C.c_int64, C.c_int32, D, D]
    lib.em_solver_backward.argtypes = [P, D, D]
    lib.em_solver_backward_device.argtypes = [P, P, P]
    lib.em_case_init.argtypes = [C.c_int32, I64, D, D, C.c_int32, D, D]
    # device pointers (u, rhs, stats) and the stream travel as plain addresses
    lib.em_rhs_noalloc.argtypes = [I64, D, D, D, I32, D, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double,
                                   P, P, P, P]
    U8 = C.POINTER(C.c_uint8)
    DP = C.POINTER(D)
    lib.em_ensemble_create.argtypes = [C.POINTER(P), C.c_int32, I64, D, D, C.c_int32, D, I32, D, C.c_int32,
                                       C.c_int32, C.c_double, C.c_double, DP]
    lib.em_ensemble_destroy.argtypes = [P]
    lib.em_ensemble_set_batched.argtypes = [P, C.c_int32]
    lib.em_ensemble_set_tiles.argtypes = [P, C.c_int32]
    lib.em_ensemble_autotune.argtypes = [P, C.c_int32, I32]
    lib.em_ensemble_run_steps.argtypes = [P, C.c_int64]
    lib.em_ensemble_upload.argtypes = [P, C.c_int32, D]
    lib.em_ensemble_download.argtypes = [P, C.c_int32, D]
    lib.em_ensemble_time.argtypes = [P, C.c_int32, D]
    lib.em_ensemble_forward.argtypes = [P, C.c_int64, C.c_int32]
    lib.em_ensemble_dts.argtypes = [P, C.c_int32, D]
    lib.em_ensemble_backward.argtypes = [P, DP, DP]
    amr_cfg = [C.POINTER(P), I64, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, D, D, D, I32, D,
               C.c_int32, C.c_int32, C.c_double, C.c_double]
    lib.em_amr_last_error.argtypes = []
    lib.em_amr_create.argtypes = amr_cfg + [C.c_int64, I32, I64, U8, D]
    lib.em_amr_create_case.argtypes = amr_cfg + [C.c_int32, D]
    lib.em_amr_create_f32.argtypes = amr_cfg + [C.c_int64, I32, I64, U8, D]
    lib.em_amr_create_case_f32.argtypes = amr_cfg + [C.c_int32, D]
    lib.em_amr_destroy.argtypes = [P]
    lib.em_amr_set_variant.argtypes = [P, C.c_int32]
    lib.em_amr_set_batched.argtypes = [P, C.c_int32]
    lib.em_amr_set_graph.argtypes = [P, C.c_int32]
    lib.em_amr_set_shapes.argtypes = [P, I32]
    lib.em_amr_autotune.argtypes = [P, C.c_int32, I32]
    lib.em_amr_set_profile.argtypes = [P, C.c_int32]
    lib.em_amr_profile.argtypes = [P, D]
    lib.em_amr_set_next_regrid.argtypes = [P, C.c_int64]
    lib.em_amr_nblocks.argtypes = [P, I64, I64]
    lib.em_amr_download.argtypes = [P, I32, I64, U8, D]
    lib.em_amr_upload.argtypes = [P, D]
    lib.em_amr_rhs.argtypes = [P, C.c_double, D, U64]
    lib.em_amr_step.argtypes = [P, C.c_double]
    lib.em_amr_run_steps.argtypes = [P, C.c_int64]
    lib.em_amr_run_to.argtypes = [P, C.c_double, I64]
    lib.em_amr_dt_cfl.argtypes = [P, D]
    lib.em_amr_details.argtypes = [P, D]
    lib.em_amr_regrid.argtypes = [P]
    lib.em_amr_counters.argtypes = [P, U64]
    lib.em_amr_time.argtypes = [P, D]
    lib.em_amr_synchronize.argtypes = [P]
    lib.em_amr_group_create.argtypes = [C.POINTER(P), C.POINTER(P), C.c_int32]
    lib.em_amr_group_destroy.argtypes = [P]
    lib.em_amr_group_run_steps.argtypes = [P, C.c_int64]
    lib.em_fiveeq_last_error.argtypes = []
    lib.em_fiveeq_create.argtypes = [C.POINTER(P), I64, D, D, C.c_int32, D, C.c_int32, I32, D, C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_double, D]
    lib.em_fiveeq_create_f32.argtypes = [C.POINTER(P), I64, D, D, C.c_int32, D, C.c_int32, I32, D, C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_double, D]
    lib.em_fiveeq_destroy.argtypes = [P]
    lib.em_fiveeq_upload.argtypes = [P, D]
    lib.em_fiveeq_download.argtypes = [P, D]
    lib.em_fiveeq_rhs.argtypes = [P, C.c_double, D, U64]
    lib.em_fiveeq_step.argtypes = [P, C.c_double]
    lib.em_fiveeq_dt_cfl.argtypes = [P, D]
    lib.em_fiveeq_run_to.argtypes = [P, C.c_double, I64]
    lib.em_fiveeq_set_tiles.argtypes = [P, I32]
    lib.em_fiveeq_autotune.argtypes = [P, C.c_int32, I32]
    lib.em_fiveeq_counters.argtypes = [P, U64]
    lib.em_fiveeq_time.argtypes = [P, D]
    lib.em_fiveeq_synchronize.argtypes = [P]
    for name in EXPORTS:
        getattr(lib, name).restype = C.c_char_p if name.endswith("last_error") else C.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise EmError(rc, load().em_last_error().decode())
    return rc


def check_fiveeq(rc):
    if rc != 0:
        raise EmError(rc, load().em_fiveeq_last_error().decode())
    return rc


def check_amr(rc):
    if rc != 0:
        raise EmError(rc, load().em_amr_last_error().decode())
    return rc


class DeviceArray:
    """A float64 buffer in device memory (HIP runtime through ctypes), for
    the entry points that take device arrays; freed with the object."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            load()  # libemhip.so brings the HIP runtime in
            try:
                hip = C.CDLL("libamdhip64.so")
            except OSError:
                hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipFree.argtypes = [C.c_void_p]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = hip
        return cls._hip

    def __init__(self, n):
        self.n = int(n)
        self.ptr = C.c_void_p()
        rc = self._rt().hipMalloc(C.byref(self.ptr), 8 * max(self.n, 1))
        if rc != 0:
            raise EmError(rc, "hipMalloc failed")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        rc = cls._rt().hipMemcpy(d.ptr, a.ctypes.data, 8 * a.size, 1)  # host to device
        if rc != 0:
            raise EmError(rc, "hipMemcpy failed")
        return d

    def to_host(self):
        a = np.empty(self.n, dtype=np.float64)
        rc = self._rt().hipMemcpy(a.ctypes.data, self.ptr, 8 * self.n, 2)  # device to host
        if rc != 0:
            raise EmError(rc, "hipMemcpy failed")
        return a

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and self._hip is not None:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
