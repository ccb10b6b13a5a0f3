"""ctypes declarations of include/rms_amd.h (librms_amd.so).

The library is the only compute path: loading fails loudly when it has not
been built, and nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "librms_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "rms_amd.h")

NC_MAX = 4
NPARAM = 16
PREC = {"f32": 0, "f64": 1}
STATUS = {
    0: "RMS_OK", 1: "RMS_ERR_INVALID_ARG", 2: "RMS_ERR_HIP", 3: "RMS_ERR_OUT_OF_MEMORY",
    4: "RMS_ERR_NO_DEVICE", 5: "RMS_ERR_STATE", 6: "RMS_ERR_UNSUPPORTED",
}

# every symbol include/rms_amd.h declares
EXPORTS = (
    "rms_create", "rms_destroy", "rms_last_error", "rms_version", "rms_add_geometry", "rms_derive",
    "rms_nval", "rms_plan_create", "rms_plan_execute", "rms_plan_destroy", "rms_run", "rms_plan_adjoint",
    "rms_run_noalloc", "rms_plan_adjoint_ex", "rms_geometry_info",
)
# D estimators of the adjoint (rms_adjoint_opts.d_estimator)
D_ESTIMATOR = {"auto": 0, "fd": 1, "boundary": 2, "pathwise": 3}
SCRATCH_BYTES = 256 << 20  # RMS_SCRATCH_BYTES


class RmsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class DeviceOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("reserved", C.c_int32)]


class Geometry(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("label_bytes", C.c_int32),
        ("labels", C.c_void_p),
        ("label_class", C.c_void_p),
        ("nlabel", C.c_int32),
        ("nclass", C.c_int32),
        ("res", C.c_double),
        ("D", C.c_double * NC_MAX),
        ("T2", C.c_double * NC_MAX),
        ("kappa", C.c_double * (NC_MAX * NC_MAX)),
        ("periodic", C.c_int32),
        ("storage", C.c_int32),
    ]


class Job(C.Structure):
    _fields_ = [
        ("geom_id", C.c_int32),
        ("nwalk", C.c_int64),
        ("seed", C.c_uint64),
        ("nbvec", C.c_int32),
        ("bval", C.c_void_p),
        ("bvec", C.c_void_p),
        ("ngrad", C.c_int32),
        ("grad", C.c_void_p),
        ("walker_base", C.c_int64),
    ]


class RunParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("TN", C.c_int64),
        ("timepoints", C.c_int32),
        ("precision", C.c_int32),
        ("replicas", C.c_int32),
        ("chunk_steps", C.c_int64),
        ("use_graph", C.c_int32),
        ("block_threads", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("xcd_split", C.c_int32),
        ("skip_counters", C.c_int32),
    ]


class JobOut(C.Structure):
    _fields_ = [
        ("rec", C.c_void_p),
        ("rec_rep", C.c_void_p),
        ("sig", C.c_void_p),
        ("sig_rep", C.c_void_p),
        ("counters", C.c_uint64 * 5),
        ("fin_g", C.c_void_p),
        ("fin_o", C.c_void_p),
        ("fin_gi", C.c_void_p),
        ("fin_oi", C.c_void_p),
        ("fin_c", C.c_void_p),
        ("fin_cnt", C.c_void_p),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("walk_ms", C.c_double),
        ("reduce_ms", C.c_double),
        ("total_ms", C.c_double),
        ("grid", C.c_int32),
        ("block", C.c_int32),
        ("launches", C.c_int32),
        ("reserved", C.c_int32),
        ("walker_steps", C.c_int64),
    ]


class JobAdj(C.Structure):
    _fields_ = [
        ("cot_rec", C.c_void_p),
        ("cot_sig", C.c_void_p),
        ("grad", C.c_void_p),
        ("grad_rep", C.c_void_p),
    ]


class AdjointOpts(C.Structure):
    _fields_ = [
        ("d_estimator", C.c_int32),
        ("fd_all_walkers", C.c_int32),
        ("fd_rel_step", C.c_double),
    ]


_lib = None


def lib():
    """Load librms_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the HIP engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    ctx_pp = C.POINTER(C.c_void_p)
    L.rms_create.argtypes = [C.POINTER(DeviceOpts), ctx_pp]
    L.rms_destroy.argtypes = [C.c_void_p]
    L.rms_last_error.argtypes = [C.c_void_p]
    L.rms_last_error.restype = C.c_char_p
    L.rms_version.restype = C.c_char_p
    L.rms_add_geometry.argtypes = [C.c_void_p, C.POINTER(Geometry), C.POINTER(C.c_int32)]
    L.rms_geometry_info.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]
    L.rms_derive.argtypes = [C.POINTER(Geometry), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    L.rms_nval.argtypes = [C.c_int32, C.c_int32]
    L.rms_nval.restype = C.c_int32
    L.rms_plan_create.argtypes = [C.c_void_p, C.POINTER(Job), C.c_int32, C.POINTER(RunParams), ctx_pp]
    L.rms_plan_execute.argtypes = [C.c_void_p, C.POINTER(JobOut), C.POINTER(Timing)]
    L.rms_plan_destroy.argtypes = [C.c_void_p]
    L.rms_run.argtypes = [C.c_void_p, C.POINTER(Job), C.c_int32, C.POINTER(RunParams), C.POINTER(JobOut),
                          C.POINTER(Timing)]
    L.rms_plan_adjoint.argtypes = [C.c_void_p, C.POINTER(JobAdj), C.POINTER(Timing)]
    L.rms_plan_adjoint_ex.argtypes = [C.c_void_p, C.POINTER(JobAdj), C.POINTER(AdjointOpts), C.POINTER(Timing)]
    L.rms_run_noalloc.argtypes = [C.c_void_p, C.POINTER(Job), C.POINTER(RunParams), C.POINTER(JobOut),
                                  C.POINTER(Timing)]
    for name in EXPORTS:
        if name not in ("rms_version", "rms_last_error", "rms_nval"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
