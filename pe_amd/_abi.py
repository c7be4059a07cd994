"""ctypes binding of the C-ABI in ``include/pe_hip.h`` (library ``pe_amd/_lib/libpehip.so``).

The HIP library is required: importing this module succeeds without it, but the
first call that needs it raises ``RuntimeError`` (there is no CPU fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libpehip.so")

OK, INVALID_ARG, OUT_OF_MEMORY, CAPACITY_RETRY, HIP_ERROR, NONFINITE, NOT_SUPPORTED = range(7)
SPHERE, BOX, PLANE, CAPSULE, UNION = 1, 2, 3, 4, 5
SWEEP_COLOURED, SWEEP_REFERENCE_ORDER = 0, 1
SCHEDULE_AUTO, SCHEDULE_COLOUR_LAUNCHES, SCHEDULE_DATAFLOW, SCHEDULE_SINGLE_BLOCK, SCHEDULE_ISLANDS = 0, 1, 2, 3, 4

STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "CAPACITY_RETRY", 4: "HIP_ERROR",
                5: "NONFINITE", 6: "NOT_SUPPORTED"}

# Every function declared in include/pe_hip.h.
EXPORTED = [
    "pehip_default_params", "pehip_create", "pehip_destroy", "pehip_set_params", "pehip_get_params",
    "pehip_set_materials", "pehip_upload_bodies", "pehip_upload_members", "pehip_download_members", "pehip_upload_state", "pehip_set_forces",
    "pehip_set_contact_order", "pehip_step", "pehip_step_noalloc", "pehip_detect", "pehip_download_state", "pehip_get_contacts",
    "pehip_get_sweep_order", "pehip_get_impulses", "pehip_get_stats", "pehip_get_step_stats", "pehip_set_profiling", "pehip_eval_trig",
    "pehip_stream", "pehip_last_error", "pehip_version",
    "pehip_batch_create", "pehip_batch_destroy", "pehip_batch_set_params", "pehip_batch_set_materials",
    "pehip_batch_upload", "pehip_batch_set_scene_friction", "pehip_batch_upload_state", "pehip_batch_step",
    "pehip_batch_download_state", "pehip_batch_scene_stats", "pehip_batch_get_contacts",
    "pehip_batch_get_sweep_order", "pehip_batch_get_stats", "pehip_batch_get_tuning", "pehip_batch_set_tuning",
    "pehip_batch_autotune", "pehip_batch_stream", "pehip_batch_last_error",
    "pehip_adjoint_forward", "pehip_adjoint_backward",
]


class Params(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_uint32), ("relaxation_model", ctypes.c_int32),
                ("erp", ctypes.c_double), ("abort_threshold", ctypes.c_double),
                ("gravity", ctypes.c_double * 3), ("damping", ctypes.c_double),
                ("relaxation_param", ctypes.c_double), ("max_sub_iterations", ctypes.c_uint32),
                ("sweep_mode", ctypes.c_int32), ("exact_trig", ctypes.c_int32),
                ("sweep_schedule", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("num_contacts", ctypes.c_uint64), ("max_penetration", ctypes.c_double),
                ("iterations", ctypes.c_uint32), ("num_colours", ctypes.c_uint32),
                ("last_delta", ctypes.c_double), ("kinetic_energy", ctypes.c_double),
                ("fixed_contacts", ctypes.c_uint64), ("sweep_bytes", ctypes.c_double),
                ("t_detection_ms", ctypes.c_double), ("t_colouring_ms", ctypes.c_double),
                ("t_setup_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
                ("t_integrate_ms", ctypes.c_double), ("prof_steps", ctypes.c_uint64),
                ("prof_sweep_launches", ctypes.c_uint64), ("prof_sweep_kernel_ms", ctypes.c_double),
                ("prof_sweep_bytes", ctypes.c_double), ("schedule_used", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("steps_per_launch", ctypes.c_uint32), ("use_graph", ctypes.c_uint32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("nscenes", ctypes.c_uint32), ("num_bodies", ctypes.c_uint64),
                ("contact_capacity", ctypes.c_uint64), ("num_contacts", ctypes.c_uint64),
                ("max_colours", ctypes.c_uint32), ("steps_done_min", ctypes.c_uint64),
                ("capacity_retries", ctypes.c_uint32), ("launches", ctypes.c_uint32),
                ("sweep_bytes", ctypes.c_double)]


class AdjointStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("total_sweeps", ctypes.c_uint64), ("max_contacts", ctypes.c_uint64),
                ("max_last_delta", ctypes.c_double)]


class PeHipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t


def load():
    """Loads the HIP library (raises RuntimeError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("pe_amd HIP library not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    st = ctypes.c_int
    sig = {
        "pehip_default_params": (None, [ctypes.POINTER(Params)]),
        "pehip_create": (st, [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(_VP)]),
        "pehip_destroy": (st, [_VP]),
        "pehip_set_params": (st, [_VP, ctypes.POINTER(Params)]),
        "pehip_get_params": (st, [_VP, ctypes.POINTER(Params)]),
        "pehip_set_materials": (st, [_VP, ctypes.c_uint32, _D, _D]),
        "pehip_upload_bodies": (st, [_VP, ctypes.c_uint32, _U8, _U64, _U32, _U8, _D, _D, _D, _D, _D, _D, _D, _D,
                                     _D, _D]),
        "pehip_upload_state": (st, [_VP, ctypes.c_uint32, _D, _D, _D, _D, _D, _D]),
        "pehip_upload_members": (st, [_VP, ctypes.c_uint32, _U32, _U64, _U32, _D, _D, _D]),
        "pehip_download_members": (st, [_VP, ctypes.c_uint32, _D]),
        "pehip_set_forces": (st, [_VP, ctypes.c_uint32, _D, _D]),
        "pehip_set_contact_order": (st, [_VP, ctypes.c_uint32, _U32, _U32]),
        "pehip_step": (st, [_VP, ctypes.c_double, ctypes.c_uint32]),
        "pehip_step_noalloc": (st, [_VP, ctypes.c_double, ctypes.c_uint32]),
        "pehip_detect": (st, [_VP]),
        "pehip_download_state": (st, [_VP, ctypes.c_uint32, _D, _D, _D, _D, _D, _D]),
        "pehip_get_contacts": (st, [_VP, _SZ, ctypes.POINTER(_SZ), _U32, _U32, _D, _D, _D, _D]),
        "pehip_get_impulses": (st, [_VP, _SZ, ctypes.POINTER(_SZ), _D]),
        "pehip_get_sweep_order": (st, [_VP, _SZ, ctypes.POINTER(_SZ), _U32, _U32, _U32]),
        "pehip_get_stats": (st, [_VP, ctypes.POINTER(Stats)]),
        "pehip_get_step_stats": (st, [_VP, ctypes.POINTER(Stats)]),
        "pehip_set_profiling": (st, [_VP, ctypes.c_int]),
        "pehip_eval_trig": (st, [_VP, ctypes.c_int, ctypes.c_int, _SZ, _D, _D, _D]),
        "pehip_stream": (_VP, [_VP]),
        "pehip_last_error": (ctypes.c_char_p, [_VP]),
        "pehip_version": (ctypes.c_char_p, []),
        "pehip_batch_create": (st, [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(_VP)]),
        "pehip_batch_destroy": (st, [_VP]),
        "pehip_batch_set_params": (st, [_VP, ctypes.POINTER(Params)]),
        "pehip_batch_set_materials": (st, [_VP, ctypes.c_uint32, _D, _D]),
        "pehip_batch_upload": (st, [_VP, ctypes.c_uint32, _U32, _U8, _U64, _U32, _U8, _D, _D, _D, _D, _D, _D, _D,
                                    _D, _D, _D, ctypes.c_double]),
        "pehip_batch_set_scene_friction": (st, [_VP, ctypes.c_uint32, _D, _D]),
        "pehip_batch_upload_state": (st, [_VP, ctypes.c_uint64, _D, _D, _D, _D, _D, _D]),
        "pehip_batch_step": (st, [_VP, ctypes.c_double, ctypes.c_uint32]),
        "pehip_batch_download_state": (st, [_VP, ctypes.c_uint64, _D, _D, _D, _D, _D, _D]),
        "pehip_batch_scene_stats": (st, [_VP, ctypes.c_uint32, _U32, _U32, _U32, _D, _U32]),
        "pehip_batch_get_contacts": (st, [_VP, ctypes.c_uint32, _SZ, ctypes.POINTER(_SZ), _U32, _U32, _D, _D, _D,
                                          _D]),
        "pehip_batch_get_sweep_order": (st, [_VP, ctypes.c_uint32, _SZ, ctypes.POINTER(_SZ), _U32, _U32, _U32]),
        "pehip_batch_get_stats": (st, [_VP, ctypes.POINTER(BatchStats)]),
        "pehip_batch_get_tuning": (st, [_VP, ctypes.POINTER(BatchTuning)]),
        "pehip_batch_set_tuning": (st, [_VP, ctypes.POINTER(BatchTuning)]),
        "pehip_batch_autotune": (st, [_VP, ctypes.c_double, ctypes.c_uint32, ctypes.POINTER(BatchTuning)]),
        "pehip_batch_stream": (_VP, [_VP]),
        "pehip_batch_last_error": (ctypes.c_char_p, [_VP]),
        "pehip_adjoint_forward": (st, [_VP, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32]),
        "pehip_adjoint_backward": (st, [_VP, _D, _D, _D, _D, _D, _D, _D, ctypes.c_uint32, ctypes.c_double,
                                        ctypes.POINTER(AdjointStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status, ctx=None):
    if status != OK:
        msg = load().pehip_last_error(ctx).decode()
        raise PeHipError(status, msg)


def check_batch(status, b=None):
    if status != OK:
        msg = load().pehip_batch_last_error(b).decode()
        raise PeHipError(status, msg)


def ptr(a, typ=_D):
    return a.ctypes.data_as(typ) if a is not None else None
