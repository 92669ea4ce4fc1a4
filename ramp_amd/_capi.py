"""ctypes mirror of include/ramp.h and the loader of libramp.so.

The library is the MI355X path; there is no CPU fallback.  load_library()
raises RampLibraryError when the built library is missing, so every product
code path fails loudly instead of silently running elsewhere.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libramp.so")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include", "ramp.h")

RAMP_NP = 48
RAMP_NI = 8
RAMP_NT = 4


def _defines():
    out = {}
    with open(INCLUDE) as f:
        for line in f:
            p = line.split()
            if len(p) >= 3 and p[0] == "#define" and p[1].startswith("RAMP_"):
                v = p[2].strip("()")
                try:
                    out[p[1]] = int(v, 0)
                except ValueError:
                    pass
    return out


DEF = _defines()
globals().update(DEF)


class PartDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("ni", C.c_uint32 * RAMP_NI),
        ("p", C.c_float * RAMP_NP),
        ("tab_off", C.c_uint64 * RAMP_NT),
        ("tab_len", C.c_uint64 * RAMP_NT),
    ]


class ComponentDesc(C.Structure):
    _fields_ = [
        ("geom", PartDesc),
        ("scat", PartDesc),
        ("comp_idx", C.c_uint32),
        ("restore", C.c_uint32),
        ("det", C.c_int32),
        ("vis", C.c_uint32),
        ("pos", C.c_float * 3),
        ("rot", C.c_float * 3),
    ]


class BlockDesc(C.Structure):
    _fields_ = [
        ("linear", C.c_uint32),
        ("multi", C.c_uint32),
        ("first", C.c_uint32),
        ("count", C.c_uint32),
    ]


class DetDesc(C.Structure):
    _fields_ = [("n_bins", C.c_uint64)]


class PlanDesc(C.Structure):
    _fields_ = [
        ("source", PartDesc),
        ("source_block", C.c_uint32),
        ("n_blocks", C.c_uint32),
        ("n_comps", C.c_uint32),
        ("n_dets", C.c_uint32),
        ("blocks", C.POINTER(BlockDesc)),
        ("comps", C.POINTER(ComponentDesc)),
        ("dets", C.POINTER(DetDesc)),
        ("tables", C.POINTER(C.c_float)),
        ("n_tables", C.c_uint64),
    ]


class RunOpts(C.Structure):
    _fields_ = [
        ("hist_mode", C.c_uint32),
        ("keep_state", C.c_uint32),
        ("wg", C.c_uint32),
        ("npt", C.c_uint32),
        ("max_segments", C.c_uint32),
        ("flags", C.c_uint32),
        ("pad", C.c_uint32),
        ("weight_bound", C.c_double),
        ("key_base", C.c_uint64),
        ("shard_total", C.c_uint64),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("n_hist", C.c_uint64),
        ("n_chunks", C.c_uint64),
        ("n_events", C.c_uint64),
        ("n_oob", C.c_uint64),
        ("n_segments", C.c_uint64),
        ("state_in", C.c_uint64),
        ("state_out", C.c_uint64),
        ("n_det_overflow", C.c_uint64),
        ("state_out_first", C.c_uint64),
        ("ms_total", C.c_double),
        ("ms_source_segment", C.c_double),
        ("ms_mcpl_compact", C.c_double),
        ("ms_mcpl_write", C.c_double),
        ("ms_rescal_append", C.c_double),
        ("ms_plugin", C.c_double),
        ("plugin_items", C.c_uint64),
        ("ms_mesh_walk", C.c_double),
        ("ms_walk_sort", C.c_double),
        ("walk_rays", C.c_uint64),
    ]


class RampLibraryError(RuntimeError):
    pass


class RampError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libramp error {code}: {msg}")
        self.code = code


_lib = None

_SIGS = {
    "ramp_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "ramp_ctx_destroy": (C.c_int, [C.c_void_p]),
    "ramp_plan_create": (C.c_int, [C.c_void_p, C.POINTER(PlanDesc), C.POINTER(C.c_void_p)]),
    "ramp_plan_destroy": (C.c_int, [C.c_void_p]),
    "ramp_run": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32,
                           C.POINTER(RunOpts)]),
    "ramp_run_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32,
                                   C.POINTER(RunOpts)]),
    "ramp_hist_read": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.c_uint64]),
    "ramp_hist_reset": (C.c_int, [C.c_void_p]),
    "ramp_state_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(C.c_uint32), C.c_uint64]),
    "ramp_stats_read": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "ramp_trace_read": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint64]),
    "ramp_mcpl_read": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint64,
                                 C.POINTER(C.c_uint64)]),
    "ramp_rescal_read": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    "ramp_plugin_compile": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "ramp_plugin_destroy": (C.c_int, [C.c_void_p]),
    "ramp_plan_bind_plugin": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "ramp_math_eval": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.c_uint64, C.c_int]),
    "ramp_rng_eval": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.c_uint64, C.c_int]),
    "ramp_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32,
                                    C.POINTER(C.c_void_p)]),
    "ramp_batch_destroy": (C.c_int, [C.c_void_p]),
    "ramp_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_double), C.POINTER(RunOpts)]),
    "ramp_batch_stats_read": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "ramp_adjoint": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32,
                               C.c_uint32, C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_double),
                               C.c_uint32, C.POINTER(RunOpts)]),
    "ramp_last_error": (C.c_char_p, []),
    "ramp_version": (C.c_char_p, []),
}

EXPORTED = sorted(_SIGS)


def load_library(path=None):
    """Load libramp.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RampLibraryError(
            f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load_library().ramp_last_error()
        raise RampError(rc, msg.decode() if msg else "")
    return rc
