"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Runs an instrument plan (ramp_amd.Instrument.plan_desc()) through
liboracle.so, the plain-C restatement of the reference kernels.  Imported
only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RAMP_ORACLE_LIB selects another build of the same source (the sanitizer
# build, liboracle_san.so)
LIB = os.environ.get("RAMP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        from ramp_amd import _capi as capi
        lib = C.CDLL(LIB)
        lib.oracle_run.restype = C.c_int
        lib.oracle_run.argtypes = [C.POINTER(capi.PlanDesc), C.c_uint64, C.c_uint64,
                                   C.POINTER(C.c_uint32), C.c_uint32,
                                   C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_uint64)]
        lib.oracle_adjoint.restype = C.c_int
        lib.oracle_adjoint.argtypes = [C.POINTER(capi.PlanDesc), C.c_uint64, C.c_uint64,
                                       C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_uint32,
                                       C.c_int]
        lib.oracle_set_plugin.restype = None
        lib.oracle_set_plugin.argtypes = [C.c_uint32, C.c_void_p]
        lib.oracle_set_rescal_out.restype = None
        lib.oracle_set_rescal_out.argtypes = [C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_uint64)]
        lib.oracle_set_trace.restype = None
        lib.oracle_set_trace.argtypes = [C.POINTER(C.c_float), C.c_uint64]
        lib.oracle_set_mcpl_dump.restype = None
        lib.oracle_set_mcpl_dump.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_uint64]
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_math_d.restype = None
        lib.oracle_math_d.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64]
        lib.oracle_opcount.restype = None
        lib.oracle_opcount.argtypes = [C.POINTER(C.c_uint64), C.c_int]
        lib.oracle_threefry4x32.restype = None
        lib.oracle_threefry4x32.argtypes = [C.POINTER(C.c_uint32)] * 3
        _lib = lib
    return _lib


OPCOUNT_NAMES = ("threefry", "div", "sqrt", "exp", "log", "pow", "tanh", "atan2", "sin", "cos", "asin", "acos",
                 "fmod", "guide_bounce", "guide_exit")


def opcount(reset=True):
    """The operation counts of the ORACLE_COUNT build since the last reset
    (zeros from the other builds), as {name: count}."""
    out = (C.c_uint64 * len(OPCOUNT_NAMES))()
    load().oracle_opcount(out, 1 if reset else 0)
    return dict(zip(OPCOUNT_NAMES, (int(v) for v in out)))


def run(inst, n_hist, seeds, chunk=0, keep_state=False, threads=0, trace=False):
    """Run `inst` (a ramp_amd.Instrument) on the oracle.

    Returns dict: 'hist' list of (w, w2) float64 per detector slot, 'oob',
    with keep_state the final AoS state (n16, isect8, iidx) of the last
    chunk, and 'mcpl': per SMCPLOut component the records (n, 11) float32
    it wrote in the last chunk, in history order.  With trace, 'trace':
    {component name: (xyz (n_hist, 3) float32, alive (n_hist,) bool)} for
    the drawn components of linear blocks."""
    lib = load()
    desc = inst.plan_desc()
    traced = [c.name for b in inst.blocks if b.linear for c in b.components.values() if c.vis]
    tr = None
    if trace and traced and n_hist:
        tr = np.zeros((len(traced), 4, int(n_hist)), np.float32)
        lib.oracle_set_trace(tr.ctypes.data_as(C.POINTER(C.c_float)), int(n_hist))
    n_out = sum(1 for c in inst._comp_list if type(c.scat).__name__ == "SMCPLOut")
    ch0 = int(chunk) if chunk else int(n_hist)
    dump = tags = None
    if n_out:
        dump = np.zeros((n_out, ch0, 11), np.float32)
        tags = np.zeros((n_out, ch0), np.uint32)
        lib.oracle_set_mcpl_dump(dump.ctypes.data_as(C.POINTER(C.c_float)),
                                 tags.ctypes.data_as(C.POINTER(C.c_uint32)), ch0)
    n_rs = sum(1 for c in inst._comp_list if type(c.scat).__name__ == "SRescal")
    if n_rs:
        rs_ev = np.zeros((n_rs, 5, max(int(n_hist), 1)), np.float32)
        rs_cnt = np.zeros(n_rs, np.uint64)
        lib.oracle_set_rescal_out(rs_ev.ctypes.data_as(C.POINTER(C.c_float)), max(int(n_hist), 1),
                                  rs_cnt.ctypes.data_as(C.POINTER(C.c_uint64)))
    dets = [int(desc.dets[i].n_bins) for i in range(desc.n_dets)]
    hw = [np.zeros(n, np.float64) for n in dets]
    hw2 = [np.zeros(n, np.float64) for n in dets]
    nd = max(1, len(dets))
    pw = (C.POINTER(C.c_double) * nd)()
    pw2 = (C.POINTER(C.c_double) * nd)()
    for i in range(len(dets)):
        pw[i] = hw[i].ctypes.data_as(C.POINTER(C.c_double))
        pw2[i] = hw2[i].ctypes.data_as(C.POINTER(C.c_double))
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    ch = int(chunk) if chunk else int(n_hist)
    last = int(n_hist) - (int(np.ceil(n_hist / ch)) - 1) * ch if n_hist else 0
    n16 = i8 = ii = None
    if keep_state:
        n16 = np.zeros((last, 16), np.float32)
        i8 = np.zeros((last, 8), np.float32)
        ii = np.zeros(last, np.uint32)
    oob = C.c_uint64(0)
    rc = lib.oracle_run(C.byref(desc), int(n_hist), int(chunk),
                        seeds.ctypes.data_as(C.POINTER(C.c_uint32)), len(seeds), pw, pw2,
                        n16.ctypes.data_as(C.POINTER(C.c_float)) if keep_state else None,
                        i8.ctypes.data_as(C.POINTER(C.c_float)) if keep_state else None,
                        ii.ctypes.data_as(C.POINTER(C.c_uint32)) if keep_state else None,
                        int(threads), C.byref(oob))
    if tr is not None:
        lib.oracle_set_trace(None, 0)
    if n_out:
        lib.oracle_set_mcpl_dump(None, None, 0)
    if n_rs:
        lib.oracle_set_rescal_out(None, 0, None)
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib.oracle_last_error().decode()}")
    out = {"hist": list(zip(hw, hw2)), "oob": oob.value}
    if keep_state:
        out["state"] = (n16, i8, ii)
    if n_out:
        n_chunks = int(np.ceil(n_hist / ch0)) if n_hist else 0
        out["mcpl"] = [dump[s][tags[s] == n_chunks] for s in range(n_out)]
    if n_rs:
        out["rescal"] = [rs_ev[r][:, :int(rs_cnt[r])] for r in range(n_rs)]
    if tr is not None:
        out["trace"] = {name: (np.ascontiguousarray(tr[k, :3].T), tr[k, 3] != 0.0)
                        for k, name in enumerate(traced)}
    return out


def adjoint(inst, n_hist, seeds, det, dL_dh, n_grad, chunk=0, threads=0):
    """FP64 reference gradient (oracle_adjoint) of a loss of detector slot
    `det` of `inst`, given dL/dh (float32, one per bin)."""
    lib = load()
    desc = inst.plan_desc()
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    dl = np.ascontiguousarray(dL_dh, dtype=np.float32).reshape(-1)
    grad = np.zeros(int(n_grad), np.float64)
    rc = lib.oracle_adjoint(C.byref(desc), int(n_hist), int(chunk),
                            seeds.ctypes.data_as(C.POINTER(C.c_uint32)), len(seeds), int(det),
                            dl.ctypes.data_as(C.POINTER(C.c_float)),
                            grad.ctypes.data_as(C.POINTER(C.c_double)), int(n_grad), int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib.oracle_last_error().decode()}")
    return grad


def build_host_plugin(source, entry, out_dir):
    """Compile a scatter plugin's source for the host (g++, the same
    arithmetic flags as the oracle) and return (library, function address)
    for oracle_set_plugin.  The source sees ramp_plugin.h as on the GPU."""
    import ctypes
    csrc = os.path.join(os.path.dirname(HERE), "ramp_amd", "csrc")
    src = os.path.join(out_dir, f"plugin_{entry}.cpp")
    so = os.path.join(out_dir, f"plugin_{entry}.so")
    with open(src, "w") as f:
        f.write('#include <stdint.h>\n#include "rrng.h"\n#include "ramp_plugin.h"\n')
        f.write(source)
        f.write(f'\nextern "C" void ramp_host_entry(float *n, float *is, uint32_t *ii, uint32_t ci, '
                f'const float *p, uint32_t tid) {{ {entry}((RampNeutron *)n, (RampIntersection *)is, '
                f'ii, ci, p, tid); }}\n')
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I", csrc, src, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    return lib, ctypes.cast(lib.ramp_host_entry, ctypes.c_void_p).value


def set_plugin(comp, fn_addr):
    load().oracle_set_plugin(int(comp), fn_addr)
