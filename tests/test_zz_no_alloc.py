"""ramp_run_noalloc (the entry point that allocates nothing: queues, region
counts and carry buffers in the library's static scratch array) against
ramp_run, through the C-ABI, on the same plans and seeds.  Deterministic
histograms and run counters must be identical after each of two calls in a
row, the second smaller than the first.  Limits and refusals are checked
too."""
import ctypes as C
import json

import numpy as np
import pytest

from ramp_amd import _capi as capi
from tests.parity_util import instrument_cases

CASES = instrument_cases()


def _long_beamline():
    """A slit that stops most of the beam, then 60 open windows and two
    monitors: enough steps behind the slit for a compaction boundary at
    chunks of 262144 histories, so the segment queues are used."""
    comps = {"source": {"source": True,
                        "moderator_kernel": {"name": "MFlat", "mod_dim": [0.04, 0.04], "target_dim": [0.04, 0.04],
                                             "target_dist": 2.0, "E_min": 2.0, "E_max": 20.0}},
             "slit": {"position": [0.0, 0.0, 1.0], "geom_kernel": {"name": "GPlane", "width": 0.01, "height": 0.01},
                      "scat_kernel": {"name": "SArm"}}}
    for i in range(60):
        comps[f"win{i:02d}"] = {"position": [0.0, 0.0, 1.1 + 0.01 * i],
                                "geom_kernel": {"name": "GPlane", "width": 0.2, "height": 0.2},
                                "scat_kernel": {"name": "SArm"}}
    comps["tof"] = {"position": [0.0, 0.0, 2.0], "geom_kernel": {"name": "GPlane", "width": 0.2, "height": 0.2},
                    "scat_kernel": {"name": "SDetector1D", "binning": [0.0, 10.0, 3000.0], "var": "tof",
                                    "restore_neutron": True}}
    comps["count"] = {"position": [0.0, 0.0, 2.01], "geom_kernel": {"name": "GPlane", "width": 0.2, "height": 0.2},
                      "scat_kernel": {"name": "SCounter"}}
    return {"beam": dict(linear=True, **comps)}


def _run(lib, inst, fn, n, chunk, seeds):
    opts = capi.RunOpts()
    opts.hist_mode = capi.RAMP_HIST_DETERMINISTIC
    opts.weight_bound = inst.weight_bound(chunk if chunk else n) if n else 0.0
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    return fn(inst._plan, n, chunk, sd.ctypes.data_as(C.POINTER(C.c_uint32)), len(sd), C.byref(opts))


def _hists(lib, inst):
    out = []
    for k, c in enumerate(x for x in inst._comp_list if getattr(x.scat, "n_bins", 0)):
        n = int(c.scat.n_bins)
        w = np.zeros(n)
        w2 = np.zeros(n)
        capi.check(lib.ramp_hist_read(inst._plan, k, w.ctypes.data_as(C.POINTER(C.c_double)),
                                      w2.ctypes.data_as(C.POINTER(C.c_double)), n))
        out.append((w, w2))
    return out


def _pair(path, kw, calls):
    import ramp_amd
    lib = capi.load_library()
    a = ramp_amd.Instrument(path, hist_mode="deterministic", **kw)
    b = ramp_amd.Instrument(path, hist_mode="deterministic", **kw)
    a._ensure_plan()
    b._ensure_plan()
    # ramp_run with no histories allocates b's histograms (no run)
    capi.check(_run(lib, b, lib.ramp_run, 0, 0, [0]))
    segs = []
    for n, chunk, seeds in calls:
        print(f"  n={n} chunk={chunk} seeds={seeds}", flush=True)
        capi.check(_run(lib, a, lib.ramp_run, n, chunk, seeds))
        capi.check(_run(lib, b, lib.ramp_run_noalloc, n, chunk, seeds))
        sa, sb = a.stats(), b.stats()
        for f in ("n_events", "n_oob", "n_segments", "state_out", "state_in", "n_det_overflow"):
            assert sa[f] == sb[f], (f, sa[f], sb[f])
        for (wa, wa2), (wb, wb2) in zip(_hists(lib, a), _hists(lib, b)):
            assert np.array_equal(wa, wb) and np.array_equal(wa2, wb2)
        segs.append(sb["n_segments"])
    a.close()
    b.close()
    return segs


@pytest.mark.gpu
def test_noalloc_equals_run_long_beamline(tmp_path):
    p = tmp_path / "long.json"
    p.write_text(json.dumps(_long_beamline()))
    print("long beamline: two chunks, then one smaller chunk", flush=True)
    segs = _pair(str(p), {}, [(524288, 262144, [5, 6]), (100000, 0, [7])])
    # the first call used the queues (each chunk compacted behind the slit)
    # and the carry buffers of two chunks
    assert segs[0] == 4


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["let_like", "minimal", "tas"])
def test_noalloc_equals_run_instruments(case):
    path, kw = CASES[case]
    print(f"case {case}", flush=True)
    # the second call's smaller chunk raises the weights 6.6-fold and the
    # w^2 sums' budget 43-fold, still inside the first call's scale (a chunk
    # of 1e4 would need a fold, which ramp_run_noalloc refuses)
    _pair(path, kw, [(262144, 0, [11]), (40000, 20000, [12, 13])])


@pytest.mark.gpu
def test_noalloc_limits():
    import ramp_amd
    lib = capi.load_library()
    path, kw = CASES["minimal"]
    b = ramp_amd.Instrument(path, hist_mode="deterministic", **kw)
    b._ensure_plan()
    # histograms absent: refused, nothing run
    assert _run(lib, b, lib.ramp_run_noalloc, 1000, 0, [1]) == capi.RAMP_E_STATE
    capi.check(_run(lib, b, lib.ramp_run, 0, 0, [0]))
    # chunk beyond the scratch array's limit
    assert _run(lib, b, lib.ramp_run_noalloc, 262145, 0, [1]) == capi.RAMP_E_RANGE
    # keep_state needs the plan's full-state buffers
    opts = capi.RunOpts()
    opts.hist_mode = capi.RAMP_HIST_DETERMINISTIC
    opts.keep_state = 1
    sd = np.array([1], dtype=np.uint32)
    assert lib.ramp_run_noalloc(b._plan, 1000, 0, sd.ctypes.data_as(C.POINTER(C.c_uint32)), 1,
                                C.byref(opts)) == capi.RAMP_E_INVALID
    # the largest chunk runs
    capi.check(_run(lib, b, lib.ramp_run_noalloc, 262144, 0, [1]))
    b.close()
