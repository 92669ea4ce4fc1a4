"""Trace runs on the GPU against the oracle, same seeds.

Bar: for every drawn component of a linear block and every history, the
"alive after the component" flag identical to the oracle's, and the trace
point (the intersection's entry point in the component frame, copied
after terminate: mcramp.py:141-161) bit for bit where the neutron lives
on; hence the same trace lines.  Covered: one chunk and several, a
tilted slit, a guide, an undrawn component, a sample and a detector that
lets neutrons through (ramp_amd/instruments/trace.json), and a plugin
step, which writes its trace points from its own kernel."""
import os

import numpy as np
import pytest

from ramp_amd import Instrument, capi
from tests.parity_util import ROOT, compare_states
from tests.test_plugins import oracle_run as plugin_oracle_run, plugin_instrument

PATH = os.path.join(ROOT, "ramp_amd", "instruments", "trace.json")


def _check_points(g, o):
    assert list(g) == list(o)
    for name in o:
        (gx, ga), (ox, oa) = g[name], o[name]
        assert np.array_equal(ga, oa), name
        assert ga.sum() > 0, name
        assert np.array_equal(gx[ga].view(np.uint32), ox[oa].view(np.uint32)), name


def _lines_equal(a, b):
    assert len(a) == len(b)
    assert np.array_equal(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("n,chunk,seeds", [(65536, None, [5]), (100000, 30000, [1, 2, 3, 4])])
def test_trace_parity(n, chunk, seeds):
    from oracle import pyoracle
    print(f"trace N={n} chunk={chunk} seeds={seeds}", flush=True)
    g = Instrument(PATH, tilt=0.05, hist_mode="deterministic", chunk=chunk)
    g.execute(n, seeds=seeds, trace=True)
    gp = g.trace_points()
    glines = g.blocks[0].trace_lines
    st = g.stats()
    g.close()
    o = pyoracle.run(Instrument(PATH, tilt=0.05), n, seeds, chunk=chunk or 0, trace=True)
    print({k: int(v[1].sum()) for k, v in gp.items()}, st["n_segments"], flush=True)
    assert "hidden" not in gp
    _check_points(gp, o["trace"])
    # the host forms the same lines from the oracle's points
    ref = Instrument(PATH, tilt=0.05, chunk=chunk)
    lines = _lines_from_points(ref, o["trace"], n, chunk or n)
    _lines_equal(glines, lines)


def _lines_from_points(inst, pts, n, chunk):
    class _Lib:
        def ramp_trace_read(self, plan, slot, out, m):
            import ctypes as C
            name = list(pts)[slot]
            a = np.ascontiguousarray(np.vstack([pts[name][0].T, pts[name][1][None, :].astype(np.float32)]),
                                     dtype=np.float32)
            C.memmove(out, a.ctypes.data, a.nbytes)
            return 0
    inst._plan = object()
    inst._trace_lines(_Lib(), n, chunk)
    inst._plan = None
    return [l for b in inst.blocks for l in b.trace_lines]


@pytest.mark.gpu
def test_trace_with_keep_state_and_later_untraced_run():
    from oracle import pyoracle
    g = Instrument(PATH, tilt=0.0, hist_mode="deterministic")
    g.execute(50000, seeds=[9], trace=True, keep_state=True)
    gp = g.trace_points()
    gs = g.state(50000)
    o = pyoracle.run(Instrument(PATH, tilt=0.0), 50000, [9], keep_state=True, trace=True)
    _check_points(gp, o["trace"])
    d = compare_states(gs, o["state"])
    assert d["dead"] == 0 and d["iidx"] == 0 and d["ctr"] == 0 and d["beyond_rows"] == 0
    # a run without trace leaves nothing to read
    g.execute(1000, seeds=[1])
    import ctypes as C
    buf = np.zeros(4 * 1000, np.float32)
    rc = capi.load_library().ramp_trace_read(g._plan, 0, buf.ctypes.data_as(C.POINTER(C.c_float)), 1000)
    assert rc == capi.RAMP_E_STATE
    g.close()


@pytest.mark.gpu
def test_trace_plugin_step_parity():
    print("trace plugin", flush=True)
    g = plugin_instrument(1, hist_mode="deterministic", chunk=30000)
    g.execute(70000, seeds=[4, 5, 6], trace=True)
    gp = g.trace_points()
    g.close()
    o = plugin_oracle_run(plugin_instrument(1), 70000, [4, 5, 6], chunk=30000, trace=True)
    assert "att2" in gp
    _check_points(gp, o["trace"])
