"""User-written scatter plugins on the GPU: the plugin source compiled at
run time with hipRTC runs as a segment of its own, and equals its host
build on the oracle - states and histograms bit for bit - with the plugin
first, in the middle and last, over one and several chunks.  A source
that does not compile reports the compiler's log."""
import numpy as np
import pytest

from ramp_amd import capi
from tests.parity_util import compare_states
from tests.test_plugins import oracle_run, plugin_instrument


@pytest.mark.gpu
@pytest.mark.parametrize("where,n,chunk,seeds", [(0, 65536, None, [5]), (1, 65536, None, [6]),
                                                 (2, 65536, None, [7]), (1, 90000, 30000, [1, 2, 3])])
def test_plugin_parity(where, n, chunk, seeds):
    print(f"plugin where={where} N={n} chunk={chunk}", flush=True)
    g = plugin_instrument(where, hist_mode="deterministic", chunk=chunk)
    g.execute(n, keep_state=True, seeds=seeds)
    last = n - (len(seeds) - 1) * g.last_chunk
    gs = g.state(last)
    gh = g.hist64("tofmon")
    st = g.stats()
    g.close()
    o = oracle_run(plugin_instrument(where), n, seeds, chunk=chunk or 0)
    d = compare_states(gs, o["state"])
    print(where, d, st["n_segments"], flush=True)
    assert d["bitdiff_rows"] == 0 and d["dead"] == 0 and d["iidx"] == 0
    ow, ow2 = o["hist"][0]
    assert np.abs(gh[0] - ow).max() <= 1e-6 * np.abs(ow).sum()
    assert np.abs(gh[1] - ow2).max() <= 1e-6 * np.abs(ow2).sum()


@pytest.mark.gpu
def test_plugin_compile_error_reports_log():
    import ctypes as C
    lib = capi.load_library()
    ctx = C.c_void_p()
    capi.check(lib.ramp_ctx_create(0, C.byref(ctx)))
    h = C.c_void_p()
    rc = lib.ramp_plugin_compile(ctx, b"RAMP_SCATTER void broken(int x) { return x; }", b"broken", C.byref(h))
    msg = lib.ramp_last_error().decode()
    assert rc == capi.RAMP_E_INVALID and "plugin compilation failed" in msg and "error" in msg
    rc = lib.ramp_plugin_compile(ctx, b"", b"not an id", C.byref(h))
    assert rc == capi.RAMP_E_INVALID
    lib.ramp_ctx_destroy(ctx)
