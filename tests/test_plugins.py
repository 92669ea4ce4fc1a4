"""User-written scatter plugins on the CPU: the instrument file resolves
the plugin class through its module, as mcramp.py:80-92 does, and the
oracle runs the plugin's host build (the same source, g++) with the
reference's per-neutron contract.  The plugin (ramp_amd/instruments/
SAttenuatorPlugin.py) multiplies the weight by f + (1-f)u and kills with
probability q: on the neutrons that reach it the kill fraction is q and the
mean factor (1+f)/2 (closed forms)."""
import os
import sys
import tempfile

import numpy as np
import pytest

from ramp_amd import Instrument, capi
from tests.parity_util import ROOT

PLUGINS = os.path.join(ROOT, "ramp_amd", "instruments")
PATH = os.path.join(ROOT, "ramp_amd", "instruments", "plugin.json")
_HOST = {}


def plugin_instrument(where, f3=0.5, q3=0.5, **kw):
    if PLUGINS not in sys.path:
        sys.path.insert(0, PLUGINS)
    return Instrument(PATH, where=where, f3=f3, q3=q3, **kw)


def host_plugin():
    """(module, host function address) of the test plugin (built once)."""
    from oracle import pyoracle
    if PLUGINS not in sys.path:
        sys.path.insert(0, PLUGINS)
    import SAttenuatorPlugin as M
    if "fn" not in _HOST:
        _HOST["lib"], _HOST["fn"] = pyoracle.build_host_plugin(M.SOURCE, M.SAttenuatorPlugin.entry,
                                                               tempfile.mkdtemp(prefix="ramp_plugin_"))
    return M, _HOST["fn"]


def oracle_run(inst, n, seeds, chunk=0, trace=False):
    from oracle import pyoracle
    _, fn = host_plugin()
    inst.plan_desc()
    for k, c in enumerate(inst._comp_list):
        if c.scat.type_code == capi.RAMP_S_PLUGIN:
            pyoracle.set_plugin(k, fn)
    return pyoracle.run(inst, n, seeds, chunk=chunk, keep_state=True, trace=trace)


def test_plugin_resolves_by_module_name():
    inst = plugin_instrument(1)
    comps = {c.name: c for b in inst.blocks for c in b.components.values()}
    p = comps["att2"].scat
    assert type(p).__name__ == "SAttenuatorPlugin" and p.type_code == capi.RAMP_S_PLUGIN
    assert p.p[0] == np.float32(0.25) and p.p[1] == np.float32(0.1)
    assert "RAMP_SCATTER void attenuate" in p.plugin_source()


def test_plugin_closed_form_on_oracle():
    """The same histories with (f, q) = (0.5, 0.5) and (1, 0) at the last
    component: the weight ratio is uniform on [0.5, 1] and half of the
    neutrons that reach it die there."""
    from scipy import stats
    n = 100000
    a = oracle_run(plugin_instrument(2, 0.5, 0.5), n, [9])["state"][0]
    b = oracle_run(plugin_instrument(2, 1.0, 0.0), n, [9])["state"][0]
    reached = b[:, 12] == 4  # prevcomp of att3 (source 1, slit 2, tofmon 3)
    assert reached.sum() > 10000 and np.array_equal(reached, a[:, 12] == 4)
    assert np.all(b[reached, 15] == 0)
    killed = a[reached, 15] == 1
    assert abs(killed.mean() - 0.5) < 4 * np.sqrt(0.25 / reached.sum())
    ratio = a[reached, 9].astype(np.float64) / b[reached, 9]
    assert ratio.min() >= 0.5 - 1e-6 and ratio.max() <= 1 + 1e-6
    assert stats.kstest((ratio - 0.5) / 0.5, "uniform").pvalue > 1e-3
    # the plugin moved every neutron it scattered to its plane: z = 3 m for
    # the survivors; the killed stay in the component's frame (z = 0)
    z = a[reached, 2]
    assert np.allclose(z[~killed], 3.0, atol=1e-5) and np.allclose(z[killed], 0.0, atol=1e-6)


def test_plugin_contract_kill_and_move():
    """The plugin moves survivors to its geometry's exit (z = 1 m for att1)
    and a killed neutron keeps prevcomp of the component before."""
    inst = plugin_instrument(0)
    n16 = oracle_run(inst, 20000, [4])["state"][0]
    hit = n16[:, 12] >= 2
    assert hit.sum() > 1000
    assert np.all(n16[:, 12] >= 0)
