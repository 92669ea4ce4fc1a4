"""User-written scatter components compiled at run time.

The reference lets a user add a component by writing a Python class with
an OpenCL kernel, found by name when the instrument file names it
(mcramp.py:13-17 build_kernel, :80-92 importlib fallback; the template in
doc/writing/scatter.rst:51-69).  Here the kernel body is a HIP device
function with the signature of ramp_amd/csrc/ramp_plugin.h, compiled for
the context's GPU with hipRTC when the instrument's plan is built, and run
by the library on every neutron that reaches the component (after its
geometry and terminate step), with the reference's per-neutron contract:

    from ramp_amd.plugins import ScatterPlugin

    class SAttenuator(ScatterPlugin):          # in a module SAttenuator.py
        entry = "attenuate"
        source = '''
        RAMP_SCATTER void attenuate(RampNeutron *n, RampIntersection *is,
                                    uint32_t *iidx, uint32_t comp_idx,
                                    const float *params, uint32_t tid)
        {
            if (*iidx != comp_idx || n->die > 0.0f) return;
            n->p *= params[0];
            *iidx = 0; n->sc = (float)comp_idx; ramp_isect_reset(is);
        }'''
        def __init__(self, factor=1.0, idx=0, ctx=None, **kw):
            super().__init__(params=[factor], idx=idx)

An instrument file naming "SAttenuator" then resolves it through the
module of that name, as the reference resolves its plugins.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from .components import _Scat

_cache = {}


class ScatterPlugin(_Scat):
    """Base class: subclasses set `entry` and `source` (or `source_file`)."""
    type_code = capi.RAMP_S_PLUGIN
    entry = None
    source = None
    source_file = None

    def __init__(self, params=(), idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        p = np.asarray(params, dtype=np.float32).ravel()
        if p.size > capi.RAMP_NP:
            raise ValueError(f"a plugin takes at most {capi.RAMP_NP} parameters")
        self.p[:p.size] = p

    def plugin_source(self):
        if self.source_file:
            with open(self.source_file) as f:
                return f.read()
        if not self.source or not self.entry:
            raise ValueError(f"{type(self).__name__}: set `entry` and `source` (or `source_file`)")
        return self.source


def build_kernel(source, entry, ctx_handle):
    """Compile `entry` of `source` for the device of a libramp context
    (cached per context, source and entry); returns the plugin handle.
    The counterpart of the reference's build_kernel (mcramp.py:13-17)."""
    key = (ctx_handle.value if isinstance(ctx_handle, C.c_void_p) else int(ctx_handle), source, entry)
    if key not in _cache:
        lib = capi.load_library()
        h = C.c_void_p()
        capi.check(lib.ramp_plugin_compile(ctx_handle, source.encode(), entry.encode(), C.byref(h)))
        _cache[key] = h
    return _cache[key]


def release(ctx_handle):
    """Destroy the plugins compiled for a context (before it is destroyed)."""
    v = ctx_handle.value if isinstance(ctx_handle, C.c_void_p) else int(ctx_handle)
    lib = capi.load_library()
    for key in [k for k in _cache if k[0] == v]:
        lib.ramp_plugin_destroy(_cache.pop(key))
