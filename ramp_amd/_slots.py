"""Parameter-slot numbers of the plan descriptors, read from csrc/rslots.h.

The header is the single definition shared by the kernels, the host library
and this package, so a slot can never mean different things on either side
of the C-ABI.
"""
import os
import re

_HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "rslots.h")


def _parse(path):
    plain = {}
    funcs = {}
    pat = re.compile(r"^#define\s+([A-Z0-9_]+)(\(c\))?\s+(.+?)\s*(/\*.*)?$")
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if not m:
                continue
            name, arg, val = m.group(1), m.group(2), m.group(3)
            if name.startswith("RAMP_RSLOTS"):
                continue
            if arg:
                # e.g. IS_RHO(c) (3 * (c) + 0)
                mm = re.match(r"\(\s*(\d+)\s*\*\s*\(c\)\s*\+\s*(\d+)\s*\)", val)
                funcs[name] = (int(mm.group(1)), int(mm.group(2)))
            else:
                plain[name] = int(val)
    return plain, funcs


_PLAIN, _FUNCS = _parse(_HDR)
globals().update(_PLAIN)


def slot(name, c=None):
    """Slot number `name`, or `name(c)` for the per-channel slots."""
    if c is None:
        return _PLAIN[name]
    a, b = _FUNCS[name]
    return a * c + b
