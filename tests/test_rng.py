"""The per-neutron random stream against the reference's own generator.

Golden vectors come from the reference's vendored Random123 headers
(threefry4x32, 20 rounds; u01fixedpt_closed_closed_32_24), compiled by
oracle/Makefile `ref` (tests/golden/make_golden.py).  A mistake in a
rotation constant, a key-injection index or the u01 map changes every
vector."""
import ctypes as C
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "threefry_golden.txt")


def _gold():
    return np.loadtxt(GOLD, dtype=np.uint64, comments="#").astype(np.uint32)


def test_threefry_block_matches_reference(oracle):
    lib = oracle.load()
    g = _gold()
    out = (C.c_uint32 * 4)()
    for row in g:
        cin = (C.c_uint32 * 4)(*[int(v) for v in row[0:4]])
        key = (C.c_uint32 * 4)(*[int(v) for v in row[4:8]])
        lib.oracle_threefry4x32(cin, key, out)
        assert list(out) == [int(v) for v in row[8:12]]


def _neutron_rows():
    g = _gold()
    m = ((g[:, 1] == 0xF00DCAFE) & (g[:, 2] == 0xDEADBEEF) & (g[:, 3] == 0xBEEFF00D) &
         (g[:, 5] == 0xDECAFBAD) & (g[:, 6] == 0xFACEBEAD) & (g[:, 7] == 0x12345678))
    rows = g[m]
    assert len(rows) == 90
    return rows


def _draws(lib, tid, ctr, on_device=0, ctx=None):
    tid = np.ascontiguousarray(tid, np.uint32)
    ctr = np.ascontiguousarray(ctr, np.float32)
    out = np.zeros(len(tid), np.float32)
    rc = lib.ramp_rng_eval(ctx, tid.ctypes.data_as(C.POINTER(C.c_uint32)),
                           ctr.ctypes.data_as(C.POINTER(C.c_float)),
                           out.ctypes.data_as(C.POINTER(C.c_float)), len(tid), on_device)
    assert rc == 0
    return out


def test_neutron_draws_match_reference_u01(lib):
    rows = _neutron_rows()
    # counters below 2^24 are exact as floats; 2^24 itself too
    keep = rows[:, 0] <= (1 << 24)
    rows = rows[keep]
    out = _draws(lib, rows[:, 4], rows[:, 0].astype(np.float32))
    assert np.array_equal(out.view(np.uint32), rows[:, 12])


def test_float_counter_semantics(lib):
    # the counter is a float incremented by 1.0f: above 2^24 the increment
    # is lost, so draws repeat (rand.h:15,18)
    a = _draws(lib, [7, 7], np.array([2.0**24, 2.0**24 + 1], np.float32))
    assert a[0] == a[1]
    b = _draws(lib, [7, 7], np.array([100.0, 101.0], np.float32))
    assert b[0] != b[1]


def test_u01_closed_interval_endpoints():
    # the map ((w&0x7fffffc0)+(w&0x40))*2^-31 reaches 0 and 1 exactly
    def u01(w):
        w = np.uint32(w)
        return np.float32(np.uint32((w & np.uint32(0x7fffffc0)) + (w & np.uint32(0x40)))) * np.float32(2.0**-31)
    assert u01(0) == 0.0
    assert u01(0x7fffffff) == 1.0
    assert u01(0x80000000) == 0.0


@pytest.mark.gpu
def test_device_draws_equal_host(lib, gctx):
    rng = np.random.default_rng(3)
    tid = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    ctr = rng.integers(0, 2**24, 100000).astype(np.float32)
    h = _draws(lib, tid, ctr)
    d = _draws(lib, tid, ctr, on_device=1, ctx=gctx)
    assert np.array_equal(h.view(np.uint32), d.view(np.uint32))
