"""SRescal on the GPU against the oracle, same seeds.

Bar: final states as in the other parity tests, and the event lists -
appended on the device after every chunk from the histories that end
detected (se > 0) - identical to the oracle's, bit for bit and in the same
order, across chunks and across runs (the reference's lists accumulate,
rescal.py:51-62); ramp_hist_reset empties them."""
import os

import numpy as np
import pytest

from ramp_amd import Instrument
from tests.parity_util import ROOT, compare_states

PATH = os.path.join(ROOT, "ramp_amd", "instruments", "rescal.json")
CASES = {"axis": dict(tx=0.0, tz=1.0, arm_rot=0.0),
         "offset": dict(tx=0.3, tz=1.0, arm_rot=float(np.arctan2(0.3, 1.0))),
         "unit": dict(tx=0.6, tz=0.8, arm_rot=float(np.arctan2(0.6, 0.8)))}


def _oracle(kw, n, seeds, chunk=0):
    from oracle import pyoracle
    return pyoracle.run(Instrument(PATH, **kw), n, seeds, chunk=chunk, keep_state=True)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("n,chunk,seeds", [(65536, None, [5]), (100000, 30000, [1, 2, 3, 4])])
def test_rescal_parity(case, n, chunk, seeds):
    kw = CASES[case]
    print(f"rescal {case} N={n} chunk={chunk} seeds={seeds}", flush=True)
    inst = Instrument(PATH, hist_mode="deterministic", chunk=chunk, **kw)
    inst.execute(n, keep_state=True, seeds=seeds)
    last = n - (len(seeds) - 1) * inst.last_chunk
    g = inst.state(last)
    ev = np.array(inst.data()["sample"])
    inst.close()
    o = _oracle(kw, n, seeds, chunk or 0)
    d = compare_states(g, o["state"])
    ref = o["rescal"][0].astype(np.float64)
    print(case, d, ev.shape, ref.shape, flush=True)
    assert d["ctr"] == 0 and d["dead"] == 0 and d["beyond_rows"] == 0
    assert ev.shape == ref.shape and ref.shape[1] > n // 2
    assert np.array_equal(ev, ref)


@pytest.mark.gpu
def test_rescal_lists_accumulate_and_reset():
    kw = CASES["offset"]
    inst = Instrument(PATH, hist_mode="deterministic", chunk=40000, **kw)
    inst.execute(40000, seeds=[7])
    a = np.array(inst.data()["sample"])
    inst.execute(80000, seeds=[8, 9])
    b = np.array(inst.data()["sample"])
    o1 = _oracle(kw, 40000, [7], 40000)["rescal"][0]
    assert np.array_equal(a, o1.astype(np.float64))
    assert np.array_equal(b[:, :a.shape[1]], a)
    assert b.shape[1] > 2.5 * a.shape[1]
    inst.reset()
    assert inst.data()["sample"][0].size == 0
    inst.execute(40000, seeds=[7])
    # after a reset the list starts again (each run starts its slots afresh)
    assert np.array_equal(np.array(inst.data()["sample"]), a)
    inst.close()
