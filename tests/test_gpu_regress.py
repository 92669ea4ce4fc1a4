"""Regression tests of host-side run bookkeeping on the GPU, each against
the oracle.

- Queue growth: a small full-state run, then a larger fast-mode run, then a
  larger full-state run.  The fast run must grow the full-state fields too
  (they are in use once a full-state run has happened), or the last run's
  queue stores pass the end of the 14 full-state fields.
- Deterministic budget: 70 deterministic runs of a counter that takes every
  history's whole weight.  A bin's int64 sum grows by 2^57 per run at the
  scale the first run sets, which would wrap at run 64; the plan folds the
  sums into FP64 host accumulators before passing 2^62.
"""
import json

import numpy as np
import pytest

from tests.parity_util import compare_states, instrument_cases

CASES = instrument_cases()

COUNTER = {
    "beam": {
        "linear": True,
        "source": {"source": True,
                   "moderator_kernel": {"name": "MFlat", "mod_dim": [0.01, 0.01], "target_dim": [0.01, 0.01],
                                        "target_dist": 2.0, "E_min": 2.0, "E_max": 20.0}},
        "counter": {"position": [0.0, 0.0, 1.0],
                    "geom_kernel": {"name": "GPlane", "width": 1.0, "height": 1.0},
                    "scat_kernel": {"name": "SCounter"}},
    }
}


@pytest.mark.gpu
def test_queue_growth_full_after_fast():
    import ramp_amd
    from oracle import pyoracle
    path, kw = CASES["let_like"]
    inst = ramp_amd.Instrument(path, hist_mode="deterministic", **kw)
    runs = [(1_000_000, True, 7), (2_000_000, False, 8), (2_000_000, True, 9)]
    for n, keep, seed in runs:
        print(f"let_like n={n} keep_state={keep} seed={seed}", flush=True)
        inst.execute(n, keep_state=keep, seeds=[seed])
        assert inst.stats()["n_segments"] > 1  # the run went through the queues
    g = inst.state(runs[-1][0])
    inst.close()
    o = pyoracle.run(ramp_amd.Instrument(path, **kw), runs[-1][0], [runs[-1][2]], keep_state=True)
    d = compare_states(g, o["state"])
    assert d["dead"] == 0 and d["iidx"] == 0 and d["ctr"] == 0 and d["beyond_rows"] == 0


@pytest.mark.gpu
def test_deterministic_sums_fold_before_int64_overflow(tmp_path):
    import ramp_amd
    from oracle import pyoracle
    p = tmp_path / "counter.json"
    p.write_text(json.dumps(COUNTER))
    n, runs = 1000, 70
    print(f"counter n={n} runs={runs}", flush=True)
    inst = ramp_amd.Instrument(str(p), hist_mode="deterministic")
    for r in range(runs):
        inst.execute(n, seeds=[r])
        assert inst.stats()["n_det_overflow"] == 0
    w, w2 = inst.hist64("counter")
    inst.close()
    ow = ow2 = 0.0
    for r in range(runs):
        o = pyoracle.run(ramp_amd.Instrument(str(p)), n, [r])
        ow += o["hist"][0][0][0]
        ow2 += o["hist"][0][1][0]
    # every history reaches the counter with weight 1/n: 1 per run
    assert ow == pytest.approx(runs, rel=1e-6)
    assert abs(w[0] - ow) <= 1e-6 * ow and abs(w2[0] - ow2) <= 1e-6 * ow2
