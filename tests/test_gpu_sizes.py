"""Ragged and tiny history counts on the GPU against the oracle (minimal,
LET-like, the guide variants and the triple-axis instrument): the source
segment's persistent grid shrinks to the fewest waves that keep its number
of 64-history rounds, so counts below one wave, just above a wave and just
above a round of the whole grid all run with partly filled waves.  Full
state (keep_state) and fast mode (histograms only) both.  Histograms are
compared in float-atomic mode: at these counts a detector can hold a single
event whose weight lies below the resolution of the deterministic mode's
fixed-point sums (2^-58 of the run's weight bound times its history count),
which those sums round to zero."""
import pytest

from tests.parity_util import compare_states, instrument_cases, run_gpu, run_oracle

CASES = instrument_cases()


def _hist_equal(gh, oh):
    for (gw, gw2), (ow, ow2) in zip(gh, oh):
        tot = max(float(abs(ow).sum()), 1e-300)
        assert abs(gw - ow).max() <= 1e-6 * tot
        assert abs(gw2 - ow2).max() <= 1e-6 * max(float(abs(ow2).sum()), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["minimal", "let_like", "guide_variants", "tas"])
@pytest.mark.parametrize("n", [1, 63, 65, 4097, 100003])
def test_ragged_counts(case, n):
    path, kw = CASES[case]
    print(f"case {case} n={n}", flush=True)
    o = run_oracle(path, kw, n, [77])
    g, gh, st = run_gpu(path, kw, n, [77], hist_mode="atomic")
    d = compare_states(g, o["state"])
    assert d["dead"] == 0 and d["iidx"] == 0 and d["ctr"] == 0 and d["beyond_rows"] == 0, d
    _hist_equal(gh, o["hist"])
    _, fh, _ = run_gpu(path, kw, n, [77], keep_state=False, hist_mode="atomic")
    _hist_equal(fh, o["hist"])
