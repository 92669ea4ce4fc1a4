"""History-range shards on the GPU: one chunk of N histories run as a few
shards in turn (ramp_run_opts.key_base / shard_total; Instrument.execute
shard=(first, N)) must give the histograms of the whole chunk bit for bit
in deterministic mode.  This is the way the path spreads over several GPUs
(each takes a range of the chunk, then the int64 moments are summed); here
the shards run on one GPU, one after the other, into the same histograms.
Covered: guides and choppers with compacted segments (LET-like), the
incoherent sample and the 2-D detector (minimal), the monochromator with
its deferred reflections (TAS), a mesh (imaging)."""
import numpy as np
import pytest

from ramp_amd import Instrument, capi
from ramp_amd._capi import RampError
from tests.parity_util import instrument_cases


def _hists(inst):
    return [inst.hist64(c.name) for c in inst.kernel_refs if getattr(c.scat, "n_bins", 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("case,n,cuts", [("let_like", 1_000_000, [0, 123_457, 500_000, 777_777]),
                                         ("minimal", 400_000, [0, 65_536, 200_001]),
                                         ("tas", 300_000, [0, 99_999, 150_000]),
                                         ("imaging", 200_000, [0, 131_072])])
def test_shards_add_up_to_the_chunk(case, n, cuts):
    path, kw = instrument_cases()[case]
    print(f"shards {case} N={n} cuts={cuts}", flush=True)
    whole = Instrument(path, hist_mode="deterministic", chunk=n, **kw)
    whole.execute(n, seeds=[17])
    hw = _hists(whole)
    events = whole.stats()["n_events"]
    whole.close()
    parts = Instrument(path, hist_mode="deterministic", **kw)
    ev = 0
    for lo, hi in zip(cuts, cuts[1:] + [n]):
        parts.execute(hi - lo, seeds=[17], shard=(lo, n))
        ev += parts.stats()["n_events"]
    hp = _hists(parts)
    parts.close()
    assert ev == events and events > 0
    assert len(hw) == len(hp) and len(hw) > 0
    for (a, a2), (b, b2) in zip(hw, hp):
        assert np.array_equal(a, b) and np.array_equal(a2, b2)
        assert a.sum() > 0


@pytest.mark.gpu
def test_shard_arguments_are_checked():
    path, kw = instrument_cases()["minimal"]
    inst = Instrument(path, hist_mode="deterministic", **kw)
    with pytest.raises(RampError):
        inst.execute(1000, seeds=[1], shard=(500, 1000))      # past the chunk's end
    with pytest.raises(RampError):
        inst.execute(1000, seeds=[1], shard=(0, 1000), keep_state=True)
    inst.close()
    mc = Instrument(instrument_cases()["polarisation"][0], hist_mode="deterministic",
                    **instrument_cases()["polarisation"][1])
    with pytest.raises(RampError):
        mc.execute(1000, seeds=[1], shard=(0, 2000))         # full-state plan
    mc.close()
    assert capi.RAMP_E_RANGE < 0
