"""Batched settings (ramp_batch_run) on the GPU: a batch changes only the
scheduling, so every setting's histograms must equal a plain run of that
setting with the same seed and count (bit for bit in deterministic mode),
on first launch and on graph replays, for every launch shape."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INSTR = os.path.join(ROOT, "ramp_amd", "instruments")
sys.path.insert(0, INSTR)
import tas_scan  # noqa: E402

TAS = os.path.join(INSTR, "tas_scan.json")


def _scan(n, nmin, nmax, seed=4):
    s, c = tas_scan.settings(n, seed=seed, nmin=nmin, nmax=nmax)
    # two elastic points so that some histograms are well filled
    for k in (0, 1):
        s[k] = dict(s[k], Ef=s[k]["Ei"], theta_a=s[k]["theta_m"])
    return s, np.array(c, dtype=np.uint64)


def _plain(settings, counts, seeds, mode):
    import ramp_amd
    out = []
    for s, c, sd in zip(settings, counts, seeds):
        inst = ramp_amd.Instrument(TAS, hist_mode=mode, chunk=int(c), **s)
        inst.execute(int(c), seeds=[int(sd)])
        out.append(inst.hist64("tube"))
        inst.close()
    return out


def _batch_hists(b):
    return [inst.hist64("tube") for inst in b.instruments]


@pytest.mark.gpu
@pytest.mark.parametrize("npt,wg", [(1, 64), (4, 256), (8, 512)])
def test_batch_equals_plain_runs_deterministic(npt, wg):
    import ramp_amd
    settings, counts = _scan(12, 2e4, 4e5)
    seeds = np.arange(100, 112, dtype=np.uint32)
    print(f"batch npt={npt} wg={wg} settings={len(settings)} total={int(counts.sum())}", flush=True)
    ref = _plain(settings, counts, seeds, "deterministic")
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="deterministic")
    st = b.execute(counts, seeds=seeds, npt=npt, wg=wg)
    assert st["n_hist"] == int(counts.sum())
    got = _batch_hists(b)
    assert sum(float(w.sum()) for w, _ in ref) > 0
    for (gw, gq), (rw, rq) in zip(got, ref):
        assert np.array_equal(gw, rw) and np.array_equal(gq, rq)
    b.close()


@pytest.mark.gpu
def test_batch_graph_replays_accumulate_exactly():
    # replays of the captured graph, with other eager GPU work and host
    # synchronisations between them, keep adding the same fixed-point sums
    import ramp_amd
    settings, counts = _scan(6, 2e4, 2e5, seed=9)
    seeds = np.arange(7, 13, dtype=np.uint32)
    ref = _plain(settings, counts, seeds, "deterministic")
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="deterministic")
    for k in range(1, 4):
        print(f"replay {k}", flush=True)
        b.execute(counts, seeds=seeds, npt=2, wg=128)
        _plain(settings[:1], counts[:1], seeds[:1], "atomic")  # eager work in between
        for (gw, gq), (rw, rq) in zip(_batch_hists(b), ref):
            assert np.array_equal(gw, k * rw) and np.array_equal(gq, k * rq)
    b.close()


@pytest.mark.gpu
def test_batch_atomic_mode_close_to_plain():
    import ramp_amd
    settings, counts = _scan(8, 5e4, 3e5, seed=5)
    seeds = np.arange(40, 48, dtype=np.uint32)
    ref = _plain(settings, counts, seeds, "deterministic")
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="atomic")
    b.execute(counts, seeds=seeds)
    for (gw, gq), (rw, rq) in zip(_batch_hists(b), ref):
        tot = max(float(rw.sum()), 1e-300)
        assert np.abs(gw - rw).max() <= 1e-5 * tot
    b.close()


@pytest.mark.gpu
def test_batch_matches_oracle():
    import ramp_amd
    from oracle import pyoracle
    settings, counts = _scan(3, 3e4, 6e4, seed=6)
    seeds = np.array([5, 6, 7], dtype=np.uint32)
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="deterministic")
    b.execute(counts, seeds=seeds, npt=4, wg=256)
    for s, c, sd, (gw, gq) in zip(settings, counts, seeds, _batch_hists(b)):
        o = pyoracle.run(ramp_amd.Instrument(TAS, **s), int(c), [int(sd)])
        ow, oq = o["hist"][0]
        tot = max(float(ow.sum()), 1e-300)
        assert np.abs(gw - ow).max() <= 1e-6 * tot
        assert np.abs(gq - oq).max() <= 1e-6 * max(float(oq.sum()), 1e-300)
    b.close()


@pytest.mark.gpu
def test_batch_uneven_and_empty_settings():
    # a setting with no histories and one with a count not a multiple of any
    # tile size; the others are unaffected
    import ramp_amd
    settings, counts = _scan(4, 1e4, 5e4, seed=8)
    counts[1] = 0
    counts[2] = 12345
    seeds = np.array([1, 2, 3, 4], dtype=np.uint32)
    ref = _plain(settings, counts, seeds, "deterministic")
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="deterministic")
    b.execute(counts, seeds=seeds, npt=16, wg=256)
    got = _batch_hists(b)
    assert float(got[1][0].sum()) == 0.0
    for k in (0, 2, 3):
        assert np.array_equal(got[k][0], ref[k][0])
    b.close()


@pytest.mark.gpu
def test_autotuner_picks_a_measured_shape_and_resets():
    import ramp_amd
    settings, counts = _scan(8, 1e5, 1e6, seed=3)
    b = ramp_amd.BatchInstrument(TAS, settings, hist_mode="atomic")
    best, res = b.tune(counts, scale=1 / 8, npts=(1, 4), wgs=(128, 256), repeats=1)
    assert best in res and res[best] == min(res.values()) and len(res) == 4
    assert all(float(w.sum()) == 0.0 for w, _ in _batch_hists(b))
    b.close()
