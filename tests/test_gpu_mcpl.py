"""SMCPLIn / SMCPLOut on the GPU against the oracle, same seeds.

Bar: final states as in the other parity tests (flags, indices and
counters identical; FP32 values within 1e-5), deterministic histograms
within 1e-6 of the total, and the SMCPLOut records - packed in the
segment kernel, compacted by ramp_mcpl_read - identical to the oracle's,
bit for bit and in history order, in the file with the reference's header.
"""
import numpy as np
import pytest

from ramp_amd import Instrument
from ramp_amd.mcpl import read_mcpl, reference_header
from tests.parity_util import compare_states, mcpl_cases, run_gpu, run_oracle


def _hist_close(gh, oh, tol_rel=1e-6):
    for (gw, gw2), (ow, ow2) in zip(gh, oh):
        assert np.abs(gw - ow).max() <= tol_rel * max(float(np.abs(ow).sum()), 1e-300)
        assert np.abs(gw2 - ow2).max() <= tol_rel * max(float(np.abs(ow2).sum()), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("n,chunk,seeds", [(4096, None, [7]), (3000, 1000, [3, 9, 4096]),
                                           (65536, None, [11])])
def test_mcpl_in_parity(n, chunk, seeds):
    path, kw = mcpl_cases()["mcpl_in"]
    print(f"mcpl_in N={n} chunk={chunk} seeds={seeds}", flush=True)
    g, gh, st = run_gpu(path, kw, n, seeds, chunk=chunk)
    o = run_oracle(path, kw, n, seeds, chunk=chunk or 0)
    d = compare_states(g, o["state"])
    print(d, st, flush=True)
    assert d["ctr"] == 0 and d["dead"] == 0 and d["iidx"] == 0 and d["beyond_rows"] == 0
    _hist_close(gh, o["hist"])
    assert sum(float(w.sum()) for w, _ in gh) > 0


def _gpu_out(case, n, seeds, chunk=None):
    path, kw = mcpl_cases()[case]
    inst = Instrument(path, hist_mode="deterministic", chunk=chunk, **kw)
    inst.execute(n, seeds=seeds)
    c = inst._mcpl_out[0].scat
    hists = [inst.hist64(x.name) for x in inst._comp_list if getattr(x.scat, "n_bins", 0)]
    stats = inst.stats()
    inst.close()
    return c, hists, stats


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["mcpl_out", "mcpl_out_std"])
@pytest.mark.parametrize("n,chunk,seeds", [(20000, None, [3]), (50000, 20000, [1, 2, 3]),
                                           (400000, None, [5])])
def test_mcpl_out_parity(case, n, chunk, seeds):
    path, kw = mcpl_cases()[case]
    print(f"{case} N={n} chunk={chunk} seeds={seeds}", flush=True)
    c, gh, st = _gpu_out(case, n, seeds, chunk)
    o = run_oracle(path, kw, n, seeds, chunk=chunk or 0)
    ref = o["mcpl"][0]
    print(case, c.records.shape, ref.shape, st["ms_mcpl_compact"], st["ms_mcpl_write"], flush=True)
    assert c.records.shape == ref.shape and ref.shape[0] > 0
    assert np.array_equal(c.records.view(np.uint32), ref.view(np.uint32))
    _hist_close(gh, o["hist"])
    # the file: the reference's header, then the records
    raw = open(c.filename, "rb").read()
    assert raw[:56] == reference_header(ref.shape[0])
    hdr, rec = read_mcpl(c.filename)
    assert np.array_equal(rec.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_mcpl_in_reads_gpu_written_file():
    """A file the GPU wrote (MCPL's own packing) read back by SMCPLIn on the
    GPU and on the oracle."""
    import os
    c, _, _ = _gpu_out("mcpl_out_std", 20000, [4])
    out = c.records
    path, kw = mcpl_cases()["mcpl_in"]
    kw = dict(kw, mcpl_file=repr(c.filename))
    g, _, _ = run_gpu(path, kw, out.shape[0], [4])
    o = run_oracle(path, kw, out.shape[0], [4])
    d = compare_states(g, o["state"])
    assert d["ctr"] == 0 and d["dead"] == 0 and d["beyond_rows"] == 0
    assert os.path.getsize(c.filename) == 56 + 44 * out.shape[0]
