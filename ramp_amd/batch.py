"""BatchInstrument: many settings of one instrument in one persistent launch.

The reference runs an instrument scan as one Instrument per setting and a
loop of execute() calls (mcramp.py:271, 403-455).  BatchInstrument keeps
that model - each setting is an ordinary Instrument built from the same
JSON with its own $variables$, and its data() is read the same way - but
runs all settings together through ramp_batch_run: one launch with a device
work queue over (setting, history range) tiles, captured once as a HIP graph
and replayed.  Setting s runs counts[s] histories as one chunk with its own
seed, so its histograms equal those of a plain run of that setting (bit for
bit in deterministic mode); only the scheduling differs.

tune() is the launch-shape autotuner: it times the batch on a scaled-down
copy of the counts for every (histories per lane per tile, workgroup size)
pair and keeps the fastest.
"""
import ctypes as C
import itertools

import numpy as np

from . import _capi as capi
from .instrument import Context, Instrument

NPT_CHOICES = (1, 2, 4, 8)
WG_CHOICES = (64, 128, 256, 512)


class BatchInstrument:
    def __init__(self, fn, settings, hist_mode="atomic", device=0, seed=None, **common):
        """settings: list of {variable: value} dicts, one per setting;
        common: variables shared by all settings."""
        if not settings:
            raise ValueError("a batch needs at least one setting")
        self.context = Context(device)
        self.hist_mode = hist_mode
        self.instruments = [Instrument(fn, self.context, None, hist_mode=hist_mode, device=device,
                                       **{**common, **s}) for s in settings]
        self._seed_rng = np.random.default_rng(seed)
        self._batch = None
        self.npt = 0
        self.wg = 0
        self.tuning = None
        self.last_stats = None

    def __len__(self):
        return len(self.instruments)

    def _ensure_batch(self):
        lib = capi.load_library()
        if self._batch is None:
            for inst in self.instruments:
                inst._ensure_plan()
            plans = (C.c_void_p * len(self.instruments))(*[i._plan.value for i in self.instruments])
            h = C.c_void_p()
            capi.check(lib.ramp_batch_create(self.context.handle, plans, len(self.instruments),
                                             C.byref(h)))
            self._batch = h
        return lib

    def _opts(self, npt, wg):
        o = capi.RunOpts()
        o.hist_mode = (capi.RAMP_HIST_DETERMINISTIC if self.hist_mode == "deterministic"
                       else capi.RAMP_HIST_ATOMIC)
        o.npt = int(npt)
        o.wg = int(wg)
        return o

    def execute(self, counts, seeds=None, npt=None, wg=None):
        """Run counts[s] histories of every setting s in one launch."""
        lib = self._ensure_batch()
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        if counts.shape != (len(self),):
            raise ValueError(f"need {len(self)} counts, got shape {counts.shape}")
        if seeds is None:
            seeds = self._seed_rng.integers(0, 4097, size=len(self))
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        wb = np.array([inst.source.weight_bound(max(int(c), 1))
                       for inst, c in zip(self.instruments, counts)], dtype=np.float64)
        opts = self._opts(npt if npt is not None else self.npt, wg if wg is not None else self.wg)
        capi.check(lib.ramp_batch_run(self._batch, counts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      seeds.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      wb.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts)))
        for inst, c in zip(self.instruments, counts):
            inst.last_chunk = int(c)
            inst._hist_stale = True
        self.last_seeds = seeds
        self.last_stats = self.stats()
        return self.last_stats

    def execute_sequential(self, counts, seeds=None):
        """The reference's way: one plain run per setting, in turn (for
        comparison; same histories as execute with the same seeds)."""
        counts = np.asarray(counts, dtype=np.uint64)
        if seeds is None:
            seeds = self._seed_rng.integers(0, 4097, size=len(self))
        for inst, c, s in zip(self.instruments, counts, seeds):
            if int(c) == 0:
                continue
            inst.chunk = int(c)
            inst.execute(int(c), seeds=[int(s)])

    def stats(self):
        lib = capi.load_library()
        st = capi.RunStats()
        capi.check(lib.ramp_batch_stats_read(self._batch, C.byref(st)))
        return {f: getattr(st, f) for f, _ in capi.RunStats._fields_}

    def tune(self, counts, scale=1.0 / 4, npts=NPT_CHOICES, wgs=WG_CHOICES, repeats=3, seed=1):
        """Pick (npt, wg) by timing the batch (device time of the graph
        replay, the least of `repeats`) on counts scaled by `scale`;
        histograms are reset after.  The shapes differ by a few per cent,
        about as much as one timing of a sixteenth of the counts varies, so
        the default times a quarter of them three times: picks made from a
        sixteenth ran the full scan up to 5 % slower than the best shape."""
        c = np.maximum(np.asarray(counts, dtype=np.float64) * scale, 1).astype(np.uint64)
        seeds = np.random.default_rng(seed).integers(0, 4097, size=len(self))
        results = {}
        for npt, wg in itertools.product(npts, wgs):
            self.execute(c, seeds=seeds, npt=npt, wg=wg)  # capture + warm
            t = min(self.execute(c, seeds=seeds, npt=npt, wg=wg)["ms_total"] for _ in range(repeats))
            results[(npt, wg)] = t
        best = min(results, key=results.get)
        self.npt, self.wg = best
        self.tuning = results
        self.reset()
        return best, results

    def reset(self):
        for inst in self.instruments:
            inst.reset()

    def data(self):
        """[inst.data() for each setting]."""
        return [inst.data() for inst in self.instruments]

    def close(self):
        if self._batch is not None:
            capi.load_library().ramp_batch_destroy(self._batch)
            self._batch = None
        for inst in self.instruments:
            inst.close()
        self.context.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
