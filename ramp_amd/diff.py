"""Gradients of histogram losses through the MI355X path (no reference
counterpart).

DiffInstrument wraps an instrument whose sample is an SPowderDiff and whose
source is a single-temperature MCont.  Its parameter vector is

    theta = [T1, sigma_inc, sigma_abs, U, d_0, ..., d_{n-1}]

The proposal the sample draws from stays at theta0 (the values the
instrument was built with), so every history follows the same trajectory
at any theta and only its weight changes; `histogram(theta)` is one forward
run (ramp_run) and `adjoint(theta, dL_dh)` one adjoint pass (ramp_adjoint)
over the same histories.  RampHistogram is the torch.autograd.Function
joining the two; torch is used only at this boundary.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from . import _slots as S
from .components import SPowderDiff
from .instrument import Context, Instrument


class DiffInstrument:
    def __init__(self, fn, sample="sample", detector=None, n_hist=1_000_000, seeds=(1,),
                 chunk=None, device=0, **variables):
        self.fn = fn
        self.variables = dict(variables)
        self.sample_name = sample
        self.n_hist = int(n_hist)
        self.chunk = int(chunk) if chunk else self.n_hist
        self.seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        n_chunks = (self.n_hist + self.chunk - 1) // self.chunk
        if len(self.seeds) < n_chunks:
            raise ValueError(f"need {n_chunks} seeds")
        self.context = Context(device)
        base = self._instrument()
        comp, samp = self._sample(base)
        if not isinstance(samp, SPowderDiff):
            raise TypeError(f"component {sample!r} is not an SPowderDiff")
        self.kw = samp.kw
        self.n_lines = int(samp.num_lines)
        self.theta0 = np.concatenate([[float(base.source.p[S.SP_T1]), samp.model["sigma_inc"],
                                       samp.model["sigma_abs"], samp.model["U"]],
                                      samp.model["d_spacing"]]).astype(np.float64)
        self.proposal = {k: samp.model[k] for k in ("d_spacing", "F2", "U", "sigma_inc", "sigma_abs")}
        self.F2 = samp.model["F2"]
        base.plan_desc()
        names = [c.name for c in base._comp_list if getattr(c.scat, "n_bins", 0)]
        if detector is None:
            if len(names) != 1:
                raise ValueError(f"name the loss detector among {names}")
            detector = names[0]
        self.detector = detector
        self.det_slot = names.index(detector)
        self.n_bins = int(next(c.scat.n_bins for c in base._comp_list if c.name == detector))
        base.close()
        self._cache = None
        self.last_forward_ms = self.last_adjoint_ms = None
        self.last_forward_seg_ms = None   # the forward's source segment alone

    @property
    def n_params(self):
        return 4 + self.n_lines

    def param_names(self):
        return ["T1", "sigma_inc", "sigma_abs", "U"] + [f"d{j}" for j in range(self.n_lines)]

    def _instrument(self):
        return Instrument(self.fn, self.context, None, hist_mode="deterministic", chunk=self.chunk,
                          **self.variables)

    def _sample(self, inst):
        for b in inst.blocks:
            if self.sample_name in b.components:
                c = b.components[self.sample_name]
                return c, c.scat
        raise KeyError(self.sample_name)

    def build(self, theta):
        """An Instrument at model parameters theta (proposal theta0)."""
        theta = np.asarray(theta, dtype=np.float64)
        if theta.shape != (self.n_params,):
            raise ValueError(f"theta has shape {theta.shape}, expected ({self.n_params},)")
        inst = self._instrument()
        inst.source.p[S.SP_T1] = np.float32(theta[0])
        comp, old = self._sample(inst)
        kw = dict(self.kw)
        kw.update(d_spacing=theta[4:], F2=self.F2, U=float(theta[3]), sigma_inc=float(theta[1]),
                  sigma_abs=float(theta[2]), proposal=self.proposal, idx=old.idx)
        comp.scat = SPowderDiff(**kw)
        return inst

    def _inst_at(self, theta):
        key = tuple(np.asarray(theta, dtype=np.float64).tolist())
        if self._cache is not None and self._cache[0] == key:
            return self._cache[1]
        if self._cache is not None:
            self._cache[1].close()
        inst = self.build(theta)
        self._cache = (key, inst)
        return inst

    def histogram(self, theta):
        """(w, w2) float64 sums of the loss detector at theta (one forward run)."""
        inst = self._inst_at(theta)
        inst.reset()
        inst.execute(self.n_hist, seeds=self.seeds)
        self.last_forward_ms = inst.last_stats["ms_total"]
        self.last_forward_seg_ms = inst.last_stats["ms_source_segment"]
        return inst.hist64(self.detector)

    def adjoint(self, theta, dL_dh):
        """dL/dtheta (float64) given dL/dh of the loss detector at theta."""
        inst = self._inst_at(theta)
        lib = inst._ensure_plan()
        dl = np.ascontiguousarray(np.asarray(dL_dh, dtype=np.float32).reshape(-1))
        if dl.size != self.n_bins:
            raise ValueError(f"dL/dh has {dl.size} values, the detector {self.n_bins} bins")
        grad = np.zeros(self.n_params, dtype=np.float64)
        capi.check(lib.ramp_adjoint(inst._plan, self.n_hist, self.chunk,
                                    self.seeds.ctypes.data_as(C.POINTER(C.c_uint32)), len(self.seeds),
                                    self.det_slot, dl.ctypes.data_as(C.POINTER(C.c_float)),
                                    self.n_bins, grad.ctypes.data_as(C.POINTER(C.c_double)),
                                    self.n_params, None))
        self.last_adjoint_ms = inst.stats()["ms_total"]
        return grad

    def close(self):
        if self._cache is not None:
            self._cache[1].close()
            self._cache = None
        self.context.close()


def weighted_mse(h, t, var, eps):
    """sum_b (h_b - t_b)^2 / (var_b + eps) and its gradient in h."""
    den = var + eps
    r = h - t
    return float(np.sum(r * r / den)), 2.0 * r / den


def _torch_function():
    import torch

    class RampHistogram(torch.autograd.Function):
        """h = RampHistogram.apply(theta, dinst): the loss detector's
        histogram (float64, on theta's device) at parameters theta; the
        backward pass is one ramp_adjoint run."""

        @staticmethod
        def forward(ctx, theta, dinst):
            th = theta.detach().cpu().double().numpy()
            w, _ = dinst.histogram(th)
            ctx.dinst = dinst
            ctx.theta = th
            return torch.from_numpy(w).to(theta.device)

        @staticmethod
        def backward(ctx, grad_h):
            g = ctx.dinst.adjoint(ctx.theta, grad_h.detach().cpu().float().numpy())
            return torch.from_numpy(g).to(grad_h.device), None

    return RampHistogram


_RH = None


def __getattr__(name):
    global _RH
    if name == "RampHistogram":
        if _RH is None:
            _RH = _torch_function()
        return _RH
    raise AttributeError(name)
