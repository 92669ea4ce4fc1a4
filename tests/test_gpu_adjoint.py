"""Gradients through the GPU path (ramp_adjoint): against the FP64 CPU
reference gradient of the oracle on the same histories, against same-seed
central finite differences of the GPU forward, and through the
torch.autograd.Function."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INSTR = os.path.join(ROOT, "ramp_amd", "instruments")
sys.path.insert(0, INSTR)
import powder_diff  # noqa: E402

PD = os.path.join(INSTR, "powder_diff.json")


def _dinst(n, seeds=(11,), tail_join=0.0):
    from ramp_amd.diff import DiffInstrument
    d, F2 = powder_diff.lines()
    return DiffInstrument(PD, n_hist=n, seeds=seeds,
                          **powder_diff.variables(powder_diff.theta0(), F2, tail_join))


def _target(di, seed):
    # the same simulator at theta0 * 1.02, other seed (its own histogram and variance)
    from ramp_amd.diff import DiffInstrument
    d, F2 = powder_diff.lines()
    tv = powder_diff.variables(powder_diff.theta0() * 1.02, F2)
    tgt = DiffInstrument(PD, n_hist=di.n_hist, seeds=(seed,), **tv)
    t, t2 = tgt.histogram(tgt.theta0)
    tgt.close()
    return t, t2


@pytest.mark.gpu
@pytest.mark.parametrize("tail_join", [0.0, 2.0])
def test_adjoint_equals_fp64_cpu_reference(tail_join):
    # tail_join 2: the source's 1/E tail above 2 k_B T (52 meV at 300 K)
    from oracle import pyoracle
    n = 200_000
    di = _dinst(n, tail_join=tail_join)
    print(f"adjoint vs FP64 oracle N={n} tail_join={tail_join}", flush=True)
    rng = np.random.default_rng(0)
    dl = rng.normal(size=di.n_bins).astype(np.float32)
    g = di.adjoint(di.theta0, dl)
    ref = pyoracle.adjoint(di.build(di.theta0), n, di.seeds, di.det_slot, dl, di.n_params)
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.all(np.abs(g - ref) <= 2e-4 * np.abs(ref) + 1e-5 * scale), np.abs(g - ref) / scale
    di.close()


@pytest.mark.gpu
def test_forward_histogram_equals_oracle_at_shifted_parameters():
    # away from the proposal the weights carry the model/proposal ratios
    from oracle import pyoracle
    n = 100_000
    di = _dinst(n)
    th = di.theta0 * np.concatenate([[1.03, 0.9, 1.2, 1.1], np.full(di.n_lines, 1.001)])
    w, w2 = di.histogram(th)
    o = pyoracle.run(di.build(th), n, di.seeds)
    ow, ow2 = o["hist"][di.det_slot]
    assert np.abs(w - ow).max() <= 1e-6 * ow.sum()
    assert np.abs(w2 - ow2).max() <= 1e-6 * ow2.sum()
    di.close()


@pytest.mark.gpu
def test_adjoint_matches_central_finite_differences():
    from ramp_amd.diff import weighted_mse
    n = 2_000_000
    di = _dinst(n)
    t, t2 = _target(di, 12)
    eps = 1e-3 * float(np.mean(t2[t2 > 0]))
    h0, _ = di.histogram(di.theta0)
    L0, dldh = weighted_mse(h0, t, t2, eps)
    g = di.adjoint(di.theta0, dldh)
    # the scalar parameters and the four lines with the largest gradients
    idx = [0, 1, 2, 3] + list(4 + np.argsort(-np.abs(g[4:]))[:4])
    for i in idx:
        step = 1e-3 * di.theta0[i]
        e = np.zeros(di.n_params)
        e[i] = step
        Lp, _ = weighted_mse(di.histogram(di.theta0 + e)[0], t, t2, eps)
        Lm, _ = weighted_mse(di.histogram(di.theta0 - e)[0], t, t2, eps)
        fd = (Lp - Lm) / (2 * step)
        print(f"{di.param_names()[i]}: adjoint {g[i]:.6e} fd {fd:.6e}", flush=True)
        assert abs(g[i] - fd) <= 0.02 * abs(fd), (i, g[i], fd)
    di.close()


@pytest.mark.gpu
def test_torch_autograd_function():
    import torch
    from ramp_amd.diff import RampHistogram
    n = 300_000
    di = _dinst(n)
    t, t2 = _target(di, 13)
    eps = 1e-3 * float(np.mean(t2[t2 > 0]))
    theta = torch.tensor(di.theta0, dtype=torch.float64, requires_grad=True)
    h = RampHistogram.apply(theta, di)
    tt = torch.from_numpy(t)
    loss = (((h - tt) ** 2) / (torch.from_numpy(t2) + eps)).sum()
    loss.backward()
    _, dldh = __import__("ramp_amd.diff", fromlist=["weighted_mse"]).weighted_mse(
        h.detach().numpy(), t, t2, eps)
    g = di.adjoint(di.theta0, dldh)
    assert np.allclose(theta.grad.numpy(), g, rtol=1e-5, atol=0)
    di.close()
