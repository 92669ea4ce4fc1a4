"""Parameters of the differentiable powder instrument (instruments/powder_diff.json).

40 Bragg lines from numpy.random.default_rng(5): d ~ U[0.8, 3.5] A,
|F|^2 = 10^U[-1, 2] barn, multiplicity 1; source temperature 300 K,
incoherent Sigma_s = 0.2 /cm, absorption Sigma_a = 0.1 /cm at 1.8 A,
Debye-Waller U = 0.005 A^2.  theta = [T, sigma_inc, sigma_abs, U, d_0..d_39].
"""
import numpy as np


def lines(n=40, seed=5):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.8, 3.5, size=n)
    F2 = 10.0 ** rng.uniform(-1.0, 2.0, size=n)
    return d, F2


def theta0(n=40, seed=5):
    d, _ = lines(n, seed)
    return np.concatenate([[300.0, 0.2, 0.1, 0.005], d])


def variables(theta, F2, tail_join=0.0):
    """JSON variables of powder_diff.json for parameter vector theta;
    tail_join > 0 adds the source's 1/E tail above tail_join * k_B * T."""
    theta = np.asarray(theta, dtype=np.float64)
    return {"T": float(theta[0]), "sigma_inc": float(theta[1]), "sigma_abs": float(theta[2]),
            "U": float(theta[3]), "d": [float(x) for x in theta[4:]], "F2": [float(x) for x in F2],
            "tail_join": float(tail_join)}
