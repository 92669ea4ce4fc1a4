"""Settings of the batched triple-axis scan (instruments/tas_scan.json).

256 scan points from numpy.random.default_rng(4):
  Ei ~ U[3, 15] meV, energy transfer dE ~ U[0, min(5, Ei - EF_MIN)] meV,
  Ef = Ei - dE; Bragg angles of PG(002), d = 3.355 A:
  theta = asin(lambda / (2 d)), lambda = 9.0446 / sqrt(E) A (E in meV);
  neutron counts log-uniform in [1e5, 4e6]: floor(10^U[5, log10(4e6)]).
The sample rotation phi_s is fixed at 0.7 rad.  PG(002) reflects no
wavelength above 2d = 6.71 A (E < 1.82 meV), so the final energy is kept at
or above EF_MIN = 2.0 meV.
"""
import math

import numpy as np

D_PG002 = 3.355
# lambda [A] = h / sqrt(2 m E) = 9.0446 / sqrt(E [meV])
LAMBDA_OF_E = 9.0446
EF_MIN = 2.0


def bragg(E, d=D_PG002):
    return math.asin(LAMBDA_OF_E / math.sqrt(E) / (2.0 * d))


def settings(n=256, seed=4, nmin=1e5, nmax=4e6, phi_s=0.7):
    """(list of variable dicts, counts) of n scan points."""
    rng = np.random.default_rng(seed)
    out, counts = [], []
    for _ in range(n):
        Ei = float(rng.uniform(3.0, 15.0))
        dE = float(rng.uniform(0.0, min(5.0, Ei - EF_MIN)))
        Ef = Ei - dE
        out.append({"Ei": Ei, "Ef": Ef, "theta_m": bragg(Ei), "theta_a": bragg(Ef), "phi_s": phi_s})
        counts.append(int(10 ** rng.uniform(math.log10(nmin), math.log10(nmax))))
    return out, counts
