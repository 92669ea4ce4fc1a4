"""Fit the polynomial kernels of the project FP32 math library (rmath.h).

Method: discrete minimax fit by Lawson's iteratively re-weighted least
squares on a dense Chebyshev-distributed grid, in float64, for each
reduced-argument kernel below.  The coefficients are then rounded to float32
and printed as C hex-float literals, which are pasted into rmath.h.

Kernels (z = r*r unless stated):
  sin   sin(r)   = r + r^3 * S(z)            r in [0, pi/4 + 0.03]
  cos   cos(r)   = 1 - z/2 + z^2 * C(z)      r in [0, pi/4 + 0.03]
  exp   e^r      = 1 + r + r^2 * E(r)        r in [-0.36, 0.36]
  log   log1p(f) = f - f^2/2 + s*(f^2/2 + z*L(z)),  s = f/(2+f), z = s^2
  tanh  tanh(x)  = x + x^3 * T(z)            x in [0, 0.625]
  atan  atan(v)  = v + v^3 * A(z)            v in [0, tan(pi/8) + 0.01]
  asin  asin(x)  = x + x^3 * B(z)            x in [0, 0.5]

The weight of each fit converts the error of the fitted factor into the
relative error of the final function value, so the reported figure is the
truncation error of the kernel in units of float32 ulps (2^-24 relative).

Run:  python gen_rmath_coeffs.py
"""
import numpy as np


def lawson(x, f, w, deg, iters=400):
    """Weighted minimax polynomial p (in x, ascending coeffs) for f on grid x.

    Minimises max |w*(p - f)| by Lawson's algorithm."""
    V = np.vander(x, deg + 1, increasing=True)
    lam = np.full(x.size, 1.0 / x.size)
    for _ in range(iters):
        sw = np.sqrt(lam) * w
        c, *_ = np.linalg.lstsq(V * sw[:, None], f * sw, rcond=None)
        err = np.abs(w * (V @ c - f))
        lam = lam * err
        s = lam.sum()
        if s == 0:
            break
        lam /= s
    return c


def cheb(a, b, n=20000):
    k = np.arange(n)
    t = np.cos(np.pi * (2 * k + 1) / (2 * n))
    return np.sort(0.5 * (a + b) + 0.5 * (b - a) * t)


def report(name, c, err):
    c32 = c.astype(np.float32)
    print(f"/* {name}: degree {len(c) - 1}, kernel error {err / 2.0**-24:.3g} ulp */")
    for i, v in enumerate(c32):
        mant, ex = float(v).hex().split("p")
        mant = mant.rstrip("0").rstrip(".")
        print(f"#define RM_{name.upper()}{i} {mant}p{ex}f  /* {float(v)!r} */")


def fit_sin():
    r = cheb(1e-4, np.pi / 4 + 0.03)
    z = r * r
    f = (np.sin(r) - r) / r**3
    w = r**3 / np.sin(r)
    c = lawson(z, f, w, 3)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("sin", c, err)


def fit_cos():
    r = cheb(1e-3, np.pi / 4 + 0.03)
    z = r * r
    f = (np.cos(r) - 1 + z / 2) / z**2
    w = z**2 / np.cos(r)
    c = lawson(z, f, w, 3)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("cos", c, err)


def fit_exp():
    r = cheb(-0.36, 0.36)
    r = r[np.abs(r) > 1e-5]
    f = (np.exp(r) - 1 - r) / r**2
    w = r**2 / np.exp(r)
    c = lawson(r, f, w, 5)
    err = np.max(np.abs(w * (np.polyval(c[::-1], r) - f)))
    report("exp", c, err)


def fit_log():
    s = cheb(1e-4, 0.1716 + 0.002)
    z = s * s
    f = (2 * np.arctanh(s) / s - 2) / z
    fval = 2 * s / (1 - s)
    w = s * z / np.log1p(fval)
    c = lawson(z, f, w, 3)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("log", c, err)


def fit_tanh():
    x = cheb(1e-4, 0.625)
    z = x * x
    f = (np.tanh(x) - x) / x**3
    w = x**3 / np.tanh(x)
    c = lawson(z, f, w, 5)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("tanh", c, err)


def fit_atan():
    v = cheb(1e-4, np.tan(np.pi / 8) + 0.01)
    z = v * v
    f = (np.arctan(v) - v) / v**3
    w = v**3 / np.arctan(v)
    c = lawson(z, f, w, 5)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("atan", c, err)


def fit_asin():
    x = cheb(1e-4, 0.5)
    z = x * x
    f = (np.arcsin(x) - x) / x**3
    w = x**3 / np.arcsin(x)
    c = lawson(z, f, w, 5)
    err = np.max(np.abs(w * (np.polyval(c[::-1], z) - f)))
    report("asin", c, err)


if __name__ == "__main__":
    for fn in (fit_sin, fit_cos, fit_exp, fit_log, fit_tanh, fit_atan, fit_asin):
        fn()
