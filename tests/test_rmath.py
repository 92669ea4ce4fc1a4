"""The shared FP32 math library: accuracy against float64 references and
bit-identity of host and device evaluation.

The reference kernels use the OpenCL runtime's builtins, whose results are
implementation-defined; both the oracle and the HIP kernels use this one
library, so its accuracy is checked here against the mathematics (float64
numpy), with known special values."""
import ctypes as C

import numpy as np
import pytest

from ramp_amd import _capi

FN = {"sin": 1, "cos": 2, "exp": 3, "log": 4, "pow": 5, "tanh": 6, "atan2": 7, "asin": 8,
      "acos": 9, "fmod": 10, "round": 11, "sqrt": 12, "div": 13}


def ev(lib, name, x, y=None, on_device=0, ctx=None):
    x = np.ascontiguousarray(x, np.float32)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    out = np.zeros_like(x)
    rc = lib.ramp_math_eval(ctx, FN[name], x.ctypes.data_as(C.POINTER(C.c_float)),
                            None if yy is None else yy.ctypes.data_as(C.POINTER(C.c_float)),
                            out.ctypes.data_as(C.POINTER(C.c_float)), x.size, on_device)
    assert rc == 0
    return out


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    spacing = np.abs(np.spacing(np.abs(ref32).astype(np.float32))).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).smallest_subnormal))
    return np.abs(got.astype(np.float64) - ref) / spacing


RNG = np.random.default_rng(1)

CASES = {
    "sin": (lambda: RNG.uniform(-40, 40, 200000), None, np.sin, 1.5),
    "cos": (lambda: RNG.uniform(-40, 40, 200000), None, np.cos, 1.5),
    "exp": (lambda: RNG.uniform(-100, 88, 200000), None, np.exp, 1.5),
    "log": (lambda: np.exp(RNG.uniform(-80, 80, 200000)), None, np.log, 1.5),
    "tanh": (lambda: RNG.uniform(-12, 12, 200000), None, np.tanh, 2.0),
    "asin": (lambda: RNG.uniform(-1, 1, 200000), None, np.arcsin, 2.5),
    "acos": (lambda: RNG.uniform(-1, 1, 200000), None, np.arccos, 2.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_unary_accuracy(lib, name):
    gen, _, ref, bound = CASES[name]
    x = gen().astype(np.float32)
    got = ev(lib, name, x)
    e = ulp_err(got, ref(x.astype(np.float64)))
    assert e.max() <= bound, (name, e.max(), x[np.argmax(e)])


def test_atan2_accuracy(lib):
    y = RNG.uniform(-5, 5, 200000).astype(np.float32)
    x = RNG.uniform(-5, 5, 200000).astype(np.float32)
    got = ev(lib, "atan2", y, x)
    e = ulp_err(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 2.5


def test_pow_accuracy(lib):
    x = np.exp(RNG.uniform(-60, 60, 200000)).astype(np.float32)
    y = RNG.uniform(-2, 2, 200000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    ok = np.isfinite(ref) & (ref < 3e38) & (ref > 1e-37)
    got = ev(lib, "pow", x, y)
    assert ulp_err(got[ok], ref[ok]).max() <= 2.0
    # the Maxwellian site: pow(1/(k_B T), 1.5) for T in 1..1000 K
    T = np.linspace(1, 1000, 5000).astype(np.float32)
    ok_ = (np.float32(1.0) / (np.float32(1.38066e-23) * T)).astype(np.float32)
    got = ev(lib, "pow", ok_, np.full_like(ok_, 1.5))
    assert ulp_err(got, np.power(ok_.astype(np.float64), 1.5)).max() <= 2.0
    # y == 2 is the correctly rounded square (detector energy, detector1D.cl:51)
    xs = RNG.uniform(0, 5000, 10000).astype(np.float32)
    assert np.array_equal(ev(lib, "pow", xs, np.full_like(xs, 2.0)), xs * xs)


def test_fmod_exact(lib):
    x = RNG.uniform(-1e4, 1e4, 100000).astype(np.float32)
    y = RNG.uniform(0.01, 10, 100000).astype(np.float32)
    got = ev(lib, "fmod", x, y)
    assert np.array_equal(got, np.fmod(x, y))


def test_round_half_away(lib):
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, 2.4999998, 1e8, -3.5, 0.49999997], np.float32)
    assert np.array_equal(ev(lib, "round", x), np.array([1, 2, 3, -1, -2, 2, 1e8, -4, 0], np.float32))


def test_special_values(lib):
    assert ev(lib, "exp", [0.0])[0] == 1.0
    assert ev(lib, "exp", [-200.0])[0] == 0.0
    assert np.isinf(ev(lib, "exp", [89.0])[0])
    assert ev(lib, "log", [1.0])[0] == 0.0
    assert np.isinf(ev(lib, "log", [0.0])[0])
    assert np.isnan(ev(lib, "log", [-1.0])[0])
    assert ev(lib, "sin", [0.0])[0] == 0.0 and ev(lib, "cos", [0.0])[0] == 1.0
    assert ev(lib, "acos", [1.0])[0] == 0.0
    assert np.isnan(ev(lib, "acos", [1.5])[0])
    assert ev(lib, "tanh", [20.0])[0] == 1.0
    a = ev(lib, "atan2", [0.0, -0.0, 0.0, 1.0], [-1.0, -1.0, 1.0, 0.0])
    assert a[0] == np.float32(np.pi) and a[1] == -np.float32(np.pi) and a[2] == 0.0
    assert a[3] == np.float32(np.pi / 2)


def test_sqrt_div_correctly_rounded(lib):
    x = RNG.uniform(0, 1e6, 100000).astype(np.float32)
    y = RNG.uniform(0.1, 100, 100000).astype(np.float32)
    assert np.array_equal(ev(lib, "sqrt", x), np.sqrt(x))
    assert np.array_equal(ev(lib, "div", x, y), x / y)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FN))
def test_device_bit_identical_to_host(lib, gctx, name):
    rng = np.random.default_rng(7)
    x = rng.uniform(-50, 50, 200000).astype(np.float32)
    y = rng.uniform(-5, 5, 200000).astype(np.float32)
    if name in ("log", "sqrt", "pow"):
        x = np.abs(x) * 37.0
    if name in ("asin", "acos"):
        x = rng.uniform(-1, 1, 200000).astype(np.float32)
    two = name in ("pow", "atan2", "fmod", "div")
    h = ev(lib, name, x, y if two else None)
    d = ev(lib, name, x, y if two else None, on_device=1, ctx=gctx)
    same = (h.view(np.uint32) == d.view(np.uint32)) | (np.isnan(h) & np.isnan(d))
    assert same.all(), (name, int((~same).sum()))


def test_sin_cos_of_signed_zero_are_exact(lib):
    # rmat3_euler skips the transcendentals for a zero angle on this basis
    import ctypes as C
    from ramp_amd import _capi
    x = np.array([0.0, -0.0], np.float32)
    y = np.zeros(2, np.float32)
    for fn, want in ((_capi.RAMP_FN_SIN, 0), (_capi.RAMP_FN_COS, 0x3F800000)):
        o = np.zeros(2, np.float32)
        _capi.check(lib.ramp_math_eval(None, fn, x.ctypes.data_as(C.POINTER(C.c_float)),
                                       y.ctypes.data_as(C.POINTER(C.c_float)),
                                       o.ctypes.data_as(C.POINTER(C.c_float)), 2, 0))
        assert (o.view(np.uint32) == want).all()


def _math_d(oracle, fn, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    oracle.load().oracle_math_d(fn, x.ctypes.data_as(C.POINTER(C.c_double)),
                                out.ctypes.data_as(C.POINTER(C.c_double)), x.size)
    return out


def _ulp64(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_fp64_exp_and_cos_accuracy(oracle):
    # rd_exp and rd_cos (rmath.h): the reference's double-precision sites
    # (detailed balance, Box-Muller); within 2 ulp of numpy's float64
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-700, 700, 100000), rng.uniform(-2, 2, 100000),
                        rng.uniform(-1e-6, 1e-6, 1000)])
    e = _ulp64(_math_d(oracle, 0, x), np.exp(x))
    assert e.max() <= 2.0, (e.max(), x[np.argmax(e)])
    # Box-Muller's argument 2 pi u, u in [0, 1]; and a wider range
    u = rng.uniform(0, 1, 200000)
    a = np.concatenate([2.0 * np.pi * u, rng.uniform(-1e3, 1e3, 100000)])
    got, ref = _math_d(oracle, 1, a), np.cos(a)
    # absolute error near the zeros of cos, relative elsewhere
    err = np.abs(got - ref) / np.maximum(np.spacing(np.abs(ref)), np.spacing(1.0) * 0.5)
    assert err.max() <= 2.0, (err.max(), a[np.argmax(err)])
    sp = _math_d(oracle, 0, np.array([0.0, 1.0, -np.inf, np.inf, np.nan, 710.0, -746.0]))
    assert sp[0] == 1.0 and abs(sp[1] - np.e) <= np.spacing(np.e) and sp[2] == 0.0 and np.isinf(sp[3]) and np.isnan(sp[4])
    assert np.isinf(sp[5]) and sp[6] == 0.0
    cp = _math_d(oracle, 1, np.array([0.0, np.inf, np.nan, np.pi]))
    assert cp[0] == 1.0 and np.isnan(cp[1]) and np.isnan(cp[2]) and cp[3] == -1.0
