"""Pins of the oracle functions that no earlier closed-form test covered
(CPU only).

Each test fixes a quantity independently of the restated kernel: the
published formula or special values of the physics (Maxwellian energy
weights, the McStas supermirror reflectivity, Bragg's law, Beer-Lambert),
the binning arithmetic the reference documents (round for 2D detectors,
floor for the tube bank), exact geometric classifications (a beamstop's
shadow, a mesh cube against the analytic box), and the reference's own data
file for the moderator spectrum parser.  A dropped term, a swapped index or
a wrong branch in the restatement moves these numbers."""
import json
import math
import os

import numpy as np
import pytest

import ramp_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

K_B = 1.38066e-23        # J/K, consts.h:4 (the reference's k_B)
MEV2J = 1.6021765e-22    # J/meV, cont_mod.cl:38
VS2E = 5.22703725e-6     # meV s^2/m^2, consts.h
SE2V = 437.393377        # m/s / sqrt(meV), consts.h
V2K = 1.58825361e-3      # 1/A per m/s, consts.h
K2V = 629.622368


def _inst(tmp_path, comps, name="t.json", **kw):
    d = {"linear": True}
    d.update(comps)
    p = tmp_path / name
    p.write_text(json.dumps({"b": d}))
    return ramp_amd.Instrument(str(p), **kw)


def _parallel(w, h, lmin, lmax=None, pos=(0.0, 0.0, 0.0)):
    lmax = lmin * (1 + 1e-7) if lmax is None else lmax
    return {"source": True, "position": list(pos),
            "moderator_kernel": {"name": "MParallel", "width": w, "height": h,
                                 "lambda_min": lmin, "lambda_max": lmax}}


def _counter(z, restore=False):
    c = {"position": [0, 0, z], "geom_kernel": {"name": "GPlane", "width": 10.0, "height": 10.0},
         "scat_kernel": {"name": "SCounter"}}
    if restore:
        c["restore_neutron"] = True
    return c


def _energy(n16):
    v2 = (n16[:, 3:6].astype(np.float64) ** 2).sum(axis=1)
    return VS2E * v2


# ---------------------------------------------------------------- beamstop

def test_beamstop_kills_exactly_its_shadow(tmp_path, oracle):
    # scat/beamstop.cl:3-32: a neutron whose geometry hit the stop dies
    # where it stands; with restore_neutron the misses pass.  In a parallel
    # beam along +z the crossing point is the neutron's (x, y).
    w, h = 0.02, 0.01
    inst = _inst(tmp_path, {
        "src": _parallel(0.04, 0.04, 3.0),
        "stop": {"position": [0, 0, 1.0], "restore_neutron": True,
                 "geom_kernel": {"name": "GPlane", "width": w, "height": h},
                 "scat_kernel": {"name": "SBeamstop"}},
        "cnt": _counter(2.0)})
    N = 50000
    res = oracle.run(inst, N, [5], keep_state=True)
    n16 = res["state"][0]
    dead = n16[:, 15] > 0
    inside = (np.abs(n16[:, 0]) < w / 2) & (np.abs(n16[:, 1]) < h / 2)
    assert np.array_equal(dead, inside)
    assert np.all(n16[dead, 12] == 2.0)            # prevcomp: the stop (comp 2)
    assert abs(dead.mean() - (w * h) / (0.04 * 0.04)) < 4 * math.sqrt(0.125 * 0.875 / N)
    # the counter sees every survivor at weight 1/N
    assert res["hist"][0][0].sum() == pytest.approx((~dead).sum() / N, rel=1e-5)


# ---------------------------------------------------------------- MCont

def _mcont(tmp_path, E0, E1, T, I, join=0.0, name="m.json"):
    src = {"source": True, "moderator_kernel": {
        "name": "MCont", "mod_dim": [0.01, 0.01], "target_dim": [0.01, 0.01], "target_dist": 1.0,
        "E_min": E0, "E_max": E1, "T1": T[0], "I1": I[0], "T2": T[1], "I2": I[1], "T3": T[2], "I3": I[2]}}
    if join:
        src["moderator_kernel"]["tail_join"] = join
    return _inst(tmp_path, {"src": src, "cnt": _counter(0.5, restore=True)}, name=name)


def _maxwell_ratio(E_meV, T):
    # M(E;T)/M(kT;T) of cont_mod.cl:4-9,37-47: M ~ sqrt(E) e^{-E/kT}, so the
    # ratio is sqrt(E/kT) e^{1 - E/kT}, energies in joules
    x = (MEV2J * E_meV) / (K_B * T)
    return np.sqrt(x) * np.exp(1.0 - x)


def test_mcont_weight_is_the_maxwellian_sum(tmp_path, oracle):
    T, I = (40.0, 120.0, 300.0), (1.0, 0.5, 0.25)
    inst = _mcont(tmp_path, 1.0, 60.0, T, I)
    N = 20000
    n16 = oracle.run(inst, N, [3], keep_state=True)["state"][0]
    E = _energy(n16)
    assert E.min() >= 0.999 and E.max() <= 60.01
    expect = sum(i * _maxwell_ratio(E, t) for t, i in zip(T, I)) / N
    np.testing.assert_allclose(n16[:, 9], expect, rtol=2e-5)


def test_mcont_single_term_and_unit_peak(tmp_path, oracle):
    # one Maxwellian: the weight peaks at E = kT/2 with value
    # sqrt(1/2) e^{1/2} / N and equals exactly 1/N at E = kT
    T = 40.0
    kT = K_B * T / MEV2J
    inst = _mcont(tmp_path, 0.5 * kT, 0.5 * kT * (1 + 1e-6), (T, T, T), (1.0, 0.0, 0.0))
    n16 = oracle.run(inst, 1000, [4], keep_state=True)["state"][0]
    np.testing.assert_allclose(n16[:, 9] * 1000, math.sqrt(0.5) * math.exp(0.5), rtol=1e-5)
    inst = _mcont(tmp_path, kT, kT * (1 + 1e-6), (T, T, T), (1.0, 0.0, 0.0), name="m2.json")
    n16 = oracle.run(inst, 1000, [4], keep_state=True)["state"][0]
    np.testing.assert_allclose(n16[:, 9] * 1000, 1.0, rtol=1e-5)


def test_mcont_zero_temperature_gives_nan(tmp_path, oracle):
    # cont_mod.cl:39-45: M(kT;0) is 0/0 and the I2 = 0 term is 0*NaN
    inst = _mcont(tmp_path, 1.0, 10.0, (40.0, 0.0, 40.0), (1.0, 0.0, 0.0))
    n16 = oracle.run(inst, 1000, [5], keep_state=True)["state"][0]
    assert np.isnan(n16[:, 9]).all()


def test_mcont_epithermal_tail(tmp_path, oracle):
    # above E_j = join k_B T1 the first term is replaced by its value at E_j
    # times E_j / E: the weight falls as 1/E and is continuous at E_j
    T, join = 40.0, 2.0
    Ej = join * K_B * T / MEV2J
    inst = _mcont(tmp_path, 1.1, 30.0, (T, T, T), (1.0, 0.0, 0.0), join=join)
    N = 40000
    n16 = oracle.run(inst, N, [6], keep_state=True)["state"][0]
    E = _energy(n16)
    w = n16[:, 9].astype(np.float64) * N
    lo, hi = E < Ej * (1 - 1e-5), E > Ej * (1 + 1e-5)
    assert lo.sum() > 1000 and hi.sum() > 1000
    np.testing.assert_allclose(w[lo], _maxwell_ratio(E[lo], T), rtol=2e-5)
    r_j = math.sqrt(join) * math.exp(1.0 - join)
    np.testing.assert_allclose(w[hi] * E[hi], r_j * Ej, rtol=2e-5)


# ---------------------------------------------------------------- mesh

def _cube_triangles(s):
    """12 outward-facing triangles (counter-clockwise seen from outside) of
    the cube [-s/2, s/2]^3."""
    h = s / 2
    v = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)])
    idx = lambda x, y, z: (x > 0) * 4 + (y > 0) * 2 + (z > 0)  # noqa: E731
    faces = []
    for axis in range(3):
        for sgn in (-1, 1):
            corners = []
            for a in (-1, 1):
                for b in (-1, 1):
                    c = [0, 0, 0]
                    c[axis] = sgn
                    o = [k for k in range(3) if k != axis]
                    c[o[0]], c[o[1]] = a, b
                    corners.append(idx(*c))
            p00, p01, p10, p11 = corners
            quad = [p00, p10, p11, p01]
            n = np.zeros(3)
            n[axis] = sgn
            t1, t2 = [quad[0], quad[1], quad[2]], [quad[0], quad[2], quad[3]]
            for t in (t1, t2):
                a, b, c = v[t[0]], v[t[1]], v[t[2]]
                if np.dot(np.cross(b - a, c - a), n) < 0:
                    t = [t[0], t[2], t[1]]
                faces.append(np.concatenate([v[k] for k in t]))
    return np.array(faces, dtype=np.float32)


def test_mesh_cube_equals_analytic_box(tmp_path, oracle):
    # mesh.cl:52-118 brute force on a closed cube: the path between the
    # nearest entering and exiting faces is the box chord, so Beer-Lambert
    # attenuation equals GBox's, neutron by neutron
    from ramp_amd.components import write_stl_binary
    s, mu = 0.02, 1.3
    tri = _cube_triangles(s)
    assert tri.shape == (12, 9)
    stl = tmp_path / "cube.stl"
    write_stl_binary(str(stl), tri)
    beam = _parallel(0.03, 0.03, 2.5)
    common = {"restore_neutron": True, "position": [0.0, 0.0, 1.0], "rotation": [0.3, 0.5, 0.2]}
    mesh = _inst(tmp_path, {"src": beam,
                            "obj": dict(common, geom_kernel={"name": "GMesh", "filename": str(stl)},
                                        scat_kernel={"name": "SAbsorber", "sigma_a": mu}),
                            "cnt": _counter(2.0)}, name="mesh.json")
    box = _inst(tmp_path, {"src": beam,
                           "obj": dict(common, geom_kernel={"name": "GBox", "width": s, "height": s, "depth": s},
                                       scat_kernel={"name": "SAbsorber", "sigma_a": mu}),
                           "cnt": _counter(2.0)}, name="box.json")
    N = 40000
    a = oracle.run(mesh, N, [7], keep_state=True)["state"][0]
    b = oracle.run(box, N, [7], keep_state=True)["state"][0]
    hit = b[:, 9] < 1.0 / N * (1 - 1e-6)
    assert hit.sum() > 0.2 * N
    np.testing.assert_array_equal(a[:, 9] < 1.0 / N * (1 - 1e-6), hit)
    # the two chords differ by float rounding of positions near z = 1 m
    # (about 1e-7 m), i.e. mu * 1e-7 m in the exponent
    np.testing.assert_allclose(a[:, 9], b[:, 9], rtol=1e-4)
    assert abs(a[:, 9].sum() / b[:, 9].sum() - 1) < 1e-6


# ---------------------------------------------------------------- detectors

def _pencil(lam=2.0):
    # a source emits from the instrument origin (its component position is
    # not applied); the detectors below are offset instead
    return _parallel(1e-7, 1e-7, lam)


def test_detector2d_round_binning_and_alias(tmp_path, oracle):
    # detector2D.cl:17-24, 169: idx = round((v - min)/step) on each axis,
    # flat = a1 * n2 + a2 with n2 = ceil((max - min)/step); v = max gives
    # a2 = n2, which lands in row a1 + 1, column 0
    b1, b2 = [-0.05, 0.01, 0.05], [-0.02, 0.005, 0.02]
    n1 = int(math.ceil((b1[2] - b1[0]) / b1[1]))
    n2 = int(math.ceil((b2[2] - b2[0]) / b2[1]))
    cases = [(0.0137, -0.0093), (-0.0412, 0.0149), (0.0231, 0.02)]   # the last sits at y = max
    for k, (x0, y0) in enumerate(cases):
        inst = _inst(tmp_path, {
            "src": _pencil(),
            "psd": {"position": [-x0, -y0, 1.0], "geom_kernel": {"name": "GPlane", "width": 0.2, "height": 0.2},
                    "scat_kernel": {"name": "SDetector2D", "axis1_var": "x", "axis1_binning": b1,
                                    "axis2_var": "y", "axis2_binning": b2}}}, name=f"d{k}.json")
        w = oracle.run(inst, 100, [1])["hist"][0][0]
        assert w.size == n1 * n2
        a1 = int(round((x0 - b1[0]) / b1[1]))
        a2 = int(round((y0 - b2[0]) / b2[1]))
        flat = a1 * n2 + a2
        assert np.flatnonzero(w).tolist() == [flat] and w[flat] > 0
    assert a2 == n2   # the alias case was exercised


def test_detector_tubes_floor_binning(tmp_path, oracle):
    # SDetectorTubes (no reference counterpart): tube = floor((theta - th0) /
    # (th1 - th0) * n_tubes), pixel and TOF bin likewise, flat index
    # (tube * n_pix + pixel) * n_tof + tof_bin.  A pencil beam through the
    # axis of a banana turned by beta about y hits it at theta = -beta.
    th0, th1, nt = -40.0, 140.0, 72
    y0r, y1r, npx = -0.2, 0.2, 16
    t0, t1, nf = 0.0, 3000.0, 30
    lam = 2.0
    v = 2 * math.pi / (lam * V2K)
    R = 1.5   # the source sits 1 m from the axis, inside the bank
    for k, (beta, y0) in enumerate([(-37.3, 0.0373), (-101.1, -0.1234), (12.6, 0.013), (45.0, 0.1)]):
        inst = _inst(tmp_path, {
            "src": _pencil(lam),
            "bank": {"position": [0, -y0, 1.0], "rotation": [0.0, math.radians(beta), 0.0],
                     "geom_kernel": {"name": "GBanana", "radius": R, "height": 0.5, "mintheta": -180.0,
                                     "maxtheta": 180.0},
                     "scat_kernel": {"name": "SDetectorTubes", "theta_range": [th0, th1], "y_range": [y0r, y1r],
                                     "tof_range": [t0, t1], "n_tubes": nt, "n_pixels": npx, "n_tof": nf}}},
            name=f"t{k}.json")
        res = oracle.run(inst, 100, [1])
        w = res["hist"][0][0]
        theta = -beta
        tof_us = 1e6 * (1.0 + R) / v
        it = math.floor((theta - th0) / (th1 - th0) * nt)
        ip = math.floor((y0 - y0r) / (y1r - y0r) * npx)
        iff = math.floor((tof_us - t0) / (t1 - t0) * nf)
        if 0 <= it < nt:
            assert np.flatnonzero(w).tolist() == [(it * npx + ip) * nf + iff] and w.sum() > 0
        else:
            assert w.sum() == 0.0     # outside the tube range: dropped


# ---------------------------------------------------------------- guide

def _mcstas_reflectivity(q, R0, Qc, alpha, m, W):
    """The McStas supermirror reflectivity (McStas component manual, Guide;
    ref.h:3-54 restates it): R0 up to Qc, then
    R0/2 (1 - tanh((q - m Qc)/W)) (1 - alpha (q - Qc) + beta (q - Qc)^2),
    zero when (q - m Qc)/W > 10; with W = alpha = 0 the m-only
    parametrisation (m' = 0.9853 m + 0.1978, W = 0.0022 - 0.0002 m',
    alpha = 0.2304 m' + 5.0944, beta = 68.1137 - 7.6251 m', and for m' <= 3
    alpha = m', beta = 0).  m < 1 scales Qc by m."""
    beta = 0.0
    if W == 0.0 and alpha == 0.0:
        m = 0.9853 * m + 0.1978
        W = 0.0022 - 0.0002 * m
        alpha, beta = (m, 0.0) if m <= 3 else (0.2304 * m + 5.0944, 68.1137 - 7.6251 * m)
    if W <= 0 or (q - m * Qc) / W > 10 or m <= 0 or Qc <= 0 or R0 <= 0:
        return 0.0
    if m < 1:
        Qc, m = Qc * m, 1.0
    if q <= Qc:
        return R0
    R = R0 * 0.5 * (1 - math.tanh((q - m * Qc) / W)) * (1 - alpha * (q - Qc) + beta * (q - Qc) ** 2)
    return R if 1e-12 <= R <= 1.0 else 0.0


@pytest.mark.parametrize("params", [
    dict(R0=0.99, Qc=0.0219, alpha=6.07, m=2.0, W=0.003),      # the LET supermirror
    dict(R0=0.95, Qc=0.0217, alpha=0.0, m=2.5, W=0.0),         # m-only parametrisation, m' <= 3
    dict(R0=0.98, Qc=0.0217, alpha=0.0, m=3.5, W=0.0),         # m-only, m' > 3 (beta term)
    dict(R0=0.9, Qc=0.03, alpha=2.0, m=0.6, W=0.004),          # m < 1 rescales Qc
])
def test_guide_single_bounce_reflectivity(tmp_path, oracle, params):
    # A narrow parallel beam enters a straight guide turned by theta about y,
    # so every neutron meets one side wall once at grazing angle theta:
    # q = 4 pi sin(theta) / lambda, and it leaves with weight R(q) / N
    lam, w = 4.0, 0.02
    qs = [0.5, 0.98, 1.3, 1.9, 2.2, 2.6, 3.5, 6.0]
    for k, qr in enumerate(qs):
        q = qr * params["Qc"]
        theta = math.asin(q * lam / (4 * math.pi))
        l = 2 * (w / 2) / math.tan(theta)
        inst = _inst(tmp_path, {
            "src": _parallel(0.001, 0.001, lam),
            "g": {"position": [0, 0, 0.1], "rotation": [0.0, theta, 0.0],
                  "geom_kernel": {"name": "GPlane", "width": w, "height": 0.05},
                  "scat_kernel": dict({"name": "SGuide", "w1": w, "h1": 0.05, "w2": w, "h2": 0.05, "l": l},
                                      **params)},
            "cnt": _counter(0.2 + l * 1.2)}, name=f"g{k}.json")
        N = 200
        res = oracle.run(inst, N, [k], keep_state=True)
        n16 = res["state"][0]
        expect = _mcstas_reflectivity(q, **params)
        # float32 arithmetic: 1 - tanh(arg) carries an absolute error of an
        # ulp of 1 (6e-8), which is all that is left of R near the cut-off
        np.testing.assert_allclose(n16[:, 9] * N, expect, rtol=2e-4, atol=1e-7)


# ---------------------------------------------------------------- powder

def test_powdern_lines_at_bragg_angles_with_sigma_ratios(tmp_path, oracle):
    # powderN.cl:77-103 chooses line i with probability sigma_i / sum sigma
    # and scatters it to 2 theta_i = 2 asin(lambda / 2 d_i); the weight
    # factor does not depend on the line.  powderN.py:59-64:
    # sigma_i ~ mult_i |F_i|^2 / q_i = mult_i |F_i|^2 d_i / (2 pi)
    lam = 1.8
    E = 81.80420 / lam ** 2
    d = [1.25, 1.8, 2.5, 3.1]
    F2 = [40.0, 25.0, 90.0, 10.0]
    mult = [6, 12, 2, 8]
    inst = _inst(tmp_path, {
        "src": {"source": True, "moderator_kernel": {
            "name": "MFlat", "mod_dim": [0.0005, 0.0005], "target_dim": [0.0005, 0.0005],
            "target_dist": 1.0, "E_min": E, "E_max": E}},
        "sample": {"position": [0, 0, 1.0], "restore_neutron": True,
                   "geom_kernel": {"name": "GCylinderExt", "radius": 0.0005, "height": 0.02},
                   "scat_kernel": {"name": "SPowderN", "d_spacing": d, "pack": 1.0, "vc": 120.0,
                                   "sigma_abs": 0.5, "multiplicity": mult, "DW": 1.0, "F2": F2,
                                   "d_phi": 2.0}},
        "det": {"position": [0, 0, 1.0],
                "geom_kernel": {"name": "GBanana", "radius": 1.0, "height": 0.5,
                                "mintheta": -40.0, "maxtheta": 140.0},
                "scat_kernel": {"name": "SDetector1D", "var": "theta", "binning": [-180.0, 0.05, 180.0]}}})
    N = 400000
    res = oracle.run(inst, N, [9], keep_state=True)
    w = res["hist"][0][0]
    axis = np.abs(-180.0 + 0.05 * (np.arange(w.size) + 0.5))
    sig = np.array([m_ * f * dd for m_, f, dd in zip(mult, F2, d)])
    frac = sig / sig.sum()
    # the transmitted beam (restore_neutron on the sample) arrives at 0 deg
    scat = axis > 5.0
    tot = w[scat].sum()
    got = []
    for dd in d:
        tt = math.degrees(2 * math.asin(lam / (2 * dd)))
        sel = np.abs(axis - tt) < 0.6
        got.append(w[sel].sum() / tot)
        peak = np.sum(axis[sel] * w[sel]) / w[sel].sum()
        assert abs(peak - tt) < 0.05
    got = np.array(got)
    assert got.sum() == pytest.approx(1.0, abs=1e-6)      # nothing off the four cones
    # multinomial scatter of the line choice: n_eff scattered histories of
    # (nearly) equal weight
    n_eff = tot ** 2 / res["hist"][0][1][scat].sum()
    assert n_eff > 2000
    assert np.all(np.abs(got - frac) < 4 * np.sqrt(frac * (1 - frac) / n_eff) + 1e-4)


# ---------------------------------------------------------------- MISIS

LET_CUT = os.path.join(ROOT, "tests", "golden", "let_base_cut.mcstas.gz")


@pytest.fixture(scope="module")
def let_cut(tmp_path_factory):
    import gzip
    p = tmp_path_factory.mktemp("spec") / "let_cut.mcstas"
    p.write_text(gzip.open(LET_CUT, "rt").read())
    return str(p)


def test_tally_reader_on_the_reference_spectrum(let_cut):
    # the blocks of examples/Let_Base.mcstas: each block's rows sum to the
    # file's own 'total' line, the edges are its 'energy bin' lines (MeV ->
    # meV), the time offset its 'TimeOffset' comment
    import re
    from ramp_amd.components import read_tally_spectrum
    text = open(let_cut).read()
    spec = read_tally_spectrum(let_cut)
    edges = [(float(a), float(b)) for a, b in re.findall(r"energy bin:\s*(\S+)\s+to\s+(\S+)", text)]
    totals = [float(t) for t in re.findall(r"^\s*total\s+(\S+)", text, re.M)]
    assert len(edges) == 10 and spec["values"].shape == (10, 352)
    np.testing.assert_array_equal(spec["e_lower"], [a * 1e9 for a, _ in edges])
    np.testing.assert_array_equal(spec["e_upper"], [b * 1e9 for _, b in edges])
    np.testing.assert_allclose(spec["values"].sum(axis=1), totals, rtol=2e-5)
    np.testing.assert_array_equal(spec["totals"], totals)
    assert spec["time_offset"] == 157.0
    assert np.all(np.diff(spec["time"]) > 0) and spec["time"][0] == 1e-8


def test_misis_window_weights_and_total(let_cut):
    # the kept total is the overlap-weighted sum of the blocks' own totals
    from ramp_amd.components import MISIS, read_tally_spectrum
    spec = read_tally_spectrum(let_cut)
    for E0, E1 in [(2.3, 5.0), (2.1544, 6.0), (3.0, 3.3)]:
        src = MISIS(spec_file=let_cut, mod_dim=(0.04, 0.09), target_dim=(0.04, 0.09), target_dist=1.7,
                    E_min=E0, E_max=E1)
        lo, hi = spec["e_lower"], spec["e_upper"]
        e0, e1 = float(np.float32(E0)), float(np.float32(E1))
        overlap = np.clip(np.minimum(hi, e1) - np.maximum(lo, e0), 0, None) / (hi - lo)
        expect = float((overlap * spec["values"].sum(axis=1)).sum())
        assert float(src.total) == pytest.approx(expect, rel=1e-6)
        assert src.e_bins[0] == np.float32(E0) and src.e_bins[-1] == np.float32(E1)
        inner = spec["e_lower"][(spec["e_lower"] > e0) & (spec["e_lower"] < e1)]
        np.testing.assert_array_equal(src.e_bins[1:-1], inner.astype(np.float32))
        assert src.e_int[-1] == pytest.approx(float(src.total), rel=1e-6)


def test_misis_solid_angle_factor_far_field():
    from ramp_amd.components import _target_solid_angle_factor
    # sizes much smaller than the distance: A_mod * A_target[cm^2] / d^2
    a = _target_solid_angle_factor((0.01, 0.02), (0.03, 0.04), 100.0)
    assert float(a) == pytest.approx(0.01 * 0.02 * 0.03 * 0.04 * 1e4 / 100.0 ** 2, rel=1e-6)


def test_misis_sampled_time_and_energy_follow_the_table(tmp_path, oracle, let_cut):
    # isis_mod.cl:35-74: a deviate u * total picks the time row i of the
    # cumulative flux and the time is interpolated inside [t_i, t_{i+1}), so
    # P(t in [t_i, t_{i+1})) is the (window-weighted) tally value of row i+1
    # summed over the kept blocks; the energy lands in block k with
    # probability (E_int[k+1] - E_int[k]) / total; every weight is
    # str_area * total * 6.2415093e12 / N
    from ramp_amd.components import MISIS, read_tally_spectrum
    E0, E1 = 2.3, 5.0
    src = {"source": True, "moderator_kernel": {
        "name": "MISIS", "spec_file": let_cut, "mod_dim": [0.04, 0.09], "target_dim": [0.04, 0.09],
        "target_dist": 1.7, "E_min": E0, "E_max": E1}}
    inst = _inst(tmp_path, {"src": src, "cnt": _counter(0.5, restore=True)})
    N = 200000
    n16 = oracle.run(inst, N, [8], keep_state=True)["state"][0]
    m = MISIS(spec_file=let_cut, mod_dim=(0.04, 0.09), target_dim=(0.04, 0.09), target_dist=1.7,
              E_min=E0, E_max=E1)
    np.testing.assert_allclose(n16[:, 9], np.float32(m.str_area * m.total * np.float32(6.2415093e12) / N),
                               rtol=1e-6)
    spec = read_tally_spectrum(let_cut)
    e0, e1 = float(np.float32(E0)), float(np.float32(E1))
    lo, hi = spec["e_lower"], spec["e_upper"]
    overlap = np.clip(np.minimum(hi, e1) - np.maximum(lo, e0), 0, None) / (hi - lo)
    row_w = (overlap[:, None] * spec["values"]).sum(axis=0)          # per time row
    t = spec["time"].astype(np.float32)
    p_t = row_w[1:] / row_w[1:].sum()                                 # interval i <- row i + 1
    tof = n16[:, 10]
    assert tof.min() >= t[0] and tof.max() <= t[-1]
    counts = np.histogram(tof, bins=t)[0]
    big = p_t * N > 50
    chi2 = float((((counts[big] - N * p_t[big]) ** 2) / (N * p_t[big])).sum())
    dof = int(big.sum())
    assert chi2 < dof + 5 * math.sqrt(2 * dof)
    # energy blocks: counts between the kept edges
    E = _energy(n16)
    edges = m.e_bins.astype(np.float64)
    pk = np.diff(m.e_int.astype(np.float64)) / float(m.total)
    ck = np.histogram(E, bins=edges)[0]
    assert ck.sum() >= 0.999 * N
    assert np.all(np.abs(ck - N * pk) < 5 * np.sqrt(N * pk * (1 - pk)) + 0.002 * N)


# ---------------------------------------------------------------- SIsotropic

def _write_sqw(path, q, w, S, header, first_row=None):
    with open(path, "w") as f:
        for k, v in header.items():
            f.write(f"# {k} {v}\n")
        f.write(" ".join(repr(float(x)) for x in q) + "\n")
        f.write(" ".join(repr(float(x)) for x in w) + "\n")
        row0 = first_row if first_row is not None else [9e9] * len(w)
        f.write(" ".join(repr(float(x)) for x in row0) + "\n")
        for r in S:
            f.write(" ".join(repr(float(x)) for x in r) + "\n")


def test_sqw_reader_keeps_what_the_reference_loader_keeps(tmp_path):
    # isotropic.py:171-196: '# key value' parameters; the q line loses its
    # last value; the first S row is consumed; S gains float64 epsilon
    from ramp_amd.components import read_sqw
    p = tmp_path / "s.sqw"
    S = np.arange(12, dtype=float).reshape(4, 3) + 1
    _write_sqw(p, [0.5, 1.0, 1.5, 2.0, 2.5], [-1.0, 0.0, 1.0], S,
               {"V_rho": 0.07, "sigma_coh": 2.5, "sigma_abs": 0.3, "T": "300 K"})
    hdr, q, w, sqw = read_sqw(str(p))
    assert hdr["V_rho"] == "0.07" and hdr["sigma_coh"] == "2.5" and hdr["sigma_abs"] == "0.3" and hdr["T"] == "300"
    np.testing.assert_array_equal(q, np.float32([0.5, 1.0, 1.5, 2.0]))
    np.testing.assert_array_equal(w, np.float32([-1.0, 0.0, 1.0]))
    assert sqw.dtype == np.float32 and sqw.shape == (4, 3)
    np.testing.assert_array_equal(sqw, S.astype(np.float32) + np.float32(2.220446049250313e-16))


def test_legacy_simpson_rule():
    from scipy.integrate import simpson, trapezoid
    from ramp_amd.components import _simps_legacy
    # odd count, equal spacing: exact on cubics
    x = np.linspace(0.3, 2.1, 9)
    y = 2 * x ** 3 - x ** 2 + 0.5 * x - 3
    F = lambda t: 0.5 * t ** 4 - t ** 3 / 3 + 0.25 * t ** 2 - 3 * t  # noqa: E731
    assert _simps_legacy(y, x) == pytest.approx(F(2.1) - F(0.3), rel=1e-13)
    # odd count, unequal spacing: SciPy's simpson (same rule for odd counts), exact on quadratics
    xn = np.sort(np.random.default_rng(1).uniform(0, 3, 11))
    yn = np.cos(xn) + xn
    assert _simps_legacy(yn, xn) == pytest.approx(simpson(yn, x=xn), rel=1e-12)
    assert _simps_legacy(xn ** 2, xn) == pytest.approx((xn[-1] ** 3 - xn[0] ** 3) / 3, rel=1e-12)
    # even count: the mean of the two Simpson + end-trapezoid compositions
    xe = np.sort(np.random.default_rng(2).uniform(0, 3, 10))
    ye = np.sin(xe) + 2
    first = simpson(ye[:-1], x=xe[:-1]) + trapezoid(ye[-2:], x=xe[-2:])
    last = simpson(ye[1:], x=xe[1:]) + trapezoid(ye[:2], x=xe[:2])
    assert _simps_legacy(ye, xe) == pytest.approx(0.5 * (first + last), rel=1e-12)
    assert _simps_legacy(3 * xe + 1, xe) == pytest.approx(1.5 * (xe[-1] ** 2 - xe[0] ** 2) + xe[-1] - xe[0], rel=1e-12)
    # float32 in, float32 out; either axis
    Y = np.random.default_rng(3).random((4, 9)).astype(np.float32)
    X = np.linspace(0, 1, 9).astype(np.float32)
    assert _simps_legacy(Y, X, axis=1).dtype == np.float32
    np.testing.assert_array_equal(_simps_legacy(Y.T, X, axis=0), _simps_legacy(Y, X, axis=1))


def test_sqw_tables_of_a_separable_s():
    # S(q, w) = f(q) g(w): p(q | w) = q g(w) / int g, whose running CDF is
    # (q_i^2 - q_0^2) / (q_n^2 - q_0^2) at every w (the rule is exact on
    # linear integrands); with g linear in w the w CDF is exact too
    from ramp_amd.components import _sqw_distributions
    q = np.linspace(0.2, 4.0, 21).astype(np.float32)
    w = np.linspace(-2.0, 2.0, 9).astype(np.float32)
    f = np.exp(-0.3 * q.astype(np.float64) ** 2)
    g = 1.5 + 0.25 * w.astype(np.float64)
    S = (f[:, None] * g[None, :]).astype(np.float32)
    pw_cdf, pq_cdf = _sqw_distributions(q, w, S)
    qd = q.astype(np.float64)
    expect_q = (qd ** 2 - qd[0] ** 2) / (qd[-1] ** 2 - qd[0] ** 2)
    for j in range(len(w)):
        np.testing.assert_allclose(pq_cdf[j], expect_q, rtol=2e-5, atol=2e-7)
    # p(w) is g(w) * int q f dq, normalised: CDF of a linear density
    G = lambda t: 1.5 * t + 0.125 * t * t  # noqa: E731
    wd = w.astype(np.float64)
    expect_w = (G(wd) - G(wd[0])) / (G(wd[-1]) - G(wd[0]))
    np.testing.assert_allclose(pw_cdf, expect_w, rtol=2e-5, atol=2e-7)


def _isotropic_instrument(tmp_path, sqw_path, E, T, name="iso.json"):
    return _inst(tmp_path, {
        "src": {"source": True, "moderator_kernel": {
            "name": "MFlat", "mod_dim": [1e-6, 1e-6], "target_dim": [1e-6, 1e-6],
            "target_dist": 1.0, "E_min": E, "E_max": E}},
        "sample": {"position": [0, 0, 1.0], "restore_neutron": True,
                   "geom_kernel": {"name": "GCylinderExt", "radius": 0.005, "height": 0.02},
                   "scat_kernel": {"name": "SIsotropic", "fn_inc": sqw_path, "temperature": T}}},
        name=name)


def test_isotropic_kinematics_and_q_sampling(tmp_path, oracle):
    # isotropic.cl:5-39, 226-291: w and Q are taken by the nearest-CDF rule
    # (the first entry of 1.0 ends the scan) and drawn uniformly inside the
    # interval below the chosen grid point (from 0 for the first); then
    # kf^2 = ki^2 -/+ E2KS w, and the direction turns by theta with
    # Q^2 = ki^2 + kf^2 - 2 ki kf cos(theta), the Q stored in the record
    q = np.linspace(0.2, 4.0, 21)
    w = np.array([-1e-4, 0.0, 1e-4])
    S = np.exp(-0.3 * q[:-1, None] ** 2) * np.ones((1, 3))
    p = tmp_path / "el.sqw"
    _write_sqw(p, q, w, S, {"V_rho": 0.07, "sigma_coh": 20.0, "sigma_abs": 0.0})
    E = 20.0
    ki = V2K * SE2V * math.sqrt(E)
    inst = _isotropic_instrument(tmp_path, str(p), E, 300.0)
    N = 300000
    n16 = oracle.run(inst, N, [12], keep_state=True)["state"][0]
    # scattered (not transmitted): alive after the sample with a Q recorded
    sc = (n16[:, 12] == 2.0) & (n16[:, 15] == 0) & (n16[:, 13] != 0)
    v = n16[sc, 3:6].astype(np.float64)
    Q = n16[sc, 13].astype(np.float64)
    om = n16[sc, 14].astype(np.float64)
    assert sc.sum() > 3000 and np.abs(om).max() <= 1e-4 * (1 + 1e-6)
    kf = V2K * np.linalg.norm(v, axis=1)
    E2KS = 1 / 2.072
    np.testing.assert_allclose(kf ** 2, ki ** 2 - E2KS * om, rtol=2e-6)
    cos_t = v[:, 2] / np.linalg.norm(v, axis=1)
    Q_geom = np.sqrt(np.maximum(ki ** 2 + kf ** 2 - 2 * ki * kf * cos_t, 0))
    np.testing.assert_allclose(Q_geom, Q, rtol=1e-4, atol=2e-4)
    # Q: piecewise uniform over [0, q0], [q0, q1], ... with the nearest-CDF
    # probabilities of c_i = (q_i^2 - q_0^2) / (q_n^2 - q_0^2)
    qk = q[:-1].astype(np.float32).astype(np.float64)
    c = (qk ** 2 - qk[0] ** 2) / (qk[-1] ** 2 - qk[0] ** 2)
    m = len(c) - 1                           # entries before the 1.0
    mids = (c[:m - 1] + c[1:m]) / 2
    P = np.diff(np.concatenate([[0.0], mids, [1.0]]))
    edges = np.concatenate([[0.0], qk[:m]])
    counts = np.histogram(Q, bins=edges)[0]
    n = counts.sum()
    assert n == sc.sum()
    chi2 = float(((counts - n * P) ** 2 / (n * P)).sum())
    assert chi2 < m + 5 * math.sqrt(2 * m)


def test_isotropic_detailed_balance_branch(tmp_path, oracle):
    # isotropic.cl:247: the neutron keeps kf^2 = ki^2 - E2KS w (se = w) when
    # a draw exceeds 0.5 exp(-w / (kB T)) with the reference's kB =
    # 8.6173 * 10e-2 (consts.h:36), else gains kf^2 = ki^2 + E2KS w (se = -w).
    # Here every sampled w is <= 0, so P(se = -w) = min(1, 0.5 exp(|w| / kB T)).
    q = np.linspace(0.2, 3.0, 15)
    w = np.array([-2.0, 0.0, 2.0])
    S = np.exp(-0.3 * q[:-1, None] ** 2) * np.ones((1, 3))
    p = tmp_path / "db.sqw"
    _write_sqw(p, q, w, S, {"V_rho": 0.07, "sigma_coh": 20.0, "sigma_abs": 0.0})
    T = 10.0
    kBT = 8.6173 * float(np.float32(10e-2)) * T
    inst = _isotropic_instrument(tmp_path, str(p), 25.0, T)
    N = 400000
    n16 = oracle.run(inst, N, [13], keep_state=True)["state"][0]
    sc = (n16[:, 12] == 2.0) & (n16[:, 15] == 0) & (n16[:, 13] != 0)
    se = n16[sc, 14].astype(np.float64)
    om = -np.abs(se)                          # the sampled w (all <= 0)
    gain = se > 0                              # se = -w > 0: second branch
    bins = np.linspace(-2.0, -0.2, 7)
    for lo, hi in zip(bins[:-1], bins[1:]):
        sel = (om >= lo) & (om < hi)
        p_exp = np.minimum(1.0, 0.5 * np.exp(-om[sel] / kBT)).mean()
        k = sel.sum()
        assert k > 1000
        assert abs(gain[sel].mean() - p_exp) < 5 * math.sqrt(p_exp * (1 - p_exp) / k) + 1e-3


# ---------------------------------------------------------------- monochromator

def test_monochromator_quadrature_table_is_gauss_legendre():
    # monochromator.py:60-69: the 15-point Gauss-Legendre rule on [-1, 1]
    from ramp_amd.components import _GAUSS_W, _GAUSS_X
    x, w = np.polynomial.legendre.leggauss(15)
    np.testing.assert_allclose(_GAUSS_X, x.astype(np.float32), rtol=0, atol=1e-7)
    np.testing.assert_allclose(_GAUSS_W, w.astype(np.float32), rtol=0, atol=1e-7)


def _mono_inst(tmp_path, lam, alpha, r0=0.7, mosaic=30.0, name="mono.json"):
    return _inst(tmp_path, {
        "src": _parallel(0.002, 0.002, lam),
        "mono": {"position": [0, 0, 1.0], "rotation": [0.0, alpha, 0.0],
                 "geom_kernel": {"name": "GPlane", "width": 0.1, "height": 0.1, "orientation": "yz"},
                 "scat_kernel": {"name": "SMonochromator", "slab_width": 0.1, "slab_height": 0.1, "gap": 0.0,
                                 "n_horizontal": 1, "n_vertical": 1, "mosaic_horizontal": mosaic,
                                 "mosaic_vertical": mosaic, "r0": r0, "d_spacing": 3.355}}}, name=name)


@pytest.mark.parametrize("delta_fwhm", [0.0, 0.6, 1.2])
def test_monochromator_bragg_reflection(tmp_path, oracle, delta_fwhm):
    # monochromator.cl:84-170, PG(002) d = 3.355 A, mosaic 30' FWHM.  A
    # parallel beam meets the slab at the glancing angle theta_B + delta:
    #  - a neutron reflects with probability r0 exp(-delta^2 / 2 sigma^2),
    #    sigma = 30' / sqrt(8 ln 2) (the beam lies in the horizontal plane);
    #  - a reflected neutron keeps |k| and turns by exactly 2 theta_B,
    #    theta_B = asin(lambda / 2d) (Bragg's law, first order);
    #  - its weight factor G(q) / (total G(phi)) averages the integral of
    #    the mosaic distribution over the reflection ring divided by the
    #    15-point quadrature of it, i.e. 1
    lam, d, r0 = 4.0, 3.355, 0.7
    thB = math.asin(lam / (2 * d))
    sigma = math.radians(30.0 / 60.0) / math.sqrt(8 * math.log(2))
    delta = delta_fwhm * sigma * math.sqrt(8 * math.log(2))
    inst = _mono_inst(tmp_path, lam, thB + delta, r0)
    N = 200000
    n16 = oracle.run(inst, N, [21], keep_state=True)["state"][0]
    # Box-Muller's log(u1) with u1 = 0, which the closed [0, 1] draws allow
    # (about 3e-8 per draw), leaves a NaN velocity and weight
    # (monochromator.cl:16-19); such histories are set aside
    fin = np.isfinite(n16[:, 3:6]).all(axis=1)
    assert (~fin).sum() <= 2 and np.isnan(n16[~fin, 9]).all()
    n16 = n16[fin]
    v = n16[:, 3:6].astype(np.float64)
    speed = np.linalg.norm(v, axis=1)
    ang = np.arccos(np.clip(v[:, 2] / speed, -1, 1))
    refl = ang > 1e-3
    p = r0 * math.exp(-delta ** 2 / (2 * sigma ** 2))
    assert abs(refl.mean() - p) < 5 * math.sqrt(p * (1 - p) / N)
    v0 = 2 * math.pi / (lam * V2K)
    np.testing.assert_allclose(speed, v0, rtol=2e-6)
    np.testing.assert_allclose(ang[refl], 2 * thB, rtol=0, atol=2e-5)
    ratio = n16[refl, 9].astype(np.float64) * N
    assert abs(ratio.mean() - 1.0) < 5 * ratio.std() / math.sqrt(refl.sum()) + 2e-3
    assert np.all(n16[~refl, 9] * N == pytest.approx(1.0, rel=1e-6))
