"""Trace and visualisation on the CPU: the oracle's trace points pinned by
closed forms (straight rays between slits, the slit bounds, the component
frame), the host's assembly of the reference's trace lines
(mcramp.py:118-122, 222-251) from trace points, the component outlines
(lines(), geom/*.py, scat/guide.py:109-123) and the figure
(mcramp/visualisation.py:10-251) drawn with matplotlib's Agg backend."""
import ctypes as C
import math

import numpy as np
import pytest

import ramp_amd
from ramp_amd.instrument import _rotate_points, frame_rotate

TRACE_IDF = """
{
    "b" : {
        "linear" : true,
        "source" : { "source" : true, "position" : [0.0, 0.0, 0.0],
                     "moderator_kernel" : { "name" : "MFlat", "mod_dim" : [0.02, 0.02],
                                            "target_dim" : [0.03, 0.03], "target_dist" : 4.0,
                                            "E_min" : 4.0, "E_max" : 6.0 } },
        "slitA" : { "position" : [0.0, 0.0, 2.0],
                    "geom_kernel" : { "name" : "GPlane", "width" : 0.02, "height" : 0.016 },
                    "scat_kernel" : { "name" : "SArm" } },
        "skip" : { "position" : [0.0, 0.0, 3.0], "visualise" : false,
                   "geom_kernel" : { "name" : "GArm" }, "scat_kernel" : { "name" : "SArm" } },
        "slitB" : { "position" : [0.0, 0.0, 4.0], "rotation" : [0.0, 0.1, 0.0],
                    "geom_kernel" : { "name" : "GPlane", "width" : 0.02, "height" : 0.02 },
                    "scat_kernel" : { "name" : "SArm" } },
        "mon" : { "position" : [0.0, 0.0, 6.0],
                  "geom_kernel" : { "name" : "GPlane", "width" : 0.2, "height" : 0.2 },
                  "scat_kernel" : { "name" : "SDetector1D", "var" : "tof", "binning" : [0.0, 10.0, 10000.0] } }
    }
}
"""

NONLINEAR_IDF = """
{
    "first" : {
        "linear" : true,
        "source" : { "source" : true, "moderator_kernel" : { "name" : "MFlat", "mod_dim" : [0.1, 0.1],
                     "target_dim" : [0.1, 0.1], "target_dist" : 1.0, "E_min" : 1.0, "E_max" : 2.0 } },
        "arm" : { "position" : [0.0, 0.0, 1.0], "geom_kernel" : { "name" : "GArm" },
                  "scat_kernel" : { "name" : "SArm" } }
    },
    "second" : {
        "linear" : false, "multi" : 2,
        "sph" : { "position" : [0.0, 0.0, 2.0], "geom_kernel" : { "name" : "GSphere", "radius" : 0.1 },
                  "scat_kernel" : { "name" : "SArm" } }
    }
}
"""


def _make(tmp_path, text=TRACE_IDF, name="t.json", **kw):
    p = tmp_path / name
    p.write_text(text)
    return ramp_amd.Instrument(str(p), **kw)


def test_oracle_trace_points_closed_form(tmp_path):
    from oracle import pyoracle
    inst = _make(tmp_path)
    n = 20000
    o = pyoracle.run(inst, n, [11], keep_state=True, trace=True)
    tr = o["trace"]
    assert list(tr) == ["slitA", "slitB", "mon"]   # 'skip' is not drawn
    n16 = o["state"][0]
    (pa, alive_a), (pb, alive_b), (_, alive_m) = tr["slitA"], tr["slitB"], tr["mon"]
    # every neutron reaches slitA: a survivor's entry point (component frame,
    # plane z = 0) is inside the slit; a miss leaves the source's sentinel
    # intersection, whose entry point is (0, 0, 0)
    assert np.abs(pa[:, 2]).max() <= 1e-6 and 0.2 * n < alive_a.sum() < 0.95 * n
    assert ((np.abs(pa[alive_a, 0]) < 0.01) & (np.abs(pa[alive_a, 1]) < 0.008)).all()
    assert (pa[~alive_a] == 0.0).all()
    assert not (alive_b & ~alive_a).any() and alive_b.sum() < alive_a.sum()
    assert ((np.abs(pb[alive_b, 0]) < 0.01) & (np.abs(pb[alive_b, 1]) < 0.01)).all()
    assert (pb[alive_a & ~alive_b] == 0.0).all()
    # the monitor detects and kills: no neutron is alive after it
    assert not alive_m.any()
    # detected neutrons end dead on the monitor plane, in the monitor's
    # frame (untransform skips dead neutrons), with their velocity: the
    # straight ray back from there passes through both trace points
    det = alive_b & (n16[:, 15] > 0) & (n16[:, 2] == 0.0)
    assert det.sum() > 0.9 * alive_b.sum()
    p_m = n16[det, 0:3].astype(np.float64) + [0.0, 0.0, 6.0]
    v = n16[det, 3:6].astype(np.float64)
    for pts, comp_pos, rot in ((pa, 2.0, 0.0), (pb, 4.0, 0.1)):
        lab = _rotate_points(pts[det], [0.0, rot, 0.0]).astype(np.float64) + [0.0, 0.0, comp_pos]
        t = (lab[:, 2] - p_m[:, 2]) / v[:, 2]
        ray = p_m + t[:, None] * v
        assert np.abs(lab[:, :2] - ray[:, :2]).max() < 3e-6
    # slitB is tilted: its trace points lie on its own plane z = 0, and in
    # the lab frame off the plane z = 4 by x sin(0.1)
    assert np.abs(pb[alive_b, 2]).max() <= 1e-6
    lab_b = _rotate_points(pb[alive_b], [0.0, 0.1, 0.0]).astype(np.float64)
    assert np.allclose(lab_b[:, 2], -pb[alive_b, 0] * math.sin(0.1), rtol=0, atol=2e-6)
    assert np.abs(lab_b[:, 2]).max() > 5e-4


class _FakeTraceLib:
    """Stands in for libramp's ramp_trace_read with given trace points."""

    def __init__(self, pts):
        self.pts = pts   # slot -> float32 [4, n]

    def ramp_trace_read(self, plan, slot, out, n):
        a = np.ascontiguousarray(self.pts[slot], dtype=np.float32)
        assert a.shape == (4, n)
        C.memmove(out, a.ctypes.data, a.nbytes)
        return 0


TWO_COMP_IDF = """
{
    "b" : {
        "linear" : true,
        "source" : { "source" : true, "moderator_kernel" : { "name" : "MFlat", "mod_dim" : [0.1, 0.1],
                     "target_dim" : [0.1, 0.1], "target_dist" : 1.0, "E_min" : 1.0, "E_max" : 2.0 } },
        "A" : { "position" : [0.0, 0.0, 1.0], "geom_kernel" : { "name" : "GArm" },
                "scat_kernel" : { "name" : "SArm" } },
        "hid" : { "position" : [0.0, 0.0, 1.5], "visualise" : false, "geom_kernel" : { "name" : "GArm" },
                  "scat_kernel" : { "name" : "SArm" } },
        "B" : { "position" : [0.0, 0.0, 2.0], "rotation" : [0.0, 1.5707963267948966, 0.0],
                "geom_kernel" : { "name" : "GArm" }, "scat_kernel" : { "name" : "SArm" } }
    }
}
"""


def test_trace_lines_from_points(tmp_path):
    """Per chunk and block every neutron starts at the origin; after each
    drawn component each live neutron, in index order, gets a line from its
    previous point to the lab-frame point, which becomes its previous point
    (mcramp.py:121-122, 222-251)."""
    inst = _make(tmp_path, TWO_COMP_IDF)
    n = 5
    i = np.arange(n, dtype=np.float32)
    a = np.stack([i, 0 * i, 0 * i, np.array([1, 0, 1, 1, 1], np.float32)])
    b = np.stack([1 + 0 * i, i, 0 * i, np.array([1, 0, 1, 0, 1], np.float32)])
    inst._plan = object()
    inst._trace_lines(_FakeTraceLib({0: a, 1: b}), n, 3)
    lines = inst.blocks[0].trace_lines
    # B is turned a quarter about y: its local (1, i, 0) is (0, i, -1) in the lab
    exp = [((0, 0, 0), (0, 0, 1)), ((0, 0, 0), (2, 0, 1)),      # chunk 0, A: 0 and 2 live
           ((0, 0, 1), (0, 0, 1)), ((2, 0, 1), (0, 2, 1)),      # chunk 0, B
           ((0, 0, 0), (3, 0, 1)), ((0, 0, 0), (4, 0, 1)),      # chunk 1, A
           ((4, 0, 1), (0, 4, 1))]                              # chunk 1, B: 3 died
    assert len(lines) == len(exp)
    for got, (p0, p1) in zip(lines, exp):
        arr = np.array(got, dtype=np.float64)       # [[x0, x1], [y0, y1], [z0, z1]]
        assert np.allclose(arr[:, 0], p0, atol=1e-6) and np.allclose(arr[:, 1], p1, atol=1e-6)
    pts = inst.trace_points()
    assert list(pts) == ["A", "B"] and pts["B"][0].shape == (n, 3)
    assert np.array_equal(pts["B"][1], [True, False, True, False, True])


def test_trace_refuses_drawn_components_in_nonlinear_blocks(tmp_path):
    inst = _make(tmp_path, NONLINEAR_IDF)
    with pytest.raises(ValueError, match="non-linear"):
        inst.execute(10, trace=True)


def test_rotate_points_matches_frame_rotate():
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(50, 3)).astype(np.float32)
    rot = np.array([0.3, -1.1, 2.0], np.float32)
    got = _rotate_points(pts, rot)
    ref = np.array([frame_rotate(list(p), list(rot)) for p in pts], dtype=np.float64)
    assert np.abs(got - ref).max() < 1e-6


def _outline(obj):
    return [np.asarray(v, dtype=np.float64) for v in obj.lines()]


def test_geometry_outlines():
    x, y, z = _outline(ramp_amd.GBox(width=0.2, height=0.4, depth=0.6))
    assert len(x) == 17
    assert set(np.round(np.abs(x), 6)) == {0.1} and set(np.round(np.abs(y), 6)) == {0.2}
    assert set(np.round(np.abs(z), 6)) == {0.3}
    corners = {(round(a, 6), round(b, 6), round(c, 6)) for a, b, c in zip(x, y, z)}
    assert len(corners) == 8
    # consecutive points differ in one coordinate: the walk follows edges
    steps = np.abs(np.diff(np.stack([x, y, z]), axis=1)) > 1e-9
    assert (steps.sum(axis=0) == 1).all()
    for cls in (ramp_amd.GSphere, ramp_amd.GSphereInt):
        x, y, z = _outline(cls(radius=0.3))
        assert len(x) == 202 and np.allclose(np.sqrt(x * x + y * y + z * z), 0.3, atol=1e-6)
        assert np.allclose(x[:50], 0) and np.allclose(z[102:152], 0) and np.allclose(y[152:], 0)
    x, y, z = _outline(ramp_amd.GCylinderExt(radius=0.25, height=0.5))
    assert len(x) == 202 and np.allclose(np.hypot(x, z), 0.25, atol=1e-6)
    assert set(np.round(np.abs(y), 6)) == {0.25}
    assert (x[:101] >= -1e-7).all() and (x[101:201] <= 1e-7).all()
    x, y, z = _outline(ramp_amd.GBanana(radius=2.0, height=1.0, mintheta=-30.0, maxtheta=60.0))
    assert len(x) == 101 and np.allclose(np.hypot(x, z), 2.0, atol=1e-5)
    th = np.degrees(np.arctan2(x, z))
    assert th.min() == pytest.approx(-30.0, abs=1e-4) and th.max() < 60.0
    assert th.max() == pytest.approx(60.0 - 90.0 / 50, abs=1e-4)
    assert np.allclose(y[:50], 0.5) and np.allclose(y[50:100], -0.5)
    x, y, z = _outline(ramp_amd.GPlane(width=0.1, height=0.2))
    assert np.allclose(np.abs(x), 0.05) and np.allclose(np.abs(y), 0.1) and np.allclose(z, 0)
    x, y, z = _outline(ramp_amd.SGuide(w1=0.04, h1=0.06, w2=0.02, h2=0.03, l=2.0))
    assert len(x) == 16 and set(np.round(z, 6)) == {0.0, 2.0}
    assert np.allclose(np.abs(x[z == 0]), 0.02) and np.allclose(np.abs(y[z == 0]), 0.03)
    assert np.allclose(np.abs(x[z == 2]), 0.01) and np.allclose(np.abs(y[z == 2]), 0.015)
    assert np.allclose(_outline(ramp_amd.GArm()), 0) and np.allclose(_outline(ramp_amd.SArm()), 0)


def test_visualise_draws_outlines_and_trace(tmp_path):
    mpl = pytest.importorskip("matplotlib")
    mpl.use("Agg")
    import matplotlib.pyplot as plt
    inst = _make(tmp_path, TRACE_IDF)
    inst.blocks[0].trace_lines.append([[0.0, 0.001], [0.0, 0.002], [0.0, 2.0]])
    vis = inst.visualise(controls=True)
    zx = vis.axes["zx"].get_lines()
    # slitA, slitB, mon: geometry and scatter outline each, then the trace line
    assert len(zx) == 2 * 3 + 1
    # slitA's frame: horizontal axis lab z, vertical lab x
    assert np.allclose(zx[0].get_xdata(), 2.0) and np.allclose(np.abs(zx[0].get_ydata()), 0.01)
    # slitB is turned by 0.1 rad about y: its edge spans z = 4 +- 0.01 sin 0.1
    zb = np.asarray(zx[2].get_xdata())
    assert zb.max() - 4.0 == pytest.approx(0.01 * math.sin(0.1), rel=1e-4)
    red = zx[-1]
    assert red.get_color() == "r" and np.allclose(red.get_xdata(), [0.0, 2.0])
    assert np.allclose(red.get_ydata(), [0.0, 0.001])
    assert len(vis.ax3d.get_lines()) == 7
    vis.focus("slitB")
    zb = np.asarray(vis.axes["zx"].get_lines()[2].get_xdata())
    assert abs(zb.max() + zb.min()) < 1e-6 and zb.max() > 0.0
    lo, hi = vis.axes["zx"].get_xlim()
    assert lo == pytest.approx(-hi)
    vis.set_limits(1, -0.5, 0.5)
    assert vis.axes["xy"].get_ylim() == (-0.5, 0.5) and vis.axes["zy"].get_ylim() == (-0.5, 0.5)
    plt.close("all")
