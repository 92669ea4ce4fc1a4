"""Host side of the drop-in (CPU only): instrument description parsing,
component numbering and placement, parameter conversion, the C-ABI surface
and its validation errors."""
import ctypes as C
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

import ramp_amd
from ramp_amd import _capi, _slots

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

IDF = """
{
    // comment line
    "first" : {
        "linear" : true,
        "mod" : { "source" : true,
                  "moderator_kernel" : { "name" : "MFlat", "mod_dim" : [0.1, 0.1],
                                         "target_dim" : [0.1, 0.1], "target_dist" : 1.0,
                                         "E_min" : $0.95*Ei$, "E_max" : $1.05*Ei$ } },  // trailing
        "arm" : { "position" : [0.0, 0.0, 1.0], "rotation" : [0.0, $ang$, 0.0],
                  "geom_kernel" : { "name" : "GArm" }, "scat_kernel" : { "name" : "SArm" } },
        "skipped" : { "disable" : true, "position" : [0.0, 0.0, 2.0],
                      "geom_kernel" : { "name" : "GArm" }, "scat_kernel" : { "name" : "SArm" } },
        "rel" : { "position" : [0.0, 0.0, 0.5], "relative" : "arm",
                  "geom_kernel" : { "name" : "GPlane", "width" : 0.1, "height" : 0.2, "orientation" : "yz" },
                  "scat_kernel" : { "name" : "SDetector1D", "binning" : [1.0, 0.5, 11.0], "var" : "energy" } }
    },
    "second" : {
        "linear" : false, "multi" : 3,
        "a" : { "position" : [0.0, 0.0, 3.0], "restore_neutron" : true,
                "geom_kernel" : { "name" : "GSphere", "radius" : 0.01 },
                "scat_kernel" : { "name" : "SDetector2D", "axis1_var" : "x", "axis1_binning" : [-0.1, 0.01, 0.1],
                                  "axis2_var" : "y", "axis2_binning" : [-0.1, 0.02, 0.1] } }
    }
}
"""


def make(tmp_path, text=IDF, **kw):
    p = tmp_path / "idf.json"
    p.write_text(text)
    return ramp_amd.Instrument(str(p), **kw)


def test_substitution_comments_and_blocks(tmp_path):
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    assert len(inst.blocks) == 2
    src = inst.source
    assert src.E_min == np.float32(9.5) and src.E_max == np.float32(10.5)
    b0, b1 = inst.blocks
    assert b0.linear and not b1.linear and b1.max_events == 3
    assert list(b0.components) == ["arm", "rel"]


def test_component_indices_follow_reference_numbering(tmp_path):
    # mcramp.py:275,112: counter starts at 1, counts the source, skips disabled
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    idx = {c.name: c.idx for c in inst.kernel_refs}
    assert idx == {"arm": 2, "rel": 3, "a": 4}


def test_relative_placement_uses_host_frame_rotation(tmp_path):
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    rel = inst.blocks[0].components["rel"]
    ang = np.float32(0.25)
    # mcramp.py:65-78: pos rotated by the parent's angles then offset
    assert rel.pos[0] == pytest.approx(0.5 * math.sin(ang), rel=1e-6)
    assert rel.pos[2] == pytest.approx(1.0 + 0.5 * math.cos(ang), rel=1e-6)
    assert rel.rot[1] == ang


def test_plan_descriptor_contents(tmp_path):
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    d = inst.plan_desc()
    assert d.n_blocks == 2 and d.n_comps == 3 and d.n_dets == 2
    c = [d.comps[i] for i in range(3)]
    assert [x.comp_idx for x in c] == [2, 3, 4]
    assert c[1].geom.type == _capi.RAMP_G_PLANE and c[1].geom.ni[_slots.GI_ORIENT] == 1
    assert c[1].geom.p[_slots.GP_H] == np.float32(0.2)
    assert c[2].restore == 1
    assert d.blocks[1].linear == 0 and d.blocks[1].multi == 3
    assert d.dets[0].n_bins == 20 and d.dets[1].n_bins == 20 * 10


def test_detector_data_shapes(tmp_path):
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    det1 = inst.blocks[0].components["rel"].scat
    axis, h, e = det1.data()
    assert axis.shape == h.shape == e.shape == (20,)
    det2 = inst.blocks[1].components["a"].scat
    X, Y, Z, E = det2.data()
    assert Z.shape == (10, 20) and X.shape == (10, 20)


def test_powder_cross_sections_match_wrapper_formula():
    p = ramp_amd.SPowderN(d_spacing=[2.0, 3.0], pack=0.5, vc=40.0, sigma_abs=1.5,
                          multiplicity=[2, 4], DW=0.9, F2=[10.0, 20.0])
    q = 2 * np.pi / np.array([2.0, 3.0])
    expect = 4 * np.pi**3 * 0.5 * np.array([2, 4]) * np.array([10.0, 20.0]) * 100 * 0.9 / (40.0**2 * q)
    assert np.allclose(p.sigma_scat_v2, expect, rtol=1e-6)
    assert p.sigma_abs_v == np.float32(0.5 * 1.5 / 40.0 * 100)
    pu = ramp_amd.SPowderN(d_spacing=[2.0], pack=0.5, vc=40.0, multiplicity=[2], DW=1.0, F2=[10.0],
                           U=0.005)
    assert pu.sigma_scat_v2[0] == pytest.approx(
        ramp_amd.SPowderN(d_spacing=[2.0], pack=0.5, vc=40.0, multiplicity=[2], DW=1.0,
                          F2=[10.0]).sigma_scat_v2[0] * math.exp(-0.005 * (2 * math.pi / 2.0) ** 2),
        rel=1e-6)


def test_unknown_and_out_of_scope_components(tmp_path):
    bad = IDF.replace('"name" : "SArm" } },\n        "skipped"', '"name" : "SNoSuchThing" } },\n        "skipped"')
    with pytest.raises(ModuleNotFoundError):
        make(tmp_path, bad, Ei=10.0, ang=0.25)
    with pytest.raises(ValueError):
        ramp_amd.SIsotropic()


def test_stl_roundtrip(tmp_path):
    pts = np.random.default_rng(1).normal(size=(7, 9)).astype(np.float32)
    p = tmp_path / "m.stl"
    ramp_amd.components.write_stl_binary(str(p), pts)
    assert np.array_equal(ramp_amd.components.read_stl(str(p)), pts)
    q = tmp_path / "a.stl"
    with open(q, "w") as f:
        f.write("solid x\n")
        for t in pts:
            f.write(" facet normal 0 0 0\n  outer loop\n")
            for k in range(3):
                f.write(f"   vertex {float(t[3*k])!r} {float(t[3*k+1])!r} {float(t[3*k+2])!r}\n")
            f.write("  endloop\n endfacet\n")
        f.write("endsolid x\n")
    assert np.array_equal(ramp_amd.components.read_stl(str(q)), pts)


# ---- C-ABI surface -------------------------------------------------------

def test_library_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "ramp.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(ramp_\w+)\s*\(", hdr, re.M))
    assert declared == set(_capi.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert declared <= syms


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "ramp.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(ramp_part_desc), sizeof(ramp_component_desc), sizeof(ramp_plan_desc),'
                   'sizeof(ramp_run_opts), sizeof(ramp_run_stats), sizeof(ramp_block_desc));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()))
    assert got == [C.sizeof(_capi.PartDesc), C.sizeof(_capi.ComponentDesc), C.sizeof(_capi.PlanDesc),
                   C.sizeof(_capi.RunOpts), C.sizeof(_capi.RunStats), C.sizeof(_capi.BlockDesc)]


def _err(lib, rc):
    return rc, lib.ramp_last_error().decode()


def test_plan_validation_errors_without_a_gpu(lib, tmp_path):
    inst = make(tmp_path, Ei=10.0, ang=0.25)
    d = inst.plan_desc()
    h = C.c_void_p()
    # valid descriptor, no context: rejected for the context, after validation
    rc, msg = _err(lib, lib.ramp_plan_create(None, C.byref(d), C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "context" in msg
    d.comps[0].geom.type = 99
    rc, msg = _err(lib, lib.ramp_plan_create(None, C.byref(d), C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "geometry type" in msg
    d.comps[0].geom.type = _capi.RAMP_G_ARM
    d.comps[1].det = 7
    rc, msg = _err(lib, lib.ramp_plan_create(None, C.byref(d), C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "detector slot" in msg
    d.comps[1].det = 0
    d.blocks[1].count = 9
    rc, msg = _err(lib, lib.ramp_plan_create(None, C.byref(d), C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "out of range" in msg
    d.blocks[1].count = 1
    d.comps[1].pos[0] = float("nan")
    rc, msg = _err(lib, lib.ramp_plan_create(None, C.byref(d), C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "non-finite" in msg
    rc, msg = _err(lib, lib.ramp_plan_create(None, None, C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID


def test_run_argument_errors_without_a_gpu(lib):
    rc, msg = _err(lib, lib.ramp_run(None, 10, 10, None, 0, None))
    assert rc == _capi.RAMP_E_INVALID and "plan" in msg
    rc, msg = _err(lib, lib.ramp_hist_read(None, 0, None, None, 0))
    assert rc == _capi.RAMP_E_INVALID


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_capi.RampLibraryError):
        _capi.load_library(str(tmp_path / "nope.so"))


def test_tas_scan_settings_follow_bragg_law():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "ramp_amd", "instruments"))
    import tas_scan
    s, c = tas_scan.settings()
    assert len(s) == 256 and len(c) == 256
    assert min(c) >= 1e5 and max(c) <= 4e6
    for x in s:
        assert 3.0 <= x["Ei"] <= 15.0 and tas_scan.EF_MIN <= x["Ef"] <= x["Ei"]
        assert x["Ei"] - x["Ef"] <= 5.0
        # lambda = 2 d sin(theta), E = 81.8042 / lambda^2 (meV, A)
        lam = 2 * 3.355 * math.sin(x["theta_m"])
        assert 81.8042 / lam**2 == pytest.approx(x["Ei"], rel=2e-4)
    inst = ramp_amd.Instrument(os.path.join(ROOT, "ramp_amd", "instruments", "tas_scan.json"), **s[0])
    names = [k.name for k in inst.kernel_refs]
    assert names[-1] == "tube" and inst.blocks[0].components["tube"].scat.n_bins == 128


def test_batch_argument_errors_without_a_gpu(lib):
    h = C.c_void_p()
    rc, msg = _err(lib, lib.ramp_batch_create(None, None, 0, C.byref(h)))
    assert rc == _capi.RAMP_E_INVALID and "context" in msg
    rc, msg = _err(lib, lib.ramp_batch_run(None, None, None, None, None))
    assert rc == _capi.RAMP_E_INVALID and "batch" in msg
    st = _capi.RunStats()
    rc, _ = _err(lib, lib.ramp_batch_stats_read(None, C.byref(st)))
    assert rc == _capi.RAMP_E_INVALID


def test_adjoint_argument_errors_without_a_gpu(lib):
    g = (C.c_double * 4)()
    rc, msg = _err(lib, lib.ramp_adjoint(None, 10, 10, None, 0, 0, None, 0, g, 4, None))
    assert rc == _capi.RAMP_E_INVALID and "plan" in msg


def test_powderdiff_line_table_and_log_derivatives():
    d = [2.0, 3.0]
    p = ramp_amd.SPowderDiff(d_spacing=d, F2=[10.0, 20.0], pack=1.0, vc=40.0, U=0.005,
                             sigma_inc=0.2, sigma_abs=0.1, proposal={"d_spacing": [2.1, 2.9], "U": 0.004})
    L = p.lines
    q = 2 * np.pi / np.array(d)
    # sigma v^2 of SPowderN (powderN.py:59-64) with F2 -> F2 exp(-U q^2)
    s = ramp_amd.SPowderN(d_spacing=d, pack=1.0, vc=40.0, multiplicity=[1, 1], DW=1.0,
                          F2=[10.0 * np.exp(-0.005 * q[0] ** 2), 20.0 * np.exp(-0.005 * q[1] ** 2)])
    assert np.allclose(L[:, _slots.PDL_SSV2], s.sigma_scat_v2, rtol=1e-6)
    # d log(sigma)/dd by central differences of the same formula
    def logsig(dd, U=0.005):
        qq = 2 * np.pi / dd
        return np.log(dd * np.exp(-U * qq * qq))
    num = (logsig(np.array(d) + 1e-6) - logsig(np.array(d) - 1e-6)) / 2e-6
    assert np.allclose(L[:, _slots.PDL_DLS], num, rtol=1e-5)
    assert np.allclose(L[:, _slots.PDL_Q0], 2 * np.pi / np.array([2.1, 2.9]), rtol=1e-6)
    assert np.allclose(L[:, _slots.PDL_D0], [2.1, 2.9]) and np.allclose(L[:, _slots.PDL_D], d)
    with pytest.raises(ValueError):
        ramp_amd.SPowderDiff(d_spacing=[1.0, 2.0], F2=[1.0])
