"""The algorithmic operation counts behind bench.py's algorithmic roofline
(CPU only).

The oracle's op-count build counts, per history, the primitives the
algorithm needs (Threefry evaluations, divisions, square roots, builtin
transcendentals, guide iterations); tests/native/primitive_cost.py counts
each primitive's VALU instructions in gfx950 code.  bench.py carries both
tables; these tests re-derive them, and pin the counting on instruments
whose counts follow from their definition."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

INSTR = os.path.join(ROOT, "tests", "instruments")


def opcount(*args):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "opcount.py"), *args],
                       check=True, capture_output=True, text=True, cwd=ROOT)
    return json.loads(r.stdout.strip().split("\n")[-1])["per_history"]


def nonzero(d):
    return {k: v for k, v in d.items() if v}


def test_flat_source_counts():
    # MFlat: the energy and the four position draws (the polarisation's two
    # are not read by this plan); the speed's root, and the direction's
    # normalisation (one root, three divisions); its weight 1/N depends on
    # the history count only
    assert nonzero(opcount("--instrument", os.path.join(INSTR, "opcount_flat.json"), "4096", "3")) == \
        {"threefry": 5.0, "sqrt": 2.0, "div": 3.0}


def test_straight_guides_on_the_axis():
    # a parallel beam along the axis of two straight guides: three draws
    # (wavelength, x, y); per guide one plane division and one loop
    # iteration that leaves at the exit (t1 = (l - z)/vz), facing no wall
    assert nonzero(opcount("--instrument", os.path.join(INSTR, "opcount_straight.json"), "4096", "3")) == \
        {"threefry": 3.0, "div": 4.0, "guide_exit": 2.0}


def test_let_source_segment_counts_match_bench():
    got = opcount("65536", "7")
    assert got == bench.ALG_PER_HISTORY
    # the five pre-chopper draws, and one exit iteration per guide entered
    assert got["threefry"] == 5.0
    assert 3.9 < got["guide_exit"] <= 4.0


def test_primitive_costs_match_bench():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not present")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "native", "primitive_cost.py")],
                       check=True, capture_output=True, text=True)
    assert json.loads(r.stdout) == bench.PRIMITIVE_VALU


def test_algorithmic_lane_ops_is_the_weighted_sum():
    alg = bench.algorithmic_lane_ops()
    assert alg == pytest.approx(sum(bench.ALG_PER_HISTORY[k] * c for k, c in bench.PRIMITIVE_VALU.items()))
    # a lower bound: below the instructions the source segment issues
    assert alg < bench.PMC_PER_HISTORY["let_like"]["valu"] * 64
