"""The fixed-point scale and int64 budget of deterministic histograms
(ramp_amd/csrc/det_scale.h), compiled with g++ into a small driver and run
on the CPU.  The budget must fold the device sums into the host before any
bin's sum could pass 2^62, for the w and the w^2 sums alike, and must not
fold (or change the scale) for shards that add up to the chunk the scale was
chosen for."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r'''
#include <cstdio>
#include <cstdlib>
#include "det_scale.h"
/* argv: wb n_scale then triples (n_run) ...; prints per run: fold sh sh2 used used2 */
int main(int argc, char **argv)
{
    DetScale s;
    double wb = atof(argv[1]), n_scale = atof(argv[2]);
    for (int i = 3; i < argc; i++) {
        int f = det_scale_run(s, wb, n_scale, atof(argv[i]));
        printf("%d %d %d %.17g %.17g\n", f, s.sh, s.sh2, s.used, s.used2);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("detscale")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "ramp_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)

    def run(wb, n_scale, runs):
        out = subprocess.run([str(exe), repr(wb), repr(n_scale)] + [repr(r) for r in runs], check=True,
                             capture_output=True, text=True).stdout.split("\n")
        return [tuple(float(x) for x in line.split()) for line in out if line.strip()]
    return run


def test_scale_leaves_sixteen_fold_headroom(driver):
    # wb * n = 1: q sums of one run are at most 2^sh = 2^61 / 16
    (f, sh, sh2, used, used2), = driver(1e-3, 1000.0, [1000.0])
    assert f == 0 and sh == 57 and used == 2.0 ** 57
    # w^2 budget: wb^2 * n = 1e-3 -> sh2 = floor(log2(2^61 / 1.6e-2))
    assert sh2 == 66 and used2 == 1e-6 * 1000 * 2.0 ** 66


def test_equal_runs_fold_before_the_int64_limit(driver):
    rows = driver(1e-3, 1000.0, [1000.0] * 70)
    folds = [i for i, r in enumerate(rows) if r[0]]
    # 32 runs fill 2^62 exactly; the 33rd folds, and so does the 65th
    assert folds == [32, 64]
    assert max(r[3] for r in rows) <= 2.0 ** 62 and max(r[4] for r in rows) <= 2.0 ** 62
    # without the fold a bin taking every history's weight would reach
    # 64 * 2^57 = 2^63 (int64 wrap) at run 64
    assert all(r[1] == 57 for r in rows)


def test_larger_run_refolds_and_rescales(driver):
    rows = driver(1e-3, 1000.0, [1000.0, 1000.0 * 64])
    # the run 64x larger than the first (same weight bound, as an SMCPLIn
    # file's) needs 2^63 of the old scale: fold, then a scale 2^6 coarser
    assert rows[1][0] == 1 and rows[1][1] == 57 - 6
    assert rows[1][3] <= 2.0 ** 57


def test_w2_budget_alone_triggers_a_fold(driver):
    # wb = 1.2, one history per run: one run adds 1.2 * 2^56 to the w budget
    # (0.6 of 2^57) and 1.44 * 2^56 to the w^2 budget (0.72 of 2^57), so the
    # w^2 budget passes 2^62 first, at run 45 (the w budget would at run 54)
    rows = driver(1.2, 1.0, [1.0] * 60)
    assert rows[0][1] == 56 and rows[0][2] == 56
    folds = [i for i, r in enumerate(rows) if r[0]]
    assert folds[0] == 44
    assert max(r[4] for r in rows) <= 2.0 ** 62 and max(r[3] for r in rows) <= 2.0 ** 62


def test_shards_of_a_chunk_keep_the_scale(driver):
    # uneven shards of a chunk of 1e6 (scale chosen for the whole chunk)
    rows = driver(1e-6, 1e6, [3e5, 2e5, 5e5])
    assert [r[0] for r in rows] == [0, 0, 0]
    assert len({r[1] for r in rows}) == 1
    assert rows[-1][3] == pytest.approx(2.0 ** rows[0][1])
