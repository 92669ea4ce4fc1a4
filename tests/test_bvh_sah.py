"""Host SAH builder (ramp_amd/csrc/bvh_sah.h) on the CPU: the tree a GMesh's
BVH is written from must hold every triangle once, in leaves of one or two,
with consistent parents, subtree sizes and nested boxes, and stay shallow
enough for the preorder pass.  Traversal over it equals the brute-force loop
on the GPU (tests/test_gpu_parity.py)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path_factory.mktemp("sah") / "bvh_sah_check")
    subprocess.run([gxx, "-O2", "-std=c++17", os.path.join(HERE, "native", "bvh_sah_check.cpp"), "-o", exe],
                   check=True)
    return exe


def _check(exe, tris, tmp_path):
    pts = np.ascontiguousarray(np.asarray(tris, dtype=np.float32).reshape(-1, 9))
    path = tmp_path / "pts.bin"
    pts.tofile(path)
    r = subprocess.run([exe, str(path), str(len(pts))], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_sah_tree_of_the_imaging_sphere(checker, tmp_path):
    from ramp_amd import meshes
    out = _check(checker, meshes.icosphere(50, radius=0.015, sigma=0.001, seed=3), tmp_path)
    assert "ok" in out


@pytest.mark.parametrize("n", [1, 2, 3, 5, 1000])
def test_sah_tree_of_coincident_triangles(checker, tmp_path, n):
    # all centroids equal: every split falls back to the middle of the range
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], dtype=np.float32)
    assert "ok" in _check(checker, np.tile(tri, (n, 1)), tmp_path)


def test_sah_tree_of_random_triangles(checker, tmp_path):
    rng = np.random.default_rng(5)
    c = rng.uniform(-1, 1, size=(20000, 1, 3))
    tris = (c + rng.normal(scale=0.01, size=(20000, 3, 3))).reshape(-1, 9)
    assert "ok" in _check(checker, tris, tmp_path)
