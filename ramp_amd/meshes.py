"""Synthetic triangle meshes for the imaging workload.

icosphere(): class-I geodesic icosphere of frequency f (20 f^2 triangles):
every face of the regular icosahedron is divided into f^2 triangles on a
barycentric grid, grid points are projected onto the sphere of the given
radius, and each distinct vertex (shared by neighbouring triangles, so the
surface stays closed) is displaced radially by a normal deviate of the
given sigma drawn from numpy.random.default_rng(seed).
"""
import numpy as np


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def icosphere(freq=50, radius=0.015, sigma=0.001, seed=3):
    """Triangles (n, 9) float32 of a noisy closed icosphere, outward normals."""
    V, F = _icosahedron()
    key_of = {}
    verts = []

    def vid(p):
        u = p / np.linalg.norm(p)
        k = tuple(np.round(u, 9))
        if k not in key_of:
            key_of[k] = len(verts)
            verts.append(u)
        return key_of[k]

    tris = []
    for a, b, c in F:
        A, B, C = V[a], V[b], V[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (C - A) * (j / freq)
                grid[(i, j)] = vid(p)
        for i in range(freq):
            for j in range(freq - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < freq - 1:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    verts = np.array(verts)
    rng = np.random.default_rng(seed)
    r = radius + rng.normal(0.0, sigma, size=len(verts))
    P = verts * r[:, None]
    T = np.array(tris)
    pts = P[T].reshape(-1, 9)
    # orient every triangle outward (normal along the centroid direction)
    e1 = pts[:, 3:6] - pts[:, 0:3]
    e2 = pts[:, 6:9] - pts[:, 0:3]
    nrm = np.cross(e1, e2)
    cen = (pts[:, 0:3] + pts[:, 3:6] + pts[:, 6:9]) / 3.0
    flip = (nrm * cen).sum(axis=1) < 0
    pts[flip] = pts[flip][:, [0, 1, 2, 6, 7, 8, 3, 4, 5]]
    return pts.astype(np.float32)
