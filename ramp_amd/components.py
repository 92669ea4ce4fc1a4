"""Component classes of the drop-in: same names and keyword arguments as the
reference's wrappers (mcramp/moderator, mcramp/geom, mcramp/scat), resolved
by name from the instrument JSON exactly as mcramp.py:58-92 resolves them.

A component here does not launch anything: it converts its arguments the
way the reference wrapper does (np.float32 / np.uint32, the same
pre-computations) and writes them into a POD part descriptor (include/
ramp.h, slots in csrc/rslots.h).  The Instrument fuses all components into
one plan executed by libramp.so.

Classes marked "no reference counterpart" implement configuration elements
of the benchmark workloads that the reference cannot express.
"""
import math
import os
import re

import numpy as np

from . import _capi as capi
from . import _slots as S


class _Part:
    kind = None
    type_code = 0

    def _init_part(self):
        self.p = np.zeros(capi.RAMP_NP, dtype=np.float32)
        self.ni = np.zeros(capi.RAMP_NI, dtype=np.uint32)
        self.tables = [None] * capi.RAMP_NT

    def fill(self, desc, pool):
        """Write this part into a capi.PartDesc; tables are appended to pool."""
        desc.type = self.type_code
        for i in range(capi.RAMP_NP):
            desc.p[i] = float(self.p[i])
        for i in range(capi.RAMP_NI):
            desc.ni[i] = int(self.ni[i])
        for k, t in enumerate(self.tables):
            if t is None:
                desc.tab_off[k] = 0
                desc.tab_len[k] = 0
            else:
                arr = np.ascontiguousarray(t, dtype=np.float32).ravel()
                desc.tab_off[k] = pool.append(arr)
                desc.tab_len[k] = arr.size

    def lines(self):
        return [[0], [0], [0]]


class TablePool:
    """Concatenated float32 tables of a plan (aligned to 64 floats)."""

    def __init__(self):
        self.parts = []
        self.size = 0

    def append(self, arr):
        pad = (-self.size) % 64
        if pad:
            self.parts.append(np.zeros(pad, dtype=np.float32))
            self.size += pad
        off = self.size
        self.parts.append(arr)
        self.size += arr.size
        return off

    def array(self):
        if not self.parts:
            return np.zeros(64, dtype=np.float32)
        return np.concatenate(self.parts).astype(np.float32)


# ============================ sources ====================================


class _Source(_Part):
    kind = "source"
    draws_per_history = 7

    def _geom_args(self, mod_dim, target_dim, target_dist, E_min, E_max):
        # flat_mod.py:14-18: float2 / float32 conversions
        self.mod_dim = np.array(mod_dim[:2], dtype=np.float32)
        self.target_dim = np.array(target_dim[:2], dtype=np.float32)
        self.target_dist = np.float32(target_dist)
        self.E_min = np.float32(E_min)
        self.E_max = np.float32(E_max)
        self.p[S.SP_MOD_W], self.p[S.SP_MOD_H] = self.mod_dim
        self.p[S.SP_TGT_W], self.p[S.SP_TGT_H] = self.target_dim
        self.p[S.SP_TGT_D] = self.target_dist
        self.p[S.SP_EMIN] = self.E_min
        self.p[S.SP_EMAX] = self.E_max

    def weight_bound(self, num_sim):
        """Upper bound of one history's weight (deterministic histograms)."""
        return 1.0 / num_sim


class MFlat(_Source):
    """Flat energy spectrum source (mcramp/moderator/flat_mod.py:10-32)."""
    type_code = capi.RAMP_SRC_FLAT

    def __init__(self, ctx=None, spec_file="Let_Base.mcstas", mod_dim=(0.0, 0.0),
                 target_dim=(0.0, 0.0), target_dist=0.0, E_min=0.0, E_max=0.0):
        self._init_part()
        self._geom_args(mod_dim, target_dim, target_dist, E_min, E_max)


def _maxwell_ref64(E, T):
    """float64 value of maxwell_energy_distn (cont_mod.cl:4-9) for bounds."""
    kB = 1.38066e-23
    return 2.0 * np.sqrt(E) / np.sqrt(np.pi) * (1.0 / (kB * T)) ** 1.5 * np.exp(-E / (kB * T))


class MCont(_Source):
    """Sum of three Maxwellians (mcramp/moderator/cont_mod.py:10-70).

    Extension (no reference counterpart): tail_join > 0 replaces the first
    Maxwellian above E_j = tail_join*k_B*T1 by an epithermal 1/E tail joined
    continuously at E_j."""
    type_code = capi.RAMP_SRC_CONT

    def __init__(self, ctx=None, spec_file="Let_Base.mcstas", mod_dim=(0.0, 0.0),
                 target_dim=(0.0, 0.0), target_dist=0.0, E_min=0.0, E_max=0.0, T1=0.0, I1=0.0,
                 T2=0.0, I2=0.0, T3=0.0, I3=0.0, tail_join=0.0):
        self._init_part()
        self._geom_args(mod_dim, target_dim, target_dist, E_min, E_max)
        self.T1, self.I1 = np.float32(T1), np.float32(I1)
        self.T2, self.I2 = np.float32(T2), np.float32(I2)
        self.T3, self.I3 = np.float32(T3), np.float32(I3)
        self.str_area = np.float32(1)
        self.p[S.SP_T1], self.p[S.SP_I1] = self.T1, self.I1
        self.p[S.SP_T2], self.p[S.SP_I2] = self.T2, self.I2
        self.p[S.SP_T3], self.p[S.SP_I3] = self.T3, self.I3
        self.p[S.SP_STR_AREA] = self.str_area
        self.p[S.SP_TAIL_JOIN] = np.float32(tail_join)

    def weight_bound(self, num_sim):
        # each term is I * M(E)/M(kT) and M(E)/M(kT) <= sqrt(1/2) e^(1/2) < 1.17
        tot = 0.0
        for I in (self.I1, self.I2, self.I3):
            tot += abs(float(I)) * 1.2
        return tot / num_sim


def read_tally_spectrum(path):
    """Read a moderator spectrum in the McStas/MCNP tally layout that
    MISIS consumes (e.g. the reference's examples/Let_Base.mcstas).

    The file is line oriented:
      - header lines, one of which is the comment 'c TimeOffset <value>';
      - one block per energy bin, opened by the line
        'energy bin: <lower> to <upper>' (MeV) and preceded by its
        'surface' / 'segment:' / 'user bin' lines; after the 'cosine:' and
        'time' lines come the time rows '<time> <value> (<error>)', time in
        units of 1e-8 s, and the block ends with 'total <sum> (<error>)';
      - a closing line ('1end').
    Every block has the same time column.

    Returns a dict: 'time_offset' (float), 'e_lower' and 'e_upper' (meV,
    float64 arrays), 'time' (the time column in seconds, float64), 'values'
    (float64 array [n_blocks, n_times]) and 'totals' (the blocks' total
    lines, float64)."""
    time_offset = None
    e_lo, e_hi, totals, blocks = [], [], [], []
    rows = None
    with open(path, "r") as f:
        for line in f:
            words = line.split()
            if not words:
                continue
            if time_offset is None and "TimeOffset" in line:
                time_offset = float(words[-1])
            elif line.startswith("energy bin"):
                lo, hi = line.split(":", 1)[1].split("to")
                e_lo.append(float(lo) * 1e9)
                e_hi.append(float(hi) * 1e9)
                rows = []
                blocks.append(rows)
            elif words[0] == "total" and rows is not None:
                totals.append(float(words[1]))
                rows = None
            elif rows is not None and len(words) == 3 and words[2].startswith("("):
                rows.append((float(words[0]), float(words[1])))
    if time_offset is None or not blocks:
        raise ValueError(f"{path}: not a tally spectrum (no TimeOffset line or energy-bin blocks)")
    n_t = len(blocks[0])
    if any(len(b) != n_t for b in blocks) or len(totals) != len(blocks):
        raise ValueError(f"{path}: energy-bin blocks of unequal length or without total lines")
    table = np.array(blocks, dtype=np.float64)          # [block, time, (time, value)]
    return {"time_offset": time_offset, "e_lower": np.array(e_lo), "e_upper": np.array(e_hi),
            "time": table[0, :, 0] / 1e8, "values": table[:, :, 1], "totals": np.array(totals)}


def _energy_window(e_lower, e_upper, E_min, E_max):
    """The blocks MISIS keeps for the window [E_min, E_max] and their
    weights (isis_mod.py:127-150): from the block containing E_min (the one
    before the first lower edge >= E_min; when E_min is the file's first
    lower edge, that block itself) up to the block before the first lower
    edge >= E_max, each weighted by the part of its [lower, upper] inside
    the window.  Returns (first, stop, weights, edges): blocks [first,
    stop), and edges E_min, the kept lower edges after the first block,
    E_max.  The weights are evaluated in float64 against the float32
    window, as the numpy of the reference's time evaluates Python floats
    against float32 scalars."""
    start = int(np.searchsorted(e_lower, E_min))
    if E_min == e_lower[0]:
        start += 1
    stop = int(np.searchsorted(e_lower, E_max))
    first = start - 1
    if first < 0:
        raise ValueError(f"MISIS: E_min {E_min} meV lies below the spectrum's first energy bin")
    weights = []
    for lo, hi in zip(e_lower[first:stop], e_upper[first:stop]):
        width = hi - lo
        inside = (hi - E_min) / width if E_min > lo else 1.0
        if E_max < hi:
            inside -= (hi - E_max) / width
        weights.append(inside)
    edges = np.concatenate([[E_min], e_lower[start:stop], [E_max]])
    return first, stop, np.array(weights), edges


def _target_solid_angle_factor(mod_dim, target_dim, target_dist, n=50):
    """MISIS's str_area (isis_mod.py:43-62): a midpoint-like sum of 1/r^2
    between grid points of the moderator face (x and y in [-w/2, w/2]) and
    of one quadrant of the target window (x and y in [0, w/2]), times the
    two areas (the target's in cm^2).  The reference pairs each moderator x
    with the target x of the same grid index (a row of its x grid
    broadcasts against the target's x grid) and runs every moderator y
    against every target y, and repeats the sum once per row of its x grid
    (n times).  In the far field it tends to A_mod * A_target[cm^2] / d^2."""
    mw, mh = mod_dim
    tw, th = target_dim
    tx = np.linspace(0, tw / 2.0, num=n)
    ty = np.linspace(0, th / 2.0, num=n)
    mx = np.linspace(-mw / 2.0, mw / 2.0, num=n)
    my = np.linspace(-mh / 2.0, mh / 2.0, num=n)
    dx2 = (mx - tx) ** 2                                  # [j]: paired by index
    dy2 = (my[:, None] - ty[None, :]) ** 2                # [moderator y, target y]
    total = n * np.sum(1.0 / (dx2[None, None, :] + dy2[:, :, None] + target_dist ** 2))
    return np.float32(total * (mw * mh) / n ** 4 * tw * th * 10000)


class MISIS(_Source):
    """Tabulated ISIS moderator spectrum (mcramp/moderator/isis_mod.py:10-185).

    The spectrum file is read by read_tally_spectrum; the energy window,
    the cumulative flux table over (energy block, time row), the cumulative
    sums at block ends and the total follow isis_mod.py:121-166."""
    type_code = capi.RAMP_SRC_ISIS
    draws_per_history = 8

    def __init__(self, ctx=None, spec_file="Let_Base.mcstas", mod_dim=(0.0, 0.0),
                 target_dim=(0.0, 0.0), target_dist=0.0, E_min=0.0, E_max=0.0):
        self._init_part()
        self._geom_args(mod_dim, target_dim, target_dist, E_min, E_max)
        self.load_spectrum(spec_file)
        self.str_area = _target_solid_angle_factor(mod_dim, target_dim, target_dist)
        self.p[S.SP_ISIS_TOTAL] = self.total
        self.p[S.SP_ISIS_STR_AREA] = self.str_area
        self.p[S.SP_ISIS_TOFFSET] = self.time_offset
        self.ni[S.SI_ISIS_NT] = self.num_time_bins
        self.ni[S.SI_ISIS_NE] = self.num_ener_bins
        self.tables[S.ST_ISIS_FLUX] = self.flux
        self.tables[S.ST_ISIS_TBINS] = self.t_bins
        self.tables[S.ST_ISIS_EBINS] = self.e_bins
        self.tables[S.ST_ISIS_EINT] = self.e_int

    def load_spectrum(self, spec_file):
        spec = read_tally_spectrum(spec_file)
        self.time_offset = np.float32(spec["time_offset"])
        first, stop, weights, edges = _energy_window(spec["e_lower"], spec["e_upper"],
                                                     self.E_min, self.E_max)
        kept = spec["values"][first:stop] * weights[:, None]
        # cumulative flux over (block, time row) in row-major order, summed
        # in float64 and stored as float32; the cumulative value at the end
        # of each block, after a leading zero
        cum = np.cumsum(kept.ravel())
        self.flux = cum.astype(np.float32)
        self.e_int = np.concatenate([[0.0], cum.reshape(kept.shape)[:, -1]]).astype(np.float32)
        self.total = np.float32(kept.sum())
        self.num_time_bins = np.int32(kept.shape[1])
        self.num_ener_bins = np.int32(len(edges))
        self.t_bins = spec["time"].astype(np.float32)
        self.e_bins = edges.astype(np.float32)

    def weight_bound(self, num_sim):
        return float(self.str_area) * float(self.total) * 6.2415093e12 / num_sim


class MLambda(_Source):
    """Wavelength-uniform source with weight proportional to 1/lambda,
    normalised to unit expected total weight.  No reference counterpart:
    MFlat samples energy, not wavelength.  Draw order follows MFlat."""
    type_code = capi.RAMP_SRC_LAMBDA

    def __init__(self, ctx=None, mod_dim=(0.0, 0.0), target_dim=(0.0, 0.0), target_dist=0.0,
                 lambda_min=1.0, lambda_max=10.0):
        self._init_part()
        self._geom_args(mod_dim, target_dim, target_dist, 0.0, 0.0)
        lmin, lmax = float(lambda_min), float(lambda_max)
        if not (0 < lmin < lmax):
            raise ValueError("MLambda needs 0 < lambda_min < lambda_max")
        self.p[S.SP_LMIN] = np.float32(lmin)
        self.p[S.SP_LMAX] = np.float32(lmax)
        self.p[S.SP_LNORM] = np.float32((lmax - lmin) / np.log(lmax / lmin))
        self._wmax = float(self.p[S.SP_LNORM]) / lmin

    def weight_bound(self, num_sim):
        return self._wmax * 1.001 / num_sim


class MParallel(_Source):
    """Parallel beam along +z over a width x height face, wavelength uniform.
    No reference counterpart (MFlat always aims at a target window)."""
    type_code = capi.RAMP_SRC_PARALLEL
    draws_per_history = 5

    def __init__(self, ctx=None, width=0.0, height=0.0, lambda_min=1.0, lambda_max=10.0):
        self._init_part()
        self.p[S.SP_MOD_W] = np.float32(width)
        self.p[S.SP_MOD_H] = np.float32(height)
        self.p[S.SP_LMIN] = np.float32(lambda_min)
        self.p[S.SP_LMAX] = np.float32(lambda_max)


# ============================ geometry ===================================


class _Geom(_Part):
    kind = "geom"

    def __init__(self):
        self._init_part()


class GArm(_Geom):
    """geom/arm.py:7-37"""
    type_code = capi.RAMP_G_ARM

    def __init__(self, idx=0, ctx=None):
        super().__init__()
        self.idx = np.uint32(idx)


class GPlane(_Geom):
    """geom/plane.py:9-77"""
    type_code = capi.RAMP_G_PLANE

    def __init__(self, width=0, height=0, idx=0, orientation="xy", ctx=None):
        super().__init__()
        orientations = {"xy": 0, "yz": 1}
        self.orientation_str = orientation
        self.orientation = np.uint32(orientations[orientation])
        self.width = np.float32(width)
        self.height = np.float32(height)
        self.idx = np.uint32(idx)
        self.p[S.GP_W], self.p[S.GP_H] = self.width, self.height
        self.ni[S.GI_ORIENT] = self.orientation

    def lines(self):
        w, h = self.width / 2.0, self.height / 2.0
        x = [-w, w, w, -w, -w]
        y = [h, h, -h, -h, h]
        z = [0.0] * 5
        return [x, y, z] if self.orientation_str == "xy" else [z, y, x]


class GBox(_Geom):
    """geom/box.py:9-82"""
    type_code = capi.RAMP_G_BOX

    def __init__(self, width=0, height=0, depth=0, idx=0, ctx=None):
        super().__init__()
        self.width, self.height, self.depth = np.float32(width), np.float32(height), np.float32(depth)
        self.idx = np.uint32(idx)
        self.p[S.GP_W], self.p[S.GP_H], self.p[S.GP_D] = self.width, self.height, self.depth

    # corner walk that draws the twelve edges as one polyline (box.py:58-82)
    _WALK = (0, 1, 2, 3, 0, 4, 5, 1, 2, 6, 5, 6, 7, 4, 0, 3, 7)

    def lines(self):
        w, h, d = self.width / 2.0, self.height / 2.0, self.depth / 2.0
        # corners 0-3 on the z = -d face, 4-7 on z = +d, each face walked
        # (-w, h), (w, h), (w, -h), (-w, -h)
        face = [(-w, h), (w, h), (w, -h), (-w, -h)]
        corners = [(x, y, -d) for x, y in face] + [(x, y, d) for x, y in face]
        pts = [corners[k] for k in self._WALK]
        return [[p[0] for p in pts], [p[1] for p in pts], [p[2] for p in pts]]


class GSphere(_Geom):
    """geom/sphere.py:9-83"""
    type_code = capi.RAMP_G_SPHERE

    def __init__(self, radius=0, idx=0, ctx=None):
        super().__init__()
        self.radius = np.float32(radius)
        self.idx = np.uint32(idx)
        self.p[S.GP_R] = self.radius

    def lines(self):
        """Three great circles and a quarter arc as one polyline
        (sphere.py:49-83): the x = 0 circle, back to its start, the z = 0
        circle, back to the same start, a quarter of the z = 0 circle again,
        then the y = 0 circle."""
        r = self.radius
        full = np.linspace(0, 2 * np.pi)
        quarter = np.linspace(0, np.pi / 2)
        zero = np.zeros_like(full)
        parts = [
            (zero, r * np.cos(full), r * np.sin(full)),
            (np.zeros(1), np.array([r * np.cos(0.0)]), np.array([r * np.sin(0.0)])),
            (r * np.sin(full), r * np.cos(full), zero),
            (np.zeros(1), np.array([r * np.cos(0.0)]), np.array([r * np.sin(0.0)])),
            (r * np.sin(quarter), r * np.cos(quarter), np.zeros_like(quarter)),
            (r * np.cos(full), zero, r * np.sin(full)),
        ]
        return [list(np.concatenate([p[k] for p in parts])) for k in range(3)]


class GSphereInt(GSphere):
    """geom/sphere_int.py:9-81"""
    type_code = capi.RAMP_G_SPHERE_INT


class GCylinderExt(_Geom):
    """geom/cylinder_ext.py:10-97"""
    type_code = capi.RAMP_G_CYLINDER

    def __init__(self, radius=0, height=0, idx=0, ctx=None):
        super().__init__()
        self.radius, self.height = np.float32(radius), np.float32(height)
        self.idx = idx
        self.p[S.GP_R], self.p[S.GP_HEIGHT] = self.radius, self.height

    def lines(self):
        """The two half-cylinders x >= 0 and x <= 0, each as top rim, bottom
        rim walked back, and the closing point (cylinder_ext.py:57-97)."""
        r, h = self.radius, self.height / 2.0
        xs, ys, zs = [], [], []
        for sweep in (np.linspace(0, np.pi), np.linspace(0, -np.pi)):
            for a_seq, y in ((sweep, h), (sweep[::-1], -h)):
                xs += list(r * np.sin(a_seq))
                zs += list(r * np.cos(a_seq))
                ys += [y] * len(a_seq)
            xs.append(xs[0]); ys.append(ys[0]); zs.append(zs[0])
        return [xs, ys, zs]


class GBanana(_Geom):
    """geom/banana.py:10-87"""
    type_code = capi.RAMP_G_BANANA

    def __init__(self, radius=0, height=0, mintheta=0, maxtheta=0, idx=0, ctx=None):
        super().__init__()
        self.radius, self.height = np.float32(radius), np.float32(height)
        self.mintheta, self.maxtheta = np.float32(mintheta), np.float32(maxtheta)
        self.idx = idx
        self.p[S.GP_R], self.p[S.GP_HEIGHT] = self.radius, self.height
        self.p[S.GP_MINTH], self.p[S.GP_MAXTH] = self.mintheta, self.maxtheta

    def lines(self):
        """The arc between mintheta and maxtheta (degrees) in 50 steps: top
        rim walked back from the last step, bottom rim forward, closed
        (banana.py:65-87)."""
        n = 50
        lo = np.radians(self.mintheta)
        step = (np.radians(self.maxtheta) - lo) / float(n)
        ang = [lo + i * step for i in range(n)]
        top, bottom = ang[::-1], ang
        xs = [self.radius * np.sin(a) for a in top + bottom]
        zs = [self.radius * np.cos(a) for a in top + bottom]
        ys = [self.height / 2.0] * n + [-self.height / 2.0] * n
        return [xs + xs[:1], ys + ys[:1], zs + zs[:1]]


def read_stl(filename):
    """Triangles of an ASCII or binary STL file as float32 (n, 9), the
    layout of numpy-stl's Mesh.points used by geom/mesh.py:68-73."""
    with open(filename, "rb") as f:
        raw = f.read()
    head = raw[:5].lower()
    if head == b"solid" and b"facet" in raw[:1024]:
        vals = re.findall(rb"vertex\s+(\S+)\s+(\S+)\s+(\S+)", raw)
        pts = np.array(vals, dtype=np.float64).astype(np.float32).reshape(-1, 9)
        return pts
    n = int(np.frombuffer(raw[80:84], dtype="<u4")[0])
    rec = np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("attr", "<u2")])
    arr = np.frombuffer(raw[84:84 + n * 50], dtype=rec, count=n)
    return np.ascontiguousarray(arr["v"], dtype=np.float32)


def write_stl_binary(filename, points):
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 9)
    rec = np.zeros(pts.shape[0], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9),
                                                  ("attr", "<u2")]))
    rec["v"] = pts
    with open(filename, "wb") as f:
        f.write(b"\0" * 80)
        f.write(np.uint32(pts.shape[0]).tobytes())
        f.write(rec.tobytes())


class GMesh(_Geom):
    """geom/mesh.py:11-99 (triangle mesh from an STL file).  The reference
    traces every triangle; libramp builds a BVH on the GPU with identical
    nearest-hit and crossing-count results."""
    type_code = capi.RAMP_G_MESH

    def __init__(self, filename="", interior=False, idx=0, ctx=None, points=None):
        super().__init__()
        self.idx = np.uint32(idx)
        pts = read_stl(filename) if points is None else np.asarray(points, np.float32).reshape(-1, 9)
        self.points = pts
        allv = pts.reshape(-1, 3)
        lo, hi = allv.min(axis=0), allv.max(axis=0)
        self.p[S.GP_XMIN], self.p[S.GP_XMAX] = lo[0], hi[0]
        self.p[S.GP_YMIN], self.p[S.GP_YMAX] = lo[1], hi[1]
        self.p[S.GP_ZMIN], self.p[S.GP_ZMAX] = lo[2], hi[2]
        self.num_tri = np.uint32(pts.shape[0])
        self.interior = np.uint32(1 if interior else 0)
        self.ni[S.GI_NTRI] = self.num_tri
        self.ni[S.GI_INTERIOR] = self.interior
        self.tables[S.GT_POINTS] = pts.ravel()


# ============================ scattering =================================


class _Scat(_Part):
    kind = "scat"
    n_bins = 0  # histogram bins per moment (0: no histogram)

    def __init__(self):
        self._init_part()
        self._w = None
        self._w2 = None

    def _set_hist(self, w, w2):
        self._w, self._w2 = w, w2

    def data_reduce(self, queue=None):
        return 0

    def data(self, queue=None):
        return 0

    def plot(self, queue=None):
        return 0

    def save(self, queue=None):
        return 0


class SArm(_Scat):
    """scat/arm.py:7-38"""
    type_code = capi.RAMP_S_ARM

    def __init__(self, idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)


class SBeamstop(_Scat):
    """scat/beamstop.py:8-40"""
    type_code = capi.RAMP_S_BEAMSTOP

    def __init__(self, idx=0, ctx=None, filename=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)


class SChopper(_Scat):
    """scat/chopper.py:10-70 (freq is an angular frequency in rad/s)"""
    type_code = capi.RAMP_S_CHOPPER

    def __init__(self, slit_width=0.0, radius=0.0, freq=0.0, n_slits=0, phase=0.0, jitter=0.0,
                 idx=0, ctx=0, **kwargs):
        super().__init__()
        self.slit_width = np.float32(slit_width)
        self.radius = np.float32(radius)
        self.freq = np.float32(freq)
        self.n_slits = np.uint32(n_slits)
        self.phase = np.float32(phase)
        self.jitter = np.float32(jitter)
        self.idx = np.uint32(idx)
        self.p[S.CP_SLIT_W], self.p[S.CP_RADIUS] = self.slit_width, self.radius
        self.p[S.CP_FREQ], self.p[S.CP_PHASE], self.p[S.CP_JITTER] = self.freq, self.phase, self.jitter
        self.ni[S.CI_NSLITS] = self.n_slits


class SGuide(_Scat):
    """scat/guide.py:10-123 (float build; CONFIG_USE_DOUBLE is not offered)"""
    type_code = capi.RAMP_S_GUIDE

    def __init__(self, w1=0, h1=0, w2=0, h2=0, l=0, R0=0, Qc=0, alpha=0, m=1, W=0, idx=0, ctx=0,
                 max_bounces=50, **kwargs):
        super().__init__()
        vals = [w1, h1, w2, h2, l, R0, Qc, alpha, m, W]
        (self.w1, self.h1, self.w2, self.h2, self.l, self.R0, self.Qc, self.alpha, self.m,
         self.W) = [np.float32(v) for v in vals]
        self.idx = np.uint32(idx)
        self.max_bounces = np.uint32(max_bounces)
        for slot, v in zip((S.GU_W1, S.GU_H1, S.GU_W2, S.GU_H2, S.GU_L, S.GU_R0, S.GU_QC,
                            S.GU_ALPHA, S.GU_M, S.GU_W), vals):
            self.p[slot] = np.float32(v)
        self.ni[S.GUI_MAXB] = self.max_bounces

    def lines(self):
        """Outline of the tapered guide, entrance (w1 x h1) at z = 0, exit
        (w2 x h2) at z = l, as one polyline (guide.py:109-123)."""
        a, b = self.w1 / 2, self.h1 / 2      # entrance half-sizes
        c, d = self.w2 / 2, self.h2 / 2      # exit half-sizes
        l = self.l
        pts = [(-a, b, 0), (-c, d, l), (c, d, l), (a, b, 0), (-a, b, 0), (-a, -b, 0),
               (a, -b, 0), (a, b, 0), (c, d, l), (c, -d, l), (a, -b, 0), (-a, -b, 0),
               (-c, -d, l), (c, -d, l), (-c, -d, l), (-c, d, l)]
        return [[p[0] for p in pts], [p[1] for p in pts], [p[2] for p in pts]]


def _simps_legacy(y, x, axis=-1):
    """Composite Simpson integration of samples y at abscissae x along
    `axis`, as scipy.integrate.simps (SciPy <= 1.10, default even='avg';
    removed in SciPy 1.14) computes it: for an odd sample count, Simpson's
    rule on pairs of possibly unequal intervals h0, h1,
      (h0 + h1)/6 * (y0 (2 - 1/(h0/h1)) + y1 (h0 + h1)^2/(h0 h1) + y2 (2 - h0/h1)),
    summed; for an even count, the average of (Simpson on all but the last
    sample + the last interval's trapezoid) and (Simpson on all but the
    first sample + the first interval's trapezoid).  Arithmetic stays in the
    inputs' dtype (float32 tables stay float32), as in that SciPy.  The
    reference calls it in its S(q, w) loader (isotropic.py:7, 201-203,
    211); the installed SciPy no longer has it."""
    y = np.asarray(y)
    axis = axis % y.ndim
    n = y.shape[axis]
    # x as a column along `axis`, so that its differences broadcast
    shape = [1] * y.ndim
    shape[axis] = n
    x = np.asarray(x).reshape(shape)

    def take(a, start, stop, step=2):
        return a[(slice(None),) * axis + (slice(start, stop, step),)]

    def pairs(lo, hi):
        # Simpson over samples lo .. hi (hi - lo even); the sum runs along
        # `axis` in place, as that SciPy's np.sum did
        h = np.diff(x, axis=axis)
        h0, h1 = take(h, lo, hi), take(h, lo + 1, hi + 1)
        hsum = h0 + h1
        ratio = h0 / h1
        terms = hsum / 6.0 * (take(y, lo, hi) * (2 - 1.0 / ratio)
                              + take(y, lo + 1, hi + 1) * hsum * hsum / (h0 * h1)
                              + take(y, lo + 2, hi + 2) * (2 - ratio))
        return np.sum(terms, axis=axis)

    if n % 2 == 1:
        return pairs(0, n - 2)

    def one(a, i):
        return a[(slice(None),) * axis + (i,)]
    ends = 0.5 * (one(x, -1) - one(x, -2)) * (one(y, -1) + one(y, -2))
    ends = ends + 0.5 * (one(x, 1) - one(x, 0)) * (one(y, 1) + one(y, 0))
    return (pairs(0, n - 3) + pairs(1, n - 2)) / 2.0 + ends / 2.0


def _running_cdf(x, density):
    """The normalised running integral of `density` over x at every sample
    (isotropic.py:210-212): the legacy Simpson rule over samples 0..i for
    each i (zero for i = 0), in float32, divided by its last value."""
    partial = np.array([_simps_legacy(density[:i + 1], x[:i + 1]) for i in range(len(x))], dtype=np.float32)
    return partial / partial[-1]


_SQW_KEY_CHARS = set("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz_[\\]^`")


def _sqw_header_entry(line):
    """'# <key> <number>' of an S(q, w) file, or None.  The key runs over
    letters, '_' and the ASCII characters between 'Z' and 'a' (the
    reference's character range admits them); the value is the longest
    leading decimal '<digits>.<digits>' or integer of the next word."""
    at = line.find("#")
    while at >= 0:
        rest = line[at + 1:]
        words = rest.split()
        if rest[:1].isspace() and len(words) >= 2 and all(ch in _SQW_KEY_CHARS for ch in words[0]):
            m = re.match(r"[-+]?(\d*\.\d+|\d+)", words[1])
            gap = rest.lstrip()[len(words[0]):]
            if m and gap[:1].isspace():
                return words[0], m.group(0)
        at = line.find("#", at + 1)
    return None


def read_sqw(path):
    """Read an S(q, w) table in the layout of the reference's SIsotropic
    files (McStas Isotropic_Sqw style), keeping what its loader keeps
    (isotropic.py:171-196):
      - '# <key> <value>' lines set parameters (V_rho, sigma_coh, sigma_abs
        are used);
      - the first other line is the q grid, whose last value is dropped;
      - the second is the w grid;
      - the third is read and discarded;
      - the remaining lines are S rows, one per kept q, w along the row.
    S is returned as float32 plus the float64 machine epsilon rounded to
    float32 (2.220446e-16), as the loader adds it.
    Returns (header dict of strings, q, w, S[q, w])."""
    header = {}
    grids = []
    with open(path, "r") as f:
        lines = f.readlines()
    pos = 0
    for pos, line in enumerate(lines):
        entry = _sqw_header_entry(line)
        if entry:
            header[entry[0]] = entry[1]
        if line.startswith("#"):
            continue
        if len(grids) == 2:
            break   # this line is consumed, not kept
        grids.append(np.array(line.split(), dtype=np.float32))
    q, w = grids[0][:-1], grids[1]
    # data rows: text after '#' is a comment, blank lines are skipped
    rows = [np.array(l.split("#")[0].split(), dtype=np.float64) for l in lines[pos + 1:]
            if l.split("#")[0].strip()]
    sqw = np.array(rows).astype(np.float32) + np.float32(np.finfo(np.float64).eps)
    return header, q, w, sqw


def _sqw_distributions(q, w, sqw):
    """The sampling tables of SIsotropic (isotropic.py:199-207):
      p(w) proportional to the integral over q of q S(q, w), normalised
      over w, and its running CDF over w;
      for each w_j, p(q | w_j) = q S(q, w_j) / S(q) with S(q) the integral
      of S(q, w) over w, and its running CDF over q.
    Returns (pw_cdf[nw], pq_cdf[nw, nq]), float32."""
    per_w = _simps_legacy(q * sqw.T, q, axis=1)       # over q, one per w
    per_w /= _simps_legacy(per_w, w)
    s_of_q = _simps_legacy(sqw.T, w, axis=0)          # over w, one per q
    pw_cdf = _running_cdf(w, per_w)
    pq_cdf = np.array([_running_cdf(q, q * sqw[:, j] / s_of_q) for j in range(len(w))], dtype=np.float32)
    return pw_cdf, pq_cdf


class SIsotropic(_Scat):
    """S(Q,w) isotropic scatterer (scat/isotropic.py:12-212).  Files are read
    by read_sqw and turned into sampling tables by _sqw_distributions."""
    type_code = capi.RAMP_S_ISOTROPIC

    def __init__(self, fn_coh="", fn_inc="", fn_mag="", temperature=0, transmit=1, idx=0,
                 ctx=None, **kwargs):
        super().__init__()
        self.temperature = np.float32(temperature)
        self.transmit = np.uint32(transmit)
        if fn_coh == "" and fn_inc == "" and fn_mag == "":
            raise ValueError("Invalid filename combination")
        self.idx = np.uint32(idx)
        for ch, fn in enumerate((fn_coh, fn_inc, fn_mag)):
            if fn == "":
                # a channel without a file: one-element zero tables (isotropic.py:70-81)
                q = w = pw = pq = np.zeros(1, np.float32)
                nq = nw = 0
                rho = sa = ss = np.float32(0)
            else:
                header, q, w, sqw = read_sqw(fn)
                pw, pq = _sqw_distributions(q, w, sqw)
                pq = pq.ravel()
                nq, nw = len(q), len(w)
                # every channel scatters with its file's sigma_coh (isotropic.py:196)
                rho, sa, ss = (np.float32(header[k]) for k in ("V_rho", "sigma_abs", "sigma_coh"))
            self.p[S.slot("IS_RHO", ch)] = np.float32(rho)
            self.p[S.slot("IS_SABS", ch)] = np.float32(sa)
            self.p[S.slot("IS_SSCAT", ch)] = np.float32(ss)
            self.ni[S.slot("ISI_NQ", ch)] = nq
            self.ni[S.slot("ISI_NW", ch)] = nw
            self.tables[ch] = np.concatenate([np.asarray(a, np.float32).ravel()
                                              for a in (q, w, pw, pq)])
        self.p[S.IS_TEMP] = self.temperature
        self.ni[S.ISI_TRANSMIT] = self.transmit


def _powder_sigma(d_spacing, pack, vc, sigma_abs, multiplicity, DW, F2):
    # powderN.py:59-64 / powder1.py:57-63
    q = np.array((2 * np.pi) / np.array(d_spacing), dtype=np.float32)
    numer = np.array(4 * np.pi ** 3.0 * pack * np.array(multiplicity) * np.array(F2) * 100 * DW,
                     dtype=np.float32)
    denom = vc ** 2.0 * q
    return q, numer / denom, np.float32(pack * sigma_abs / vc * 100)


class SPowder1(_Scat):
    """Single Bragg line (scat/powder1.py:9-81)"""
    type_code = capi.RAMP_S_POWDER1

    def __init__(self, d_spacing=0.0, pack=0.0, vc=0.0, sigma_abs=0.0, multiplicity=0, DW=0.0,
                 F2=0.0, d_phi=180.0, idx=0, ctx=None, **kwargs):
        super().__init__()
        q = (2 * np.pi) / d_spacing
        numer = 4 * np.pi ** 3.0 * pack * multiplicity * F2 * 100 * DW
        denom = vc ** 2.0 * q
        self.sigma_scat_v2 = np.float32(numer / denom)
        self.sigma_abs_v = np.float32(pack * sigma_abs / vc * 100)
        self.q = np.float32(q)
        self.idx = np.uint32(idx)
        self.d_phi = np.float32(d_phi)
        self.p[S.P1_SSCAT_V2], self.p[S.P1_SABS_V] = self.sigma_scat_v2, self.sigma_abs_v
        self.p[S.P1_Q], self.p[S.P1_DPHI] = self.q, self.d_phi


class SPowderN(_Scat):
    """N Bragg lines (scat/powderN.py:9-91).

    Extension: U (Debye-Waller, A^2) multiplies each F2_i by exp(-U Q_i^2),
    Q_i = 2 pi/d_i (no reference counterpart; U = 0 leaves the reference
    behaviour)."""
    type_code = capi.RAMP_S_POWDERN

    def __init__(self, d_spacing=[0.0], pack=0.0, vc=0.0, sigma_abs=0.0, multiplicity=[0], DW=0.0,
                 F2=[0.0], d_phi=180.0, transmit=1, idx=0, ctx=None, U=0.0, **kwargs):
        super().__init__()
        F2 = np.asarray(F2, dtype=np.float64)
        if U:
            F2 = F2 * np.exp(-float(U) * ((2 * np.pi) / np.asarray(d_spacing, np.float64)) ** 2)
        q, ssv2, sav = _powder_sigma(d_spacing, pack, vc, sigma_abs, multiplicity, DW, F2)
        self.q_arr = q
        self.sigma_scat_v2 = ssv2.astype(np.float32)
        self.sigma_abs_v = sav
        self.idx = np.uint32(idx)
        self.d_phi = np.float32(d_phi)
        self.num_lines = np.uint32(len(d_spacing))
        self.transmit = np.uint32(transmit)
        self.p[S.PN_SABS_V], self.p[S.PN_DPHI] = self.sigma_abs_v, self.d_phi
        self.ni[S.PNI_NLINES], self.ni[S.PNI_TRANSMIT] = self.num_lines, self.transmit
        self.tables[S.PNT_SSCAT_V2] = self.sigma_scat_v2
        self.tables[S.PNT_Q] = self.q_arr


_GAUSS_X = np.array([-0.987992518020485, -0.937273392400706, -0.848206583410427,
                     -0.724417731360170, -0.570972172608539, -0.394151347077563,
                     -0.201194093997435, 0, 0.201194093997435, 0.394151347077563,
                     0.570972172608539, 0.724417731360170, 0.848206583410427,
                     0.937273392400706, 0.987992518020485]).astype(np.float32)
_GAUSS_W = np.array([0.030753241996117, 0.070366047488108, 0.107159220467172,
                     0.139570677926154, 0.166269205816994, 0.186161000115562,
                     0.198431485327111, 0.202578241925561, 0.198431485327111,
                     0.186161000115562, 0.166269205816994, 0.139570677926154,
                     0.107159220467172, 0.070366047488108, 0.030753241996117]).astype(np.float32)


class SMonochromator(_Scat):
    """Curved mosaic crystal (scat/monochromator.py:10-111).  The 15-point
    Gauss-Legendre nodes/weights are the standard quadrature table."""
    type_code = capi.RAMP_S_MONO

    def __init__(self, slab_width=0.0, slab_height=0.0, gap=0.0, n_horizontal=1, n_vertical=1,
                 mosaic_horizontal=0.0, mosaic_vertical=0.0, r0=1.0, d_spacing=0.0,
                 radius_vertical=0.0, radius_horizontal=0.0, idx=0, ctx=0, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        vals = (slab_width, slab_height, gap, mosaic_horizontal, mosaic_vertical, r0, d_spacing,
                radius_vertical, radius_horizontal)
        for slot, v in zip((S.MO_SLAB_W, S.MO_SLAB_H, S.MO_GAP, S.MO_MOS_H, S.MO_MOS_V, S.MO_R0,
                            S.MO_D, S.MO_RAD_V, S.MO_RAD_H), vals):
            self.p[slot] = np.float32(v)
        self.ni[S.MOI_NH] = np.uint32(n_horizontal)
        self.ni[S.MOI_NV] = np.uint32(n_vertical)
        self.ni[S.MOI_GLEN] = np.uint32(len(_GAUSS_X))
        self.tables[S.MOT_GX] = _GAUSS_X
        self.tables[S.MOT_GW] = _GAUSS_W


class _Hist(_Scat):
    def _moments(self):
        if self._w is None:
            return (np.zeros(self.n_bins, np.float32), np.zeros(self.n_bins, np.float32))
        return self._w.astype(np.float32), self._w2.astype(np.float32)


class SDetector1D(_Hist):
    """One-axis monitor (scat/detector1D.py:11-138)"""
    type_code = capi.RAMP_S_DET1D

    def __init__(self, binning=(0, 0, 0), logscale=False, restore_neutron=False, idx=0,
                 var="energy", ctx=None, filename=None, **kwargs):
        super().__init__()
        var_dict = {"energy": 0, "theta": 1, "tof": 2, "wavelength": 3}
        self.var_label_dict = {0: "Energy [meV]", 1: "Theta [deg]", 2: "Time-of-flight [us]",
                               3: "Wavelength [Ang]"}
        self.var = np.uint32(var_dict[var])
        self.binning = np.array(binning[:3], dtype=np.float32)
        self.idx = idx
        self.logscale = logscale
        self.num_bins = np.ceil((binning[2] - binning[0]) / binning[1]).astype(np.uint32)
        self.n_bins = int(self.num_bins)
        self.restore = np.uint32(1 if restore_neutron else 0)
        self.filename = filename
        self.axis = np.linspace(self.binning[0], self.binning[2], num=self.num_bins)
        self.p[S.D1_MIN], self.p[S.D1_STEP], self.p[S.D1_MAX] = self.binning
        self.ni[S.D1I_VAR], self.ni[S.D1I_NBINS], self.ni[S.D1I_RESTORE] = (
            self.var, self.num_bins, self.restore)

    def data(self, queue=None):
        h, e = self._moments()
        h[np.isnan(h)] = 0
        e[np.isnan(e)] = 0
        self.histo, self.histo_err = h, e
        return (self.axis, h, e)

    def get_histo(self):
        return self.data()

    def save(self, queue=None):
        axis, h, e = self.data()
        if self.filename:
            np.save(self.filename + "X", axis)
            np.save(self.filename + "Y", h)
            np.save(self.filename + "E", e)

    def plot(self, queue=None):
        import matplotlib.pyplot as plt
        axis, h, e = self.data()
        plt.figure()
        plt.errorbar(axis, np.log(h + 1e-7) if self.logscale else h,
                     0.434 * np.sqrt(e) / np.log(h + 1e-7) if self.logscale else np.sqrt(e),
                     capsize=8)
        plt.ylabel("Intensity")
        plt.xlabel(self.var_label_dict[int(self.var)])
        plt.tight_layout()


class SDetector2D(_Hist):
    """Two-axis monitor (scat/detector2D.py:11-207)"""
    type_code = capi.RAMP_S_DET2D

    var_dict = {"x": 0, "y": 1, "theta": 2, "alpha": 3, "phi": 4, "tof": 5, "divX": 6, "divY": 7,
                "wavelength": 8, "energy": 9, "angle": 10}

    def __init__(self, shape="", axis1_binning=(0, 0, 0), axis2_binning=(0, 0, 0),
                 axis1_var="theta", axis2_var="tof", restore_neutron=False, idx=0, ctx=None,
                 filename=None, logscale=False, **kwargs):
        super().__init__()
        self.axis1_var = np.uint32(self.var_dict[axis1_var])
        self.axis2_var = np.uint32(self.var_dict[axis2_var])
        self.axis1_binning = np.array(axis1_binning[:3], dtype=np.float32)
        self.axis2_binning = np.array(axis2_binning[:3], dtype=np.float32)
        self.idx = np.uint32(idx)
        self.restore_neutron = np.uint32(1 if restore_neutron else 0)
        self.filename = filename
        self.logscale = logscale
        self.axis1_num_bins = np.uint32(np.ceil((axis1_binning[2] - axis1_binning[0]) /
                                                axis1_binning[1]))
        self.axis2_num_bins = np.uint32(np.ceil((axis2_binning[2] - axis2_binning[0]) /
                                                axis2_binning[1]))
        self.num_bins = np.uint32(self.axis1_num_bins * self.axis2_num_bins)
        self.n_bins = int(self.axis1_num_bins) * int(self.axis2_num_bins)
        x = np.linspace(self.axis1_binning[0], self.axis1_binning[2], num=self.axis1_num_bins)
        y = np.linspace(self.axis2_binning[0], self.axis2_binning[2], num=self.axis2_num_bins)
        self.X, self.Y = np.meshgrid(x, y)
        self.p[S.D2_MIN1], self.p[S.D2_STEP1], self.p[S.D2_MAX1] = self.axis1_binning
        self.p[S.D2_MIN2], self.p[S.D2_STEP2], self.p[S.D2_MAX2] = self.axis2_binning
        self.ni[S.D2I_N1], self.ni[S.D2I_N2] = self.axis1_num_bins, self.axis2_num_bins
        self.ni[S.D2I_VAR1], self.ni[S.D2I_VAR2] = self.axis1_var, self.axis2_var
        self.ni[S.D2I_RESTORE] = self.restore_neutron

    def data(self, queue=None):
        h, e = self._moments()
        shape = (int(self.axis1_num_bins), int(self.axis2_num_bins))
        self.Z = h.reshape(shape).T
        self.E = e.reshape(shape).T
        return (self.X, self.Y, self.Z, self.E)

    def get_histo(self):
        return self.data()

    def save(self, queue=None):
        X, Y, Z, E = self.data()
        if self.filename:
            np.save(self.filename + "X", X)
            np.save(self.filename + "Y", Y)
            np.save(self.filename + "Z", Z)
            np.save(self.filename + "E", E)

    def plot(self, queue=None):
        import matplotlib.pyplot as plt
        X, Y, Z, _ = self.data()
        plt.figure()
        plt.pcolormesh(X, Y, np.log(Z + 1e-7) if self.logscale else Z, cmap="jet",
                       shading="auto")
        plt.colorbar()
        plt.tight_layout()


class SCounter(_Hist):
    """Sum of detected weights (scat/counter.py:11-78)"""
    type_code = capi.RAMP_S_COUNTER
    n_bins = 1

    def __init__(self, idx=0, ctx=None, filename=None, **kwargs):
        super().__init__()
        self.idx = idx
        self.filename = filename

    def data(self, queue=None):
        h, _ = self._moments()
        return h[0]

    def save(self, queue=None):
        if self.filename:
            np.save(self.filename, self.data())


class SIncoherent(_Scat):
    """Constant-cross-section incoherent sample with forced single
    scattering and weight correction.  No reference counterpart (SIsotropic
    needs an S(Q,w) file and scales sigma_s by 1/(2k^2)).

    sigma_s: macroscopic scattering cross-section (1/cm);
    sigma_a: macroscopic absorption cross-section at lambda_ref (1/cm),
    scaled linearly with wavelength;
    target, focus_r: optional focusing, as SRescal's (scat/rescal.cl:31-51):
    the outgoing direction is uniform in the cone around `target` (a vector
    in the sample's frame, m) whose half-angle theta_max has
    cos(theta_max) = |target| / sqrt(focus_r^2 + |target|^2), and the weight
    is multiplied by the cone's solid-angle fraction (1 - cos(theta_max))/2,
    so the expected weight into any direction inside the cone is that of
    isotropic scattering.  The cone frame (u, e1, e2) is built here in
    float64: u = target/|target|, e1 along SRescal's perpendicular
    (-t_z, 0, t_x) (or x when t_x = t_z = 0), e2 = u x e1."""
    type_code = capi.RAMP_S_INCOHERENT

    def __init__(self, sigma_s=0.0, sigma_a=0.0, lambda_ref=1.8, target=None, focus_r=0.0,
                 idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        self.p[S.IN_SIGMA_S] = np.float32(sigma_s)
        self.p[S.IN_SIGMA_A] = np.float32(sigma_a)
        self.p[S.IN_LREF] = np.float32(lambda_ref)
        if focus_r and target is not None and np.any(np.asarray(target, dtype=np.float64) != 0):
            t = np.asarray(target, dtype=np.float64)[:3]
            l2 = float(np.dot(t, t))
            r2 = float(focus_r) ** 2
            u = t / math.sqrt(l2)
            perp = np.array([1.0, 0.0, 0.0]) if (t[0] == 0 and t[2] == 0) else \
                np.array([-t[2], 0.0, t[0]])
            e1 = perp / np.linalg.norm(perp)
            e2 = np.cross(u, e1)
            cosmax = math.sqrt(l2 / (r2 + l2))
            omc = (r2 / (r2 + l2)) / (1.0 + cosmax)
            self.p[S.IN_FOCUS_R] = np.float32(focus_r)
            self.p[S.IN_U:S.IN_U + 3] = u.astype(np.float32)
            self.p[S.IN_E1:S.IN_E1 + 3] = e1.astype(np.float32)
            self.p[S.IN_E2:S.IN_E2 + 3] = e2.astype(np.float32)
            self.p[S.IN_OMC] = np.float32(omc)
            self.p[S.IN_WF] = np.float32(0.5 * omc)


class SPowderDiff(_Scat):
    """Powder sample for gradients: Bragg lines (cross-sections as SPowderN,
    scat/powderN.py:59-64, with the Debye-Waller factor exp(-U q^2)),
    incoherent scattering sigma_inc and absorption sigma_abs (1/cm, the
    latter at lambda_ref, scaling with wavelength), and a Gaussian spread
    of each line's d-spacing of relative width `strain`.  No reference
    counterpart.

    Every random choice is drawn from a fixed proposal - the same sample
    at `proposal` (a dict of any of d_spacing, F2, U, sigma_inc,
    sigma_abs; default: the model itself) - and the weight carries the
    exact ratio of model to proposal probabilities: depth of the single
    scattering (truncated exponential), channel (incoherent or line j),
    and the grain's d-spacing.  Trajectories therefore do not depend on the
    model parameters, the weight is a smooth function of them, and the
    gradient of any histogram loss is the sum over detector events of
    dL/dh_bin * dw/dtheta (ramp_adjoint).  Scattering is forced (single
    scattering with weight correction, no exit attenuation); a line is
    active while the proposal's q0 < 2k."""
    type_code = capi.RAMP_S_POWDERDIFF
    N_LW = 8

    def __init__(self, d_spacing=(2.0,), F2=(1.0,), multiplicity=None, pack=1.0, vc=1.0, DW=1.0,
                 U=0.0, sigma_inc=0.0, sigma_abs=0.0, lambda_ref=1.8, strain=1e-3, proposal=None,
                 idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        d = np.asarray(d_spacing, dtype=np.float64)
        F2 = np.asarray(F2, dtype=np.float64)
        mult = np.ones_like(d) if multiplicity is None else np.asarray(multiplicity, np.float64)
        prop = dict(proposal or {})
        d0 = np.asarray(prop.get("d_spacing", d), dtype=np.float64)
        F20 = np.asarray(prop.get("F2", F2), dtype=np.float64)
        U0 = float(prop.get("U", U))
        if not (len(d) == len(F2) == len(mult) == len(d0) == len(F20)) or len(d) == 0:
            raise ValueError("SPowderDiff: d_spacing, F2, multiplicity and the proposal's lines must "
                             "have the same non-zero length")
        if strain <= 0:
            raise ValueError("SPowderDiff: strain must be > 0")
        self.model = {"d_spacing": d, "F2": F2, "U": float(U), "sigma_inc": float(sigma_inc),
                      "sigma_abs": float(sigma_abs)}
        self.kw = {"multiplicity": mult, "pack": pack, "vc": vc, "DW": DW, "lambda_ref": lambda_ref,
                   "strain": strain}
        self.prop = {"d_spacing": d0, "F2": F20, "U": U0,
                     "sigma_inc": float(prop.get("sigma_inc", sigma_inc)),
                     "sigma_abs": float(prop.get("sigma_abs", sigma_abs))}

        def ssv2(dd, ff, uu):
            q = 2 * np.pi / dd
            return _powder_sigma(dd, pack, vc, 0.0, mult, DW, ff * np.exp(-uu * q * q))[1]

        q = 2 * np.pi / d
        lines = np.zeros((len(d), self.N_LW), dtype=np.float32)
        lines[:, S.PDL_SSV2] = ssv2(d, F2, float(U))
        lines[:, S.PDL_SSV2_0] = ssv2(d0, F20, U0)
        lines[:, S.PDL_Q0] = (2 * np.pi / d0).astype(np.float32)
        lines[:, S.PDL_D] = d.astype(np.float32)
        lines[:, S.PDL_D0] = d0.astype(np.float32)
        lines[:, S.PDL_DLS] = (1.0 / d + 2.0 * float(U) * q * q / d).astype(np.float32)
        lines[:, S.PDL_Q2] = (q * q).astype(np.float32)
        self.lines = lines
        self.num_lines = np.uint32(len(d))
        self.p[S.PD_SINC] = np.float32(sigma_inc)
        self.p[S.PD_SABS] = np.float32(sigma_abs)
        self.p[S.PD_LREF] = np.float32(lambda_ref)
        self.p[S.PD_U] = np.float32(U)
        self.p[S.PD_STRAIN] = np.float32(strain)
        self.p[S.PD_SINC0] = np.float32(self.prop["sigma_inc"])
        self.p[S.PD_SABS0] = np.float32(self.prop["sigma_abs"])
        self.ni[S.PDI_NLINES] = self.num_lines
        self.tables[S.PDT_LINES] = lines.reshape(-1)

    def param_names(self):
        return ["sigma_inc", "sigma_abs", "U"] + [f"d{j}" for j in range(int(self.num_lines))]


class SAbsorber(_Scat):
    """Pure Beer-Lambert absorber (weight *= exp(-sigma_a L)).  No reference
    counterpart."""
    type_code = capi.RAMP_S_ABSORBER

    def __init__(self, sigma_a=0.0, idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        self.p[S.AB_SIGMA_A] = np.float32(sigma_a)


class SDetectorTubes(_Hist):
    """Position-sensitive tube bank: n_tubes along the scattering angle
    (degrees, atan2(x, z) of the hit), n_pixels along y, n_tof TOF bins
    (microseconds).  data() returns the (n_tubes, n_pixels, n_tof) weight
    and weight^2 arrays.  No reference counterpart."""
    type_code = capi.RAMP_S_DETPIX

    def __init__(self, theta_range=(-40.0, 140.0), y_range=(-2.0, 2.0), tof_range=(10000.0, 60000.0),
                 n_tubes=768, n_pixels=128, n_tof=1000, restore_neutron=False, idx=0, ctx=None,
                 filename=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        self.shape = (int(n_tubes), int(n_pixels), int(n_tof))
        self.n_bins = int(n_tubes) * int(n_pixels) * int(n_tof)
        self.filename = filename
        self.p[S.DT_THMIN], self.p[S.DT_THMAX] = np.float32(theta_range[0]), np.float32(theta_range[1])
        self.p[S.DT_YMIN], self.p[S.DT_YMAX] = np.float32(y_range[0]), np.float32(y_range[1])
        self.p[S.DT_TMIN], self.p[S.DT_TMAX] = np.float32(tof_range[0]), np.float32(tof_range[1])
        self.ni[S.DTI_NTUBE], self.ni[S.DTI_NPIX], self.ni[S.DTI_NTOF] = self.shape
        self.ni[S.DTI_RESTORE] = 1 if restore_neutron else 0

    def data(self, queue=None):
        h, e = self._moments()
        return (h.reshape(self.shape), e.reshape(self.shape))

    def save(self, queue=None):
        if self.filename:
            h, e = self.data()
            np.save(self.filename + "Z", h)
            np.save(self.filename + "E", e)


class SPolariser(_Scat):
    """scat/polariser.py:6-47 + polariser.cl:3-32: an ideal polariser - the
    neutron's polarisation becomes `polarisation`."""
    type_code = capi.RAMP_S_POLARISER

    def __init__(self, polarisation=(0, 0, 0), idx=0, ctx=None, filename=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        for k in range(3):
            self.p[S.PO_P + k] = np.float32(polarisation[k])


class SAnalyser(SPolariser):
    """scat/analyser.py + analyser.cl:1-31 (which lacks its consts.h
    include): an ideal analyser - the weight is multiplied by
    (1 + P.pol)/2 and the polarisation becomes P.  Plans holding one run in
    full-state mode, which tracks the polarisation."""
    type_code = capi.RAMP_S_ANALYSER


class SLinearCollimator(_Scat):
    """scat/collimator_lin.py:7-63 + collimator_lin.cl:3-55 (which lacks a
    ';' at :30): neutrons whose |vx/vz| or |vy/vz| exceeds tan(divergence)
    (arcminutes) die, the others are weighted by transmission*(1 - phi/slope)
    per direction."""
    type_code = capi.RAMP_S_COLLIMATOR

    def __init__(self, length=0.0, divergence_H=0.0, divergence_V=0.0, transmission=1.0, idx=0, ctx=0,
                 **kwargs):
        super().__init__()

        def min2rad(x):
            return x * np.pi / (60.0 * 180.0)

        self.idx = np.uint32(idx)
        self.length = np.float32(length)
        self.slope_H = np.float32(np.tan(min2rad(divergence_H)))
        self.slope_V = np.float32(np.tan(min2rad(divergence_V)))
        self.transmission = np.float32(transmission)
        self.p[S.CO_LEN], self.p[S.CO_SLOPE_H] = self.length, self.slope_H
        self.p[S.CO_SLOPE_V], self.p[S.CO_TRANS] = self.slope_V, self.transmission


class SRescal(_Scat):
    """scat/rescal.cl:5-77 + rescal.py:9-66: scatters at a random depth
    into the cone of radius focus_r about `target` with a final energy
    uniform in E0 +- dE, and records the event (Q = ki - kf, energy
    transfer, weight) of every history that ends its block detected
    (se > 0, as an SCounter sets it).  data() returns (qx, qy, qz, dE, p),
    accumulated over all chunks and runs as the reference's lists are.
    The events stay on the GPU and are appended there after each chunk;
    they cross to the host only when data() is read.  The component must
    sit in the instrument's last block."""
    type_code = capi.RAMP_S_RESCAL

    def __init__(self, target=(0, 0, 0), E0=0, dE=0, focus_r=0, idx=0, ctx=None, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        for k in range(3):
            self.p[S.RS_T + k] = np.float32(target[k])
        self.E0 = np.float32(E0)
        self.dE = np.float32(dE)
        self.focus_r = np.float32(focus_r)
        self.p[S.RS_E0], self.p[S.RS_DE], self.p[S.RS_FOCUS_R] = self.E0, self.dE, self.focus_r
        self._events = tuple(np.zeros(0, np.float64) for _ in range(5))

    def _collect(self, lib, plan, slot):
        import ctypes as C
        cnt = C.c_uint64(0)
        capi.check(lib.ramp_rescal_read(plan, slot, None, 0, C.byref(cnt)))
        n = cnt.value
        ev = np.zeros((5, max(n, 1)), np.float32)
        if n:
            capi.check(lib.ramp_rescal_read(plan, slot, ev.ctypes.data_as(C.POINTER(C.c_float)), n,
                                            C.byref(cnt)))
        qx, qy, qz, dE, w = (ev[k, :n].astype(np.float64) for k in range(5))
        self._events = (qx, qy, qz, dE, w)

    def data(self, queue=None):
        return self._events


class SMCPLIn(_Scat):
    """scat/mcpl_in.py:35-76: the particles of an MCPL file (format 3,
    optionally gzipped) replace the chunk's neutrons - slot i takes
    particle i, the slots beyond the file's particles are dead - whatever
    their state, iidx or intersection, which it leaves as they are.  The
    file is parsed here (ramp_amd.mcpl); each GPU lane unpacks its own
    record in FP64 as MCPL_to_RAMP does (:9-33).  The particle's RNG counter
    is float32(r + i) with r a 30-bit random integer (:10), here the
    Threefry stream of key i at counter 2^31 + chunk seed.  Extension: with
    more particles in the file than histories in the chunk, the first
    `chunk` particles are used (the reference's loop, :66-69, overruns its
    buffer)."""
    type_code = capi.RAMP_S_MCPL_IN

    def __init__(self, filename="", idx=0, ctx=0, **kwargs):
        super().__init__()
        from .mcpl import read_mcpl
        self.idx = np.uint32(idx)
        self.filename = filename
        self.header, rec = read_mcpl(filename)
        n = rec.shape[0]
        if n > 0xffffffff:
            raise ValueError(f"{filename}: {n} particles, at most 2^32-1 are supported")
        self.n_particles = n
        self.max_weight = float(np.max(np.abs(rec[:, 10]))) if n else 0.0
        self.ni[S.MCI_N] = n
        self.ni[S.MCI_DOUBLE] = 0 if rec.dtype == np.float32 else 1
        # field-major (each lane of a wave then reads contiguous words);
        # doubles travel in the float32 table pool as pairs of words
        self.tables[S.MCT_REC] = np.ascontiguousarray(rec.T).view(np.float32).ravel()


class SMCPLOut(_Scat):
    """scat/mcpl_out.py:75-117: the neutrons alive at this component are
    written to an MCPL file - the reference's header and 44-byte records
    (polarisation, position in cm, pack_vector's direction with the kinetic
    energy in MeV, time in ms, weight) in history order - leaving the state,
    iidx and intersection unchanged.  As in the reference the file holds
    the last chunk of the last execute().  The records are packed on the GPU
    and compacted there (ramp_mcpl_read).  Extension: standard=True packs the
    unit direction as MCPL itself does (the reference's pack_vector keeps the
    velocity's components, not the unit vector's, and zeros where MCPL keeps
    one: mcpl_out.py:36-50)."""
    type_code = capi.RAMP_S_MCPL_OUT

    def __init__(self, filename="", standard=False, idx=0, ctx=0, **kwargs):
        super().__init__()
        self.idx = np.uint32(idx)
        self.filename = filename
        self.ni[S.MCO_STANDARD] = 1 if standard else 0
        self.records = np.zeros((0, 11), np.float32)

    def _collect(self, lib, plan, slot):
        import ctypes as C
        from .mcpl import write_mcpl
        cnt = C.c_uint64(0)
        capi.check(lib.ramp_mcpl_read(plan, slot, None, 0, C.byref(cnt)))
        rec = np.zeros((cnt.value, 11), np.float32)
        capi.check(lib.ramp_mcpl_read(plan, slot, rec.ctypes.data_as(C.POINTER(C.c_float)),
                                      cnt.value, C.byref(cnt)))
        self.records = rec
        if self.filename:
            write_mcpl(self.filename, rec)


SOURCES = ("MFlat", "MCont", "MISIS", "MLambda", "MParallel")
GEOMS = ("GArm", "GPlane", "GBox", "GSphere", "GSphereInt", "GCylinderExt", "GBanana", "GMesh")
SCATS = ("SArm", "SBeamstop", "SChopper", "SGuide", "SIsotropic", "SPowder1", "SPowderN",
         "SMonochromator", "SDetector1D", "SDetector2D", "SCounter", "SIncoherent", "SAbsorber",
         "SDetectorTubes", "SPowderDiff", "SMCPLIn", "SMCPLOut", "SRescal", "SPolariser", "SAnalyser",
         "SLinearCollimator")
