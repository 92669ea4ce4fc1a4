"""Instrument: the drop-in for mcramp.Instrument (mcramp/mcramp.py:254-455).

Reads the same JSON instrument description ($expr$ substitution, // comments,
linear and non-linear 'multi' blocks, position/rotation/relative, disable,
restore_neutron), resolves components by the same names, numbers them the
same way, and runs the whole instrument as one fused plan on the GPU through
libramp.so.  ctx and queue are accepted and ignored.

Additions: seed= (per-chunk source seeds become reproducible), chunk= (the
chunk size, which fixes the 1/chunk weight normalisation explicitly),
hist_mode= ("atomic" or "deterministic"), device=.
"""
import ctypes as C
import importlib
import json
import math
import random
import re
from collections import OrderedDict

import numpy as np

from . import _capi as capi
from . import _slots as S
from . import components as comps_mod
from .components import TablePool

# Notional largest single allocation used for the reference's chunk rule
# buf_max = MAX_MEM_ALLOC_SIZE/8/32 (mcramp.py:415-420): 32 GiB on MI355X.
MAX_ALLOC = 1 << 35


def frame_rotate(vec, rot):
    """Host Euler rotation (mcramp.py:457-476), numpy arithmetic as written."""
    x_a, y_a, z_a = rot[0], rot[1], rot[2]
    r00 = np.cos(y_a) * np.cos(z_a)
    r01 = -np.cos(x_a) * np.sin(z_a) + np.sin(x_a) * np.sin(y_a) * np.cos(z_a)
    r02 = np.sin(x_a) * np.sin(z_a) + np.cos(x_a) * np.sin(y_a) * np.cos(z_a)
    r10 = np.cos(y_a) * np.sin(z_a)
    r11 = np.cos(x_a) * np.cos(z_a) + np.sin(x_a) * np.sin(y_a) * np.sin(z_a)
    r12 = -np.sin(x_a) * np.cos(z_a) + np.cos(x_a) * np.sin(y_a) * np.sin(z_a)
    r20 = -np.sin(y_a)
    r21 = np.sin(x_a) * np.cos(y_a)
    r22 = np.cos(x_a) * np.cos(y_a)
    return [np.dot([r00, r01, r02], vec), np.dot([r10, r11, r12], vec),
            np.dot([r20, r21, r22], vec)]


def _rotate_points(pts, rot):
    """frame_rotate (mcramp.py:457-476) applied to an (n, 3) float32 array
    of points: the matrix formed from the float32 angles as frame_rotate
    forms it, each component summed in float32 in order x, y, z."""
    x_a, y_a, z_a = np.float32(rot[0]), np.float32(rot[1]), np.float32(rot[2])
    rows = frame_rotate(np.eye(3, dtype=np.float32), [x_a, y_a, z_a])
    R = np.asarray(rows, dtype=np.float32)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    out = np.empty_like(pts)
    for i in range(3):
        out[:, i] = (R[i, 0] * x + R[i, 1] * y) + R[i, 2] * z
    return out


def _resolve(name):
    """Class by name: the package namespace first, then a module of that
    name (mcramp.py:80-92)."""
    if hasattr(comps_mod, name):
        return getattr(comps_mod, name)
    mod = importlib.import_module(name)
    return getattr(mod, name)


class _Comp:
    def __init__(self, name, geom, scat, restore, pos, rot, vis, idx, block):
        self.name = name
        self.geom = geom
        self.scat = scat
        self.restore = restore
        self.pos = np.array(pos, dtype=np.float32)
        self.rot = np.array(rot, dtype=np.float32)
        self.vis = vis
        self.idx = idx
        self.block = block


class ExecutionBlock:
    """One block of the instrument (mcramp.py:35-116)."""

    def __init__(self, source, components, linear, max_events):
        self.source = source
        self.components = components
        self.linear = linear
        self.max_events = max_events
        self.trace_lines = []

    @staticmethod
    def fromJSON(block, parent, linear, events, bidx):
        comps = OrderedDict()
        source = None
        for (name, comp) in block.items():
            if name == "linear" or name == "multi":
                continue
            if "disable" in comp and comp["disable"]:
                continue
            if "source" in comp:
                mk = _resolve(comp["moderator_kernel"]["name"])
                args = {k: v for (k, v) in comp["moderator_kernel"].items() if k != "name"}
                args["ctx"] = None
                source = mk(**args)
            else:
                pos = comp["position"] if "position" in comp else [0, 0, 0]
                rot = comp["rotation"] if "rotation" in comp else [0, 0, 0]
                vis = comp["visualise"] if "visualise" in comp else True
                restore = bool(comp["restore_neutron"]) if "restore_neutron" in comp else False
                if "relative" in comp:
                    rc = comps[comp["relative"]]
                    pos = frame_rotate(pos, [rc.rot[0], rc.rot[1], rc.rot[2]])
                    pos = np.add(pos, [rc.pos[0], rc.pos[1], rc.pos[2]])
                    rot = np.add(rot, [rc.rot[0], rc.rot[1], rc.rot[2]])
                gk = _resolve(comp["geom_kernel"]["name"])
                gargs = {k: v for (k, v) in comp["geom_kernel"].items() if k != "name"}
                gargs["idx"] = parent.idx
                gargs["ctx"] = None
                sk = _resolve(comp["scat_kernel"]["name"])
                sargs = {k: v for (k, v) in comp["scat_kernel"].items() if k != "name"}
                sargs["idx"] = parent.idx
                sargs["ctx"] = None
                sargs["inst"] = parent
                sargs["block"] = bidx
                comps[name] = _Comp(name, gk(**gargs), sk(**sargs), restore, pos, rot, vis,
                                    parent.idx, bidx)
                parent.kernel_refs.append(comps[name])
            parent.idx += 1
        return ExecutionBlock(source, comps, linear, events)


class Context:
    """A libramp device context (one HIP device, one stream) that several
    Instruments share when it is passed as their ctx argument (extension:
    the reference's ctx is a PyOpenCL context, which is accepted and
    ignored).  Batches need their instruments on one Context."""

    def __init__(self, device=0):
        lib = capi.load_library()
        h = C.c_void_p()
        capi.check(lib.ramp_ctx_create(int(device), C.byref(h)))
        self.device = int(device)
        self.handle = h

    def close(self):
        if getattr(self, "handle", None) is not None:
            from .plugins import release
            release(self.handle)
            capi.load_library().ramp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Instrument:
    """Drop-in for mcramp.Instrument(fn, ctx, queue, **kwargs).

    Parameters beyond the reference: seed (int) makes the per-chunk source
    seeds reproducible (default: random.randint(0, 4096) per chunk, as the
    reference); chunk fixes the chunk size; hist_mode selects float-atomic
    or deterministic fixed-point histograms; device selects the HIP device.
    """

    def __init__(self, fn, ctx=None, queue=None, seed=None, chunk=None, hist_mode="atomic",
                 device=0, **kwargs):
        self.ctx = ctx
        self.queue = queue
        self.kernel_refs = []
        self.idx = 1
        self.seed = seed
        self.chunk = chunk
        self.hist_mode = hist_mode
        self.device = device
        self.blocks = self._fromJSON(fn, **kwargs)
        self._plan = None
        self._gctx = None
        self._desc_keep = None
        self._seed_rng = np.random.default_rng(seed) if seed is not None else None
        self.last_stats = None
        self._trace_pts = OrderedDict()

    # ---- JSON --------------------------------------------------------------
    def _substitute_params(self, json_str, **kwargs):
        # mcramp.py:328-355 (substring replacement, then eval, then // strip)
        pattern = r"(?:\$)(.*?)(?:\$)"
        tokens = re.findall(pattern, json_str)
        env = {"np": np, "math": math, "__builtins__": {"abs": abs, "min": min, "max": max,
                                                         "float": float, "int": int,
                                                         "round": round, "True": True,
                                                         "False": False}}
        for t in tokens:
            subbed = t
            for key, value in kwargs.items():
                subbed = subbed.replace("{}".format(key), str(value))
            json_str = json_str.replace("${}$".format(t), str(eval(subbed, env)))
        lines = json_str.split("\n")
        for i, line in enumerate(lines):
            if "//" in line:
                lines[i] = line.split("//", 1)[0]
        return "\n".join(lines)

    def _fromJSON(self, fn, **kwargs):
        with open(fn, "r") as f:
            json_str = self._substitute_params(f.read(), **kwargs)
        inst = json.loads(json_str, object_pairs_hook=OrderedDict)
        blocks = []
        for i, block in enumerate(inst.values()):
            if "linear" in block:
                linear = block["linear"]
                events = block["multi"] if (not linear and "multi" in block) else 1
            else:
                linear, events = True, 1
            blocks.append(ExecutionBlock.fromJSON(block, self, linear, events, i))
        return blocks

    # ---- plan ---------------------------------------------------------------
    @property
    def source(self):
        for b in self.blocks:
            if b.source is not None:
                return b.source
        raise ValueError("instrument has no source component")

    def plan_desc(self):
        """The instrument as a capi.PlanDesc (keeps its buffers alive)."""
        pool = TablePool()
        src_block = next(i for i, b in enumerate(self.blocks) if b.source is not None)
        comp_list, block_list, det_list = [], [], []
        for b in self.blocks:
            first = len(comp_list)
            for c in b.components.values():
                comp_list.append(c)
            block_list.append((1 if b.linear else 0, int(b.max_events), first,
                               len(comp_list) - first))
        cds = (capi.ComponentDesc * max(1, len(comp_list)))()
        self._det_of = {}
        self._mcpl_out = []
        self._rescal = []
        for c in comp_list:
            if c.scat.type_code == capi.RAMP_S_MCPL_OUT:
                c.scat.ni[S.MCO_SLOT] = len(self._mcpl_out)
                self._mcpl_out.append(c)
            if c.scat.type_code == capi.RAMP_S_RESCAL:
                c.scat.ni[S.RSI_SLOT] = len(self._rescal)
                self._rescal.append(c)
        for k, c in enumerate(comp_list):
            d = cds[k]
            c.geom.fill(d.geom, pool)
            c.scat.fill(d.scat, pool)
            d.comp_idx = int(c.idx)
            d.restore = 1 if c.restore else 0
            d.vis = 1 if c.vis else 0
            if getattr(c.scat, "n_bins", 0) > 0:
                d.det = len(det_list)
                self._det_of[k] = d.det
                det_list.append(int(c.scat.n_bins))
            else:
                d.det = -1
            for i in range(3):
                d.pos[i] = float(c.pos[i])
                d.rot[i] = float(c.rot[i])
        bds = (capi.BlockDesc * max(1, len(block_list)))()
        for k, (lin, multi, first, count) in enumerate(block_list):
            bds[k].linear, bds[k].multi, bds[k].first, bds[k].count = lin, multi, first, count
        dds = (capi.DetDesc * max(1, len(det_list)))()
        for k, nb in enumerate(det_list):
            dds[k].n_bins = nb
        plan = capi.PlanDesc()
        self.source.fill(plan.source, pool)
        plan.source_block = src_block
        plan.n_blocks = len(block_list)
        plan.n_comps = len(comp_list)
        plan.n_dets = len(det_list)
        plan.blocks = C.cast(bds, C.POINTER(capi.BlockDesc))
        plan.comps = C.cast(cds, C.POINTER(capi.ComponentDesc))
        plan.dets = C.cast(dds, C.POINTER(capi.DetDesc))
        tables = pool.array()
        plan.tables = tables.ctypes.data_as(C.POINTER(C.c_float))
        plan.n_tables = tables.size
        self._comp_list = comp_list
        self._det_bins = det_list
        self._desc_keep = (plan, cds, bds, dds, tables)
        return plan

    def weight_bound(self, chunk):
        """Upper bound of one history's weight at a detector (sets the
        fixed-point scale of deterministic histograms): the source's, or
        the largest particle weight of an SMCPLIn file if that is larger."""
        wb = self.source.weight_bound(chunk)
        for b in self.blocks:
            for c in b.components.values():
                if hasattr(c.scat, "max_weight"):
                    wb = max(wb, c.scat.max_weight)
        return wb

    def chunking(self, N, buf_max_mod=None):
        """Chunk size of a run of N histories (mcramp.py:415-420, 443-445)."""
        if self.chunk:
            return int(min(self.chunk, N))
        buf_max = int(MAX_ALLOC / 8 / 32 / buf_max_mod) if buf_max_mod else int(MAX_ALLOC / 8 / 32)
        return int(min(buf_max, N))

    def chunk_seeds(self, n_chunks):
        if self._seed_rng is not None:
            return np.asarray(self._seed_rng.integers(0, 4097, size=n_chunks), dtype=np.uint32)
        return np.array([random.randint(0, 4096) for _ in range(n_chunks)], dtype=np.uint32)

    def _ctx_handle(self):
        if isinstance(self.ctx, Context):
            return self.ctx.handle
        return self._gctx

    def _ensure_plan(self):
        lib = capi.load_library()
        if self._ctx_handle() is None:
            h = C.c_void_p()
            capi.check(lib.ramp_ctx_create(int(self.device), C.byref(h)))
            self._gctx = h
        if self._plan is None:
            desc = self.plan_desc()
            h = C.c_void_p()
            capi.check(lib.ramp_plan_create(self._ctx_handle(), C.byref(desc), C.byref(h)))
            self._plan = h
            for k, c in enumerate(self._comp_list):
                if c.scat.type_code == capi.RAMP_S_PLUGIN:
                    from .plugins import build_kernel
                    pl = build_kernel(c.scat.plugin_source(), c.scat.entry, self._ctx_handle())
                    capi.check(lib.ramp_plan_bind_plugin(self._plan, k, pl))
        return lib

    # ---- execution ----------------------------------------------------------
    def execute(self, N, buf_max_mod=None, debug=False, trace=False, keep_state=False,
                seeds=None, mesh_brute_force=False, wg=0, npt=0, shard=None):
        """Run N histories (mcramp.py:403-455).

        Extensions: `seeds` (one per chunk), `keep_state` (keep the final
        per-neutron state of the last chunk), `mesh_brute_force` (test every
        triangle of every GMesh as the reference does instead of traversing
        the BVH; for checking the BVH), `wg`/`npt` (launch shape of the
        source segment; 0 = library default).

        shard=(first, total) runs histories [first, first + N) of a single
        chunk of `total` (keys and 1/total weights as in the whole chunk), so
        that shards run in turn add up to the whole chunk's histograms
        (bit for bit in deterministic mode).

        trace=True appends, for every neutron alive after each drawn
        component of a linear block, the line from its previous point to
        the component's intersection entry point in the lab frame to that
        block's trace_lines (mcramp.py:118-165, 222-251); the points are
        recorded on the GPU (RAMP_RUN_TRACE)."""
        if trace:
            for b in self.blocks:
                if not b.linear and any(c.vis for c in b.components.values()):
                    raise ValueError("trace: a non-linear block holds drawn components; the "
                                     "reference keeps trace positions for linear blocks only "
                                     "(mcramp.py:118-122, 209-211); set 'visualise': false on them")
        lib = self._ensure_plan()
        N = int(N)
        chunk = N if shard is not None else self.chunking(N, buf_max_mod)
        n_chunks = (N + chunk - 1) // chunk if N else 0
        if seeds is None:
            seeds = self.chunk_seeds(n_chunks)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        opts = capi.RunOpts()
        opts.hist_mode = (capi.RAMP_HIST_DETERMINISTIC if self.hist_mode == "deterministic"
                          else capi.RAMP_HIST_ATOMIC)
        opts.keep_state = 1 if (keep_state or debug) else 0
        opts.weight_bound = self.weight_bound(int(shard[1]) if shard is not None else chunk) if N else 0.0
        if shard is not None:
            opts.key_base = int(shard[0])
            opts.shard_total = int(shard[1])
        opts.flags = ((capi.RAMP_RUN_MESH_BRUTE_FORCE if mesh_brute_force else 0)
                      | (capi.RAMP_RUN_TRACE if trace else 0))
        opts.wg = int(wg)
        opts.npt = int(npt)
        capi.check(lib.ramp_run(self._plan, N, chunk,
                                seeds.ctypes.data_as(C.POINTER(C.c_uint32)), len(seeds),
                                C.byref(opts)))
        self.last_seeds = seeds
        self.last_chunk = chunk
        self.last_stats = self.stats()
        self._hist_stale = True
        if N:
            # SMCPLOut writes its file at every chunk (mcpl_out.py:99-117); the
            # last chunk's records are what remains
            for slot, c in enumerate(self._mcpl_out):
                c.scat._collect(lib, self._plan, slot)
        if trace and N:
            self._trace_lines(lib, N, chunk)
        if debug and N:
            # the library keeps the last chunk's states: N - (chunks - 1) * chunk
            n16, _, _ = self.state(N - (-(-N // chunk) - 1) * chunk)
            live = n16[n16[:, 15] == 0.0]
            print(live)

    def trace_points(self):
        """Trace points of the last traced run (extension): {component
        name: (xyz, alive)} with xyz the intersection entry point in the
        component frame, float32 of shape (N, 3), and alive a bool array of
        shape (N,), for every drawn component of a linear block."""
        return dict(self._trace_pts)

    def _trace_lines(self, lib, N, chunk):
        """Read the GPU trace points and form the reference's trace lines
        (_get_trace_lines, mcramp.py:222-251): per chunk and per linear
        block, every neutron starts at the origin (prev_locs is zeroed when
        the block starts, :121-122); after each drawn component each live
        neutron gets a line from its previous point to the lab-frame entry
        point frame_rotate(x1, y1, z1) + pos, in index order, and that
        point becomes its previous point."""
        pts = OrderedDict()
        slot = 0
        for b in self.blocks:
            if not b.linear:
                continue
            for c in b.components.values():
                if not c.vis:
                    continue
                buf = np.empty(4 * N, dtype=np.float32)
                capi.check(lib.ramp_trace_read(self._plan, slot, buf.ctypes.data_as(C.POINTER(C.c_float)),
                                               N))
                buf = buf.reshape(4, N)
                pts[c.name] = (np.ascontiguousarray(buf[:3].T), buf[3] != 0.0)
                slot += 1
        self._trace_pts = pts
        for lo in range(0, N, chunk):
            hi = min(N, lo + chunk)
            for b in self.blocks:
                if not b.linear:
                    continue
                prev = np.zeros((hi - lo, 3), dtype=np.float32)
                for c in b.components.values():
                    if not c.vis:
                        continue
                    xyz, alive = pts[c.name]
                    idx = np.nonzero(alive[lo:hi])[0]
                    loc = xyz[lo:hi][idx]
                    lab = _rotate_points(loc, c.rot) + c.pos
                    p0 = prev[idx].tolist()
                    p1 = lab.tolist()
                    b.trace_lines.extend([[a[0], q[0]], [a[1], q[1]], [a[2], q[2]]]
                                         for a, q in zip(p0, p1))
                    prev[idx] = lab

    def _pull_hists(self):
        if self._plan is None:
            return
        self._hist_stale = False
        lib = capi.load_library()
        for slot, c in enumerate(getattr(self, "_rescal", [])):
            c.scat._collect(lib, self._plan, slot)
        for k, c in enumerate(self._comp_list):
            if k in self._det_of:
                d = self._det_of[k]
                nb = self._det_bins[d]
                w = np.zeros(nb, dtype=np.float64)
                w2 = np.zeros(nb, dtype=np.float64)
                capi.check(lib.ramp_hist_read(self._plan, d,
                                              w.ctypes.data_as(C.POINTER(C.c_double)),
                                              w2.ctypes.data_as(C.POINTER(C.c_double)), nb))
                c.scat._set_hist(w, w2)

    def _sync_hists(self):
        if getattr(self, "_hist_stale", False):
            self._pull_hists()

    def hist64(self, name):
        """(w, w^2) float64 sums of a detector component (extension)."""
        self._sync_hists()
        for c in self._comp_list:
            if c.name == name:
                return c.scat._w, c.scat._w2
        raise KeyError(name)

    def state(self, n):
        """Final per-neutron state of the last chunk (reference AoS layouts)."""
        lib = capi.load_library()
        n16 = np.zeros((n, 16), dtype=np.float32)
        i8 = np.zeros((n, 8), dtype=np.float32)
        ii = np.zeros(n, dtype=np.uint32)
        capi.check(lib.ramp_state_read(self._plan, n16.ctypes.data_as(C.POINTER(C.c_float)),
                                       i8.ctypes.data_as(C.POINTER(C.c_float)),
                                       ii.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return n16, i8, ii

    def stats(self):
        lib = capi.load_library()
        st = capi.RunStats()
        capi.check(lib.ramp_stats_read(self._plan, C.byref(st)))
        return {f: getattr(st, f) for f, _ in capi.RunStats._fields_}

    def reset(self):
        """Zero all histograms (extension; the reference accumulates)."""
        if self._plan is not None:
            capi.check(capi.load_library().ramp_hist_reset(self._plan))
            self._hist_stale = True   # read (as zeros) when next asked for

    # ---- results ------------------------------------------------------------
    def data(self):
        """{component name: data} (mcramp.py:281-293)."""
        self._sync_hists()
        return {c.name: c.scat.data(self.queue) for c in self.kernel_refs}

    def plot(self):
        from matplotlib.pyplot import show
        self._sync_hists()
        for c in self.kernel_refs:
            c.scat.plot(self.queue)
        show()

    def save(self):
        self._sync_hists()
        for c in self.kernel_refs:
            c.scat.save(self.queue)

    def visualise(self, controls=True, xlim=None, ylim=None, zlim=None, focus=None, **kwargs):
        """Figure of the instrument's drawn components and the trace lines of
        traced runs (mcramp.py:317-325); returns the Visualisation
        (extension) after showing it."""
        from .visualisation import Visualisation
        vis = Visualisation(self, controls=controls, xlim=xlim, ylim=ylim, zlim=zlim, focus=focus,
                            **kwargs)
        vis.show()
        return vis

    def close(self):
        lib = capi.load_library() if (self._plan is not None or self._gctx is not None) else None
        if self._plan is not None:
            lib.ramp_plan_destroy(self._plan)
            self._plan = None
        if self._gctx is not None:
            from .plugins import release
            release(self._gctx)
            lib.ramp_ctx_destroy(self._gctx)
            self._gctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
