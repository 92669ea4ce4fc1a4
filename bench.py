"""Benchmark: neutron histories/s of the RAMP ray-tracing path on one MI355X.

The default workload (the headline) is the LET-like direct-geometry chopper
spectrometer (ramp_amd/instruments/let_like.json), one step = one chunk of
1e8 histories: Maxwellian source with a 1/E tail, 20-segment tapered m=2
guide, 5 disk choppers, incoherent sample and a 768-tube x 128-pixel x
1000-TOF detector bank.  `--workload imaging` runs the mesh-imaging
instrument (parallel beam, 50,000-triangle noisy icosphere absorber traced
through the GPU-built BVH, 2048x2048 panel; 1e8 histories per step) and
`--workload minimal` the minimal instrument (flat source, slit, incoherent
cylinder, 256x512 banana detector; 1e6 histories per step).  Histories are
generated on the device by the source component (synthetic data);
histograms accumulate in HBM and are not read back inside the timed region.

Prints one JSON line (rank 0, N = 1).  The CPU baseline is the project's
C/OpenMP restatement of the reference kernels (oracle/) timed on the host's
cores on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neutron histories/s, 1 MI355X (LET-like 1e8); batched-scan & fwd+adjoint n/s"
HBM_PEAK_GBS = 8000.0
# gfx950 vector issue: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3e12 32-bit
# lane-ops/s at one wave64 VALU instruction per CU per clock; the SIMDs can
# issue two VALU instructions per quad-cycle (the SQ_ACTIVE_INST_VALU2
# counter; MI355X's 157.3 TFLOP/s FP32 vector peak = 2 x 39.3e12 FMA lanes),
# so the VALU ceiling is 78.6e12 lane-ops/s.
VALU_SINGLE_ISSUE_TOPS = 256 * 64 * 2.4e9 / 1e12
VALU_PEAK_TOPS = 2 * VALU_SINGLE_ISSUE_TOPS

# Per-history counts of the dominant kernel (the LET-like source segment; the
# imaging workload's separate mesh-walk launch), from rocprofv3 --pmc runs of
# this script at the commit named in each entry
# (profiles/let_like_1e8_segment_kernels.md, profiles/imaging_mesh50k_1e8_bvh.md):
#   valu  = SQ_INSTS_VALU / histories (wave64 instructions, so x64 lane-ops)
#   lanes = SQ_THREAD_CYCLES_VALU / (64 SQ_INSTS_VALU), the active-lane share
#   fetch = FETCH_SIZE / histories, write = WRITE_SIZE / histories (bytes,
#           after the plain-copy calibration of the counters: x2.0, x1.0)
# measured SQ_INSTS_VALU of the timed batch of the 256-setting scan, and per
# history of the powder forward and adjoint kernels
TAS_VALU_PER_SCAN = {"valu": 1.38435e10, "commit": "d9e6eea"}
POWDER_VALU_PER_HISTORY = {"forward": 63.43, "adjoint": 66.93, "commit": "d9e6eea"}


# Algorithmic work of the LET-like source segment, per history: the
# primitives its histories need, counted on the oracle's op-count build
# (`python oracle/opcount.py 65536 7`: the instrument cut after its first
# chopper, single-threaded; parameter-only expressions and unread values
# excluded), and each primitive's own VALU instruction count on gfx950
# (`python tests/native/primitive_cost.py`: static count in a one-primitive
# kernel).  Their product is a lower bound on the lane-ops a history needs:
# the plain adds, multiplies and compares between the primitives are not
# counted.  tests/test_opcount.py re-derives both tables.
ALG_PER_HISTORY = {"threefry": 5.0, "div": 41.76414489746094, "sqrt": 2.73858642578125,
                   "exp": 0.73858642578125, "log": 0.0, "pow": 0.0, "tanh": 1.6868743896484375,
                   "atan2": 1.995635986328125, "sin": 0.0, "cos": 0.0, "asin": 0.0, "acos": 0.0,
                   "fmod": 0.9978179931640625, "guide_bounce": 2.9347686767578125,
                   "guide_exit": 3.99566650390625}
PRIMITIVE_VALU = {"threefry": 117, "div": 11, "sqrt": 17, "exp": 38, "log": 49, "pow": 121, "tanh": 64,
                  "atan2": 127, "sin": 30, "cos": 31, "asin": 44, "acos": 46, "fmod": 63}


def algorithmic_lane_ops():
    """Lane-ops per history of the LET-like source segment's primitives."""
    return sum(ALG_PER_HISTORY[k] * PRIMITIVE_VALU[k] for k in PRIMITIVE_VALU)


def valu_roofline(kernel, lane_ops, launch_ms, basis, extra=None):
    """VALU roofline of a kernel: lane-op slots issued (measured instruction
    count x 64) per second of its launch, against the dual-issue peak."""
    ach = lane_ops / (launch_ms * 1e-3) / 1e12 if launch_ms > 0 else 0.0
    r = {"bound": "valu", "kernel": kernel, "achieved": ach, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
         "frac": ach / VALU_PEAK_TOPS, "single_issue_peak": VALU_SINGLE_ISSUE_TOPS, "traffic": None,
         "launch_ms": launch_ms, "basis": basis}
    if extra:
        r.update(extra)
    return r


PMC_PER_HISTORY = {
    "let_like": {"kernel": "k_segment<false,true>", "valu": 75.41, "lanes": 0.666, "fetch": 7.585,
                 "write": 19.444, "commit": "0fe9900"},
    "imaging": {"kernel": "k_mesh_walk", "valu": 110.90, "lanes": 0.568, "fetch": 32.24, "write": 25.50,
                "commit": "45ade1f"},
    "minimal": None,
}

WORKLOADS = {
    "let_like": dict(file="let_like.json", histories=1e8, name="let_like_1e8",
                     bins=768 * 128 * 1000),
    "imaging": dict(file="imaging.json", histories=1e8, name="imaging_mesh50k_1e8",
                    bins=2048 * 2048),
    "minimal": dict(file="minimal.json", histories=1e6, name="minimal_1e6", bins=256 * 512),
    "tas_scan": dict(file="tas_scan.json", histories=None, name="tas_scan_256", bins=256 * 128),
    "powder_adjoint": dict(file="powder_diff.json", histories=3e7, name="powder_fwd_adjoint_3e7",
                           bins=2048 * 256),
    "mcpl_out": dict(file="mcpl_out.json", histories=2e7, name="mcpl_out_guide_dump_2e7", bins=1000),
    "mcpl_in": dict(file="mcpl_in.json", histories=2e7, name="mcpl_in_2e7_particles", bins=180 * 560),
    "rescal": dict(file="rescal.json", histories=5e7, name="rescal_events_5e7", bins=1),
    "plugin": dict(file="plugin.json", histories=5e7, name="plugin_attenuator_5e7", bins=500),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="let_like", choices=sorted(WORKLOADS))
    ap.add_argument("--histories", type=float, default=None)
    ap.add_argument("--hist-mode", default="atomic", choices=["atomic", "deterministic"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prefix", type=int, default=0,
                    help="keep only the first K entries (source included) of the instrument's "
                         "first block: the cost of an instrument prefix, e.g. LET-like cut after "
                         "its k-th component (the CPU baseline is skipped)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (HBM counter scale)")
    return ap.parse_args()


def instrument_prefix(path, kw, k, tmpdir):
    """The instrument file cut to the first k entries of its first block
    (the source counts as one), written to tmpdir; returns (path, names)."""
    import ramp_amd
    from collections import OrderedDict
    text = ramp_amd.Instrument._substitute_params(None, open(path).read(), **kw)
    desc = json.loads(text, object_pairs_hook=OrderedDict)
    bname, block = next(iter(desc.items()))
    names = [n for n in block if n not in ("linear", "multi")][:k]
    cut = OrderedDict((n, v) for n, v in block.items() if n in ("linear", "multi") or n in names)
    out = os.path.join(tmpdir, f"prefix{k}.json")
    with open(out, "w") as f:
        json.dump({bname: cut}, f)
    return out, names


def instrument_prefix_text(path, k, tmpdir):
    """A batch instrument file (with per-setting $...$ expressions, so not
    parsed here) cut as text to its first k component entries: the entries
    are the keys indented by eight spaces whose value opens a brace."""
    import re
    lines = open(path).read().splitlines()
    starts = [i for i, ln in enumerate(lines) if re.match(r'^ {8}"[^"]+" : \{', ln)]
    names = [re.match(r'^ {8}"([^"]+)"', lines[i]).group(1) for i in starts][:k]
    if k >= len(starts):
        return path, names
    keep = lines[:starts[k]]
    while keep and not keep[-1].strip():
        keep.pop()
    keep[-1] = keep[-1].rstrip().rstrip(",")
    out = os.path.join(tmpdir, f"prefix{k}.json")
    with open(out, "w") as f:
        f.write("\n".join(keep + ["    }", "}"]) + "\n")
    return out, names


def instrument_args(workload, tmpdir):
    """Keyword arguments of the workload's instrument file; the imaging
    object (50,000-triangle icosphere, r = 15 mm, sigma = 1 mm, seed 3) is
    written as binary STL at run time."""
    if workload != "imaging":
        return {}
    from ramp_amd import meshes
    from ramp_amd.components import write_stl_binary
    path = os.path.join(tmpdir, "icosphere50k.stl")
    write_stl_binary(path, meshes.icosphere(50, radius=0.015, sigma=0.001, seed=3))
    return {"mesh": repr(path)}


def cpu_baseline(path, kw, target_s, label):
    """Oracle (C/OpenMP restatement of the reference kernels) on host cores."""
    import ramp_amd
    from oracle import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    inst = ramp_amd.Instrument(path, **kw)
    n0 = 20_000
    while True:
        t = time.perf_counter()
        pyoracle.run(inst, n0, [1], threads=threads)
        dt0 = time.perf_counter() - t
        if dt0 > 0.5 or n0 >= 4_000_000:
            break
        n0 *= 4
    r0 = n0 / dt0
    n = int(min(max(r0 * target_s, 1e4), 2e9))
    t = time.perf_counter()
    pyoracle.run(inst, n, [2], threads=threads)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{n} histories of the same {label} instrument, one chunk, {dt:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path shards by "
              "history range across GPUs, which this build does not launch", file=sys.stderr)
        sys.exit(2)
    import tempfile
    import ramp_amd
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    if a.calibrate_copy:
        import torch
        src = torch.ones(a.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    wl = WORKLOADS[a.workload]
    if a.workload == "tas_scan":
        return bench_tas_scan(a, wl, sync)
    if a.workload == "powder_adjoint":
        return bench_powder_adjoint(a, wl, sync)
    if a.workload in ("mcpl_out", "mcpl_in", "rescal", "plugin"):
        return bench_mcpl(a, wl, sync)
    tmpdir = tempfile.mkdtemp(prefix="ramp_bench_")
    kw = instrument_args(a.workload, tmpdir)
    path = os.path.join(ROOT, "ramp_amd", "instruments", wl["file"])
    prefix_names = None
    if a.prefix:
        path, prefix_names = instrument_prefix(path, kw, a.prefix, tmpdir)
        a.no_cpu = True
    N = int(a.histories if a.histories else wl["histories"])
    inst = ramp_amd.Instrument(path, hist_mode=a.hist_mode, chunk=N, **kw)
    seed = 20261203
    for w in range(a.warmup):
        inst.execute(N, seeds=[(seed + w) % 4097])
    sync()
    seg_ms, walk_ms, sort_ms, rays, out0, ev, qin, qout = 0.0, 0.0, 0.0, 0, 0, 0, 0, 0
    t0 = time.perf_counter()
    for k in range(a.steps):
        inst.execute(N, seeds=[(seed + 100 + k) % 4097])
        st = inst.last_stats
        seg_ms += st["ms_source_segment"]
        walk_ms += st["ms_mesh_walk"]
        sort_ms += st["ms_walk_sort"]
        rays += st["walk_rays"]
        out0 += st["state_out_first"]
        ev += st["n_events"]
        qin += st["state_in"]
        qout += st["state_out"]
    sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    K = max(a.steps, 1)
    seg_avg_ms = seg_ms / K
    walk_avg_ms = walk_ms / K
    # algorithmic HBM bytes of the source segment per launch: the 44-byte
    # SoA record (x,y,z,vx,vy,vz,w,t,ctr,key,status) of every survivor it
    # appends to the queue (it reads no state: histories are generated)
    bytes_seg0 = 44.0 * out0 / K
    hbm_gbs = bytes_seg0 / (seg_avg_ms * 1e-3) / 1e9 if seg_avg_ms > 0 else 0.0
    roof = {"bound": "valu", "kernel": "k_segment<false,true> (source segment)",
            "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": bytes_seg0},
            "traffic": None, "launch_ms": seg_avg_ms,
            "achieved": None, "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": None}
    pmc = None if a.prefix else PMC_PER_HISTORY.get(a.workload)
    if pmc and pmc.get("kernel") == "k_mesh_walk":
        # the dominant kernel is the separate mesh-walk launch; its
        # algorithmic HBM bytes: the 32-byte ray record read and the 16-byte
        # hit record written per walked ray (the BVH and triangles stay in
        # cache)
        walk_bytes = 48.0 * rays / K
        walk_gbs = walk_bytes / (walk_avg_ms * 1e-3) / 1e9 if walk_avg_ms > 0 else 0.0
        roof.update({"kernel": "k_mesh_walk", "launch_ms": walk_avg_ms,
                     "hbm": {"achieved": walk_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": walk_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": walk_bytes},
                     "walk_rays_per_step": rays / K, "ms_walk_sort": sort_ms / K,
                     "source_segment_ms": seg_avg_ms})
        launch_ms = walk_avg_ms
    else:
        launch_ms = seg_avg_ms
    if pmc and launch_ms > 0:
        # lane-op slots issued per second by the launch (measured instruction
        # count per history x 64 x histories / measured launch time)
        ach = pmc["valu"] * 64 * N / (launch_ms * 1e-3) / 1e12
        roof.update({"achieved": ach, "frac": ach / VALU_PEAK_TOPS,
                     "single_issue_peak": VALU_SINGLE_ISSUE_TOPS,
                     "valu_insts_per_history": pmc["valu"], "active_lane_share": pmc["lanes"],
                     "traffic": (pmc["fetch"] + pmc["write"]) * N, "pmc_commit": pmc["commit"]})
        if a.workload == "let_like":
            # the same launch priced by the algorithm's primitives instead of
            # the instructions issued (divergence, control flow, spills and
            # the arithmetic between primitives excluded)
            alg = algorithmic_lane_ops()
            alg_ach = alg * N / (launch_ms * 1e-3) / 1e12
            roof["algorithmic"] = {
                "basis": "primitives per history counted on the oracle (oracle/opcount.py) x their "
                         "gfx950 VALU instruction counts (tests/native/primitive_cost.py)",
                "lane_ops_per_history": alg, "achieved": alg_ach, "frac": alg_ach / VALU_PEAK_TOPS,
                "frac_single_issue": alg_ach / VALU_SINGLE_ISSUE_TOPS,
                "share_of_issued_slots": alg / (pmc["valu"] * 64),
                "primitives_per_history": {k: v for k, v in ALG_PER_HISTORY.items() if v}}
    out = {
        "metric": METRIC,
        "value": N * K / dt,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: histories generated on the device by the source component",
        "config": {"workload": wl["name"], "histories_per_step": N, "chunk": N,
                   "hist_mode": a.hist_mode, "instrument": "ramp_amd/instruments/" + wl["file"],
                   "detector_bins": wl["bins"], "events_per_step": ev / K,
                   "queue_in_per_step": qin / K, "queue_out_per_step": qout / K,
                   **({"prefix_components": prefix_names} if prefix_names else {})},
        "roofline": roof,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(path, kw, a.cpu_seconds, wl["name"])
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    inst.close()
    print(json.dumps(out), flush=True)


def cpu_baseline_settings(path, settings, counts, target_s):
    """Oracle on host cores over a bounded sample of the scan: the first
    settings in order, each at its own count scaled so the sample takes
    about target_s; the rate is histories/s over the sample."""
    import ramp_amd
    from oracle import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n0 = 200_000
    t = time.perf_counter()
    pyoracle.run(ramp_amd.Instrument(path, **settings[0]), n0, [1], threads=threads)
    rate0 = n0 / (time.perf_counter() - t)
    budget = rate0 * target_s
    done, used, k = 0, 0.0, 0
    t0 = time.perf_counter()
    while k < len(settings) and done < budget:
        n = int(min(int(counts[k]), max(budget - done, 1e4)))
        pyoracle.run(ramp_amd.Instrument(path, **settings[k]), n, [k + 1], threads=threads)
        done += n
        k += 1
    used = time.perf_counter() - t0
    return {"value": done / used, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{done} histories over the first {k} scan settings, {used:.1f} s"}


def bench_tas_scan(a, wl, sync):
    """256-point triple-axis scan: one batched launch per step (graph
    replay) after the autotuner has picked the launch shape, against the
    same scan run the reference's way (one plain run per setting)."""
    import numpy as np
    import ramp_amd
    sys.path.insert(0, os.path.join(ROOT, "ramp_amd", "instruments"))
    import tas_scan
    settings, counts = tas_scan.settings()
    counts = np.array(counts, dtype=np.uint64)
    if a.histories:
        counts = np.maximum(counts * (a.histories / counts.sum()), 1).astype(np.uint64)
    total = int(counts.sum())
    path = os.path.join(ROOT, "ramp_amd", "instruments", wl["file"])
    prefix_names = None
    if a.prefix:
        import tempfile
        path, prefix_names = instrument_prefix_text(path, a.prefix, tempfile.mkdtemp(prefix="ramp_bench_"))
        a.no_cpu = True
    b = ramp_amd.BatchInstrument(path, settings, hist_mode=a.hist_mode)
    t = time.perf_counter()
    best, tuning = b.tune(counts)
    tune_s = time.perf_counter() - t
    seeds = np.random.default_rng(20261203).integers(0, 4097, size=len(settings))

    def rel_err():
        r = []
        for inst in b.instruments:
            w, w2 = inst.hist64("tube")
            r.append(math.sqrt(w2.sum()) / w.sum() if w.sum() > 0 else float("nan"))
        return np.array(r)

    # one scan each way with the same seeds (also warms both paths): the
    # histories are the same, so the statistical errors must match (an
    # instrument prefix, timed for the cost of its components, has no tube)
    ok, err_dev = np.zeros(len(settings), dtype=bool), None
    if not a.prefix:
        b.reset()
        b.execute(counts, seeds=seeds)
        err_batch = rel_err()
        b.reset()
        b.execute_sequential(counts, seeds=seeds)
        err_seq = rel_err()
        b.reset()
        ok = np.isfinite(err_seq)
        err_dev = float(np.max(np.abs(err_batch[ok] / err_seq[ok] - 1.0))) if ok.any() else None
    for _ in range(a.warmup):
        b.execute(counts, seeds=seeds)
    sync()
    dev_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dev_ms += b.execute(counts, seeds=seeds)["ms_total"]
    sync()
    dt = time.perf_counter() - t0
    K = max(a.steps, 1)
    sync()
    if a.prefix:
        print(json.dumps({"metric": METRIC, "value": total * K / dt, "unit": "histories/s", "n_gpus": 1,
                          "steps": K, "warmup": a.warmup, "ms_per_step": dt / K * 1e3,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic", "config": {"workload": wl["name"], "prefix_components": prefix_names,
                                                          "histories_per_step": total,
                                                          "device_ms_per_step": dev_ms / K}}), flush=True)
        b.close()
        return
    t0 = time.perf_counter()
    b.execute_sequential(counts, seeds=seeds)
    sync()
    dt_seq = time.perf_counter() - t0
    seq_ms_dev = sum(i.last_stats["ms_total"] for i, c in zip(b.instruments, counts) if int(c))
    out = {
        "metric": METRIC,
        "value": total * K / dt,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 256 scan points from numpy default_rng(4), histories generated on the device",
        "config": {"workload": wl["name"], "settings": len(settings), "histories_per_step": total,
                   "counts": "log-uniform 1e5..4e6 per setting", "hist_mode": a.hist_mode,
                   "instrument": "ramp_amd/instruments/tas_scan.json",
                   "autotuned": {"npt": best[0], "wg": best[1], "tune_s": tune_s,
                                 "ms_by_shape": {f"{k[0]}x{k[1]}": v for k, v in tuning.items()}},
                   "device_ms_per_step": dev_ms / K},
        "sequential": {"histories_per_s": total / dt_seq, "ms": dt_seq * 1e3,
                       "device_ms_sum": seq_ms_dev,
                       "batch_speedup": (total * K / dt) / (total / dt_seq),
                       "batch_speedup_device": seq_ms_dev / (dev_ms / K) if dev_ms > 0 else None,
                       "settings_with_counts": int(ok.sum()),
                       "max_rel_error_mismatch": err_dev},
        "roofline": valu_roofline("k_batch<MONO> (the whole scan in one launch)",
                                  TAS_VALU_PER_SCAN["valu"] * 64, dev_ms / K,
                                  "wave64 VALU instructions of one scan, measured (SQ_INSTS_VALU) at commit "
                                  + TAS_VALU_PER_SCAN["commit"]),
    }
    b.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_settings(path, settings, counts, a.cpu_seconds)
    print(json.dumps(out), flush=True)


def cpu_baseline_adjoint(di, t, t2, eps, target_s):
    """Oracle forward + FP64 oracle adjoint of the same loss on host cores,
    on a bounded number of histories; histories/s of the pair."""
    import numpy as np
    from oracle import pyoracle
    from ramp_amd.diff import weighted_mse
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    inst = di.build(di.theta0)

    def pair(n):
        h = pyoracle.run(inst, n, [1], threads=threads)["hist"][di.det_slot][0]
        _, dl = weighted_mse(h, t, t2, eps)
        pyoracle.adjoint(inst, n, [1], di.det_slot, dl, di.n_params, threads=threads)

    n0 = 200_000
    t0 = time.perf_counter()
    pair(n0)
    r0 = n0 / (time.perf_counter() - t0)
    n = int(min(max(r0 * target_s, 1e5), 3e7))
    t0 = time.perf_counter()
    pair(n)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{n} histories, forward + FP64 adjoint of the same instrument, {dt:.1f} s"}


def bench_powder_adjoint(a, wl, sync):
    """Forward + adjoint of the differentiable powder instrument: one step is
    a forward run (3e7 histories), the weighted-MSE loss against a target
    histogram (the same simulator at theta0 * 1.02, other seed) and one
    adjoint pass giving dL/dtheta for the 44 parameters.  Reported beside
    a forward alone and the cost of the same gradient by central finite
    differences (two forwards per parameter, same seeds)."""
    import numpy as np
    from ramp_amd.diff import DiffInstrument, weighted_mse
    sys.path.insert(0, os.path.join(ROOT, "ramp_amd", "instruments"))
    import powder_diff
    path = os.path.join(ROOT, "ramp_amd", "instruments", wl["file"])
    N = int(a.histories if a.histories else wl["histories"])
    d, F2 = powder_diff.lines()
    di = DiffInstrument(path, n_hist=N, seeds=(20261203 % 4097,),
                        **powder_diff.variables(powder_diff.theta0(), F2))
    tgt = DiffInstrument(path, n_hist=N, seeds=(20261204 % 4097,),
                         **powder_diff.variables(powder_diff.theta0() * 1.02, F2))
    t, t2 = tgt.histogram(tgt.theta0)
    tgt.close()
    eps = 1e-3 * float(np.mean(t2[t2 > 0]))
    th = di.theta0
    for _ in range(a.warmup):
        h, _ = di.histogram(th)
        di.adjoint(th, weighted_mse(h, t, t2, eps)[1])
    sync()
    fwd_ms = fwd_seg_ms = adj_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        h, _ = di.histogram(th)
        L, dldh = weighted_mse(h, t, t2, eps)
        g = di.adjoint(th, dldh)
        fwd_ms += di.last_forward_ms
        fwd_seg_ms += di.last_forward_seg_ms
        adj_ms += di.last_adjoint_ms
    sync()
    dt = time.perf_counter() - t0
    K = max(a.steps, 1)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        di.histogram(th)
    sync()
    dt_fwd = time.perf_counter() - t0
    out = {
        "metric": METRIC,
        "value": N * K / dt,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 40 Bragg lines from numpy default_rng(5), histories generated on the device",
        "config": {"workload": wl["name"], "histories_per_step": N, "parameters": di.n_params,
                   "instrument": "ramp_amd/instruments/powder_diff.json", "detector_bins": wl["bins"],
                   "loss": float(L), "grad_norm": float(np.linalg.norm(g)),
                   "device_ms": {"forward": fwd_ms / K, "forward_source_segment": fwd_seg_ms / K,
                                 "adjoint": adj_ms / K}},
        "forward_only": {"histories_per_s": N * K / dt_fwd, "ms_per_step": dt_fwd / K * 1e3,
                         "fwd_adj_over_fwd": dt / dt_fwd},
        "finite_differences": {"forwards": 2 * di.n_params,
                               "ms": 2 * di.n_params * dt_fwd / K * 1e3,
                               "adjoint_speedup": (2 * di.n_params * dt_fwd / K) / (dt / K)},
        "roofline": valu_roofline("k_segment<false,true,2> (forward source segment)",
                                  POWDER_VALU_PER_HISTORY["forward"] * 64 * N, fwd_seg_ms / K,
                                  "wave64 VALU instructions per history, measured (SQ_INSTS_VALU) at "
                                  "commit " + POWDER_VALU_PER_HISTORY["commit"],
                                  extra={"adjoint": {"kernel": "k_adjoint<2>", "launch_ms": adj_ms / K,
                                                     "achieved": POWDER_VALU_PER_HISTORY["adjoint"] * 64 * N
                                                     / (adj_ms / K * 1e-3) / 1e12}}),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_adjoint(di, t, t2, eps, a.cpu_seconds)
    di.close()
    print(json.dumps(out), flush=True)



def bench_mcpl(a, wl, sync):
    """MCPL particle I/O.  mcpl_out: the layout of the reference's
    examples/MCPL_test.json (flat source, two straight guides, SMCPLOut at
    the guide exit, energy monitor), one chunk per step; the records of the
    ~97 % of neutrons alive at the dump are packed in the segment kernel and
    compacted on the device (ramp_mcpl_read).  mcpl_in: SMCPLIn replaces
    every history with a particle of a synthetic N-particle file (MCPL's
    own packing, written before the timed region), then a banana detector.
    value = histories / device time of the step (run + compaction); the
    wall rate with the records' copy to the host is reported beside it."""
    import tempfile
    import numpy as np
    import ramp_amd
    from ramp_amd.mcpl import write_mcpl_full
    N = int(a.histories if a.histories else wl["histories"])
    path = os.path.join(ROOT, "ramp_amd", "instruments", wl["file"])
    tmpdir = tempfile.mkdtemp(prefix="ramp_bench_")
    if a.workload == "mcpl_out":
        kw = {"mcpl_file": repr(""), "standard": 0}
    elif a.workload == "rescal":
        kw = {"tx": 0.3, "tz": 1.0, "arm_rot": math.atan2(0.3, 1.0)}
    elif a.workload == "plugin":
        # the plugin module sits beside its instrument file, as a user's would
        sys.path.insert(0, os.path.join(ROOT, "ramp_amd", "instruments"))
        kw = {"where": 1, "f3": 0.5, "q3": 0.5}
    else:
        rng = np.random.default_rng(7)
        rec = np.zeros((N, 11), np.float32)
        u = rng.normal(size=(N, 3)) * [0.02, 0.02, 1.0]
        u /= np.linalg.norm(u, axis=1)[:, None]
        rec[:, 3:5] = rng.uniform(-2, 2, size=(N, 2))
        rec[:, 6:8] = u[:, :2]
        rec[:, 8] = rng.uniform(1.5e-9, 9e-9, size=N) * np.sign(u[:, 2])
        rec[:, 9] = rng.uniform(20, 40, size=N)
        rec[:, 10] = 1.0 / N
        fn = os.path.join(tmpdir, "particles.mcpl")
        write_mcpl_full(fn, rec)
        del rec, u
        kw = {"mcpl_file": repr(fn)}
    inst = ramp_amd.Instrument(path, hist_mode=a.hist_mode, chunk=N, **kw)
    seed = 20261203
    for w in range(a.warmup):
        inst.execute(N, seeds=[(seed + w) % 4097])
    sync()
    dev_ms, run_ms, sel_ms, wr_ms, recs, app_ms = 0.0, 0.0, 0.0, 0.0, 0, 0.0
    ev0 = 0
    t0 = time.perf_counter()
    for k in range(a.steps):
        inst.execute(N, seeds=[(seed + 100 + k) % 4097])
        st = inst.stats()
        run_ms += st["ms_total"]
        if a.workload == "mcpl_out":
            sel_ms += st["ms_mcpl_compact"]
            wr_ms += st["ms_mcpl_write"]
            recs += inst._mcpl_out[0].scat.records.shape[0]
        if a.workload == "rescal":
            app_ms += st["ms_rescal_append"]
        if a.workload == "plugin":
            app_ms += st["ms_plugin"]
            ev0 += st["plugin_items"]
    sync()
    dt = time.perf_counter() - t0
    K = max(a.steps, 1)
    dev_ms = (run_ms + sel_ms) / K
    out = {
        "metric": METRIC, "value": N / (dev_ms * 1e-3), "unit": "histories/s", "n_gpus": 1,
        "steps": K, "warmup": a.warmup, "ms_per_step": dev_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic MCPL file of N particles, resident in HBM" if a.workload == "mcpl_in"
                else "synthetic: histories generated on the device",
        "config": {"workload": wl["name"], "histories_per_step": N, "chunk": N,
                   "instrument": "ramp_amd/instruments/" + wl["file"],
                   "run_ms": run_ms / K, "wall_ms_per_step_with_host_copy": dt / K * 1e3,
                   "wall_histories_per_s": N * K / dt},
    }
    if a.workload == "mcpl_out":
        R = recs / K
        w_ms = wr_ms / K
        # k_sel_write: reads every tag (4 B) and each record (44 B), writes each record
        alg = 4.0 * N + 88.0 * R
        ach = alg / (w_ms * 1e-3) / 1e9 if w_ms > 0 else 0.0
        out["config"].update({"records_per_step": R, "compaction_ms": sel_ms / K,
                              "record_bytes_per_step": 44 * R})
        out["roofline"] = {"bound": "hbm", "kernel": "k_sel_write (record compaction)",
                           "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None, "launch_ms": w_ms,
                           "algorithmic_bytes_per_launch": alg}
    elif a.workload == "rescal":
        import ctypes as C
        cnt = C.c_uint64(0)
        lib = ramp_amd._capi.load_library()
        lib.ramp_rescal_read(inst._plan, 0, None, 0, C.byref(cnt))
        E = cnt.value / (a.warmup + K)  # events per step (the list accumulates)
        a_ms = app_ms / K
        # k_ev_append and k_sel_count: 4 B tag per history read twice, 20 B
        # event read and 20 B written per kept event
        alg = 8.0 * N + 40.0 * E
        ach = alg / (a_ms * 1e-3) / 1e9 if a_ms > 0 else 0.0
        out["config"].update({"events_per_step": E, "append_ms": a_ms})
        out["roofline"] = {"bound": "hbm", "kernel": "SRescal event append (k_sel_count, k_sel_scan, "
                           "k_ev_append, k_ev_advance)", "achieved": ach, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None, "launch_ms": a_ms,
                           "algorithmic_bytes_per_launch": alg}
    elif a.workload == "plugin":
        items = ev0 / K
        p_ms = app_ms / K
        # the plugin step reads and rewrites each queued neutron's full-state
        # record: 23 floats, key and status = 100 B each way
        alg = 200.0 * items
        ach = alg / (p_ms * 1e-3) / 1e9 if p_ms > 0 else 0.0
        out["config"].update({"plugin_items_per_step": items, "plugin_ms": p_ms,
                              "plugin": "ramp_amd/instruments/SAttenuatorPlugin.py (hipRTC)"})
        out["roofline"] = {"bound": "hbm", "kernel": "ramp_plugin_kernel (SAttenuatorPlugin, in place on "
                           "the full-state queue)", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None, "launch_ms": p_ms,
                           "algorithmic_bytes_per_launch": alg}
    else:
        # the source segment reads each particle's 44-byte record
        seg_ms = inst.stats()["ms_source_segment"]
        alg = 44.0 * min(N, inst._comp_list[0].scat.n_particles)
        ach = alg / (seg_ms * 1e-3) / 1e9 if seg_ms > 0 else 0.0
        out["roofline"] = {"bound": "hbm", "kernel": "k_segment<true,true> (SMCPLIn unpack + detector)",
                           "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None, "launch_ms": seg_ms,
                           "algorithmic_bytes_per_launch": alg}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_mcpl(path, kw, a.cpu_seconds, wl["name"], a.workload)
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    inst.close()
    print(json.dumps(out), flush=True)


def cpu_baseline_mcpl(path, kw, target_s, label, workload):
    """Oracle on host cores: the same instrument on a bounded sample; for
    mcpl_out the oracle's records then go through the same numpy
    compaction the reference does on the host (alive filter in order)."""
    import numpy as np
    import ramp_amd
    from oracle import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    inst = ramp_amd.Instrument(path, **kw)
    if workload == "plugin":
        # the plugin's host build (the same source, g++) runs in the oracle
        import tempfile
        inst.plan_desc()
        for k, c in enumerate(inst._comp_list):
            if c.scat.type_code == ramp_amd.capi.RAMP_S_PLUGIN:
                _lib, fn = pyoracle.build_host_plugin(c.scat.plugin_source(), c.scat.entry,
                                                      tempfile.mkdtemp(prefix="ramp_plugin_"))
                pyoracle.set_plugin(k, fn)
    n = 200_000
    t = time.perf_counter()
    pyoracle.run(inst, n, [1], threads=threads)
    r0 = n / (time.perf_counter() - t)
    n = int(min(max(r0 * target_s, 1e4), 5e7))
    if workload == "mcpl_in":
        n = min(n, inst._comp_list[0].scat.n_particles)
    t = time.perf_counter()
    r = pyoracle.run(inst, n, [2], threads=threads)
    if workload == "mcpl_out":
        np.ascontiguousarray(r["mcpl"][0])
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{n} histories of the same {label} instrument, one chunk, {dt:.1f} s"}


if __name__ == "__main__":
    main()
