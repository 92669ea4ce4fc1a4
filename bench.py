"""Benchmark: neutron histories/s of the LET-like spectrometer on one MI355X.

A step is one run of 1e8 histories (one chunk) of the LET-like direct-
geometry chopper spectrometer (ramp_amd/instruments/let_like.json): Maxwellian
source with a 1/E tail, 20-segment tapered m=2 guide, 5 disk choppers,
incoherent sample and a 768-tube x 128-pixel x 1000-TOF detector bank.  The
histories are generated on the device by the source component (synthetic
data); histograms accumulate in HBM and are not read back inside the timed
region.

Prints one JSON line (rank 0, N = 1).  The CPU baseline is the project's
C/OpenMP restatement of the reference kernels (oracle/) timed on the host's
cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neutron histories/s, 1 MI355X (LET-like 1e8); batched-scan & fwd+adjoint n/s"
HBM_PEAK_GBS = 8000.0
# gfx950 vector issue: 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 32-bit lane-ops/s
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# VALU instructions (wave64) per history of the source segment, measured with
# rocprofv3 --pmc SQ_INSTS_VALU (profiles/let_like_source_segment_pmc.txt)
VALU_INSTS_PER_HISTORY = None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--histories", type=float, default=1e8)
    ap.add_argument("--hist-mode", default="atomic", choices=["atomic", "deterministic"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (HBM counter scale)")
    return ap.parse_args()


def cpu_baseline(path, target_s):
    """Oracle (C/OpenMP restatement of the reference kernels) on host cores."""
    import numpy as np
    import ramp_amd
    from oracle import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    inst = ramp_amd.Instrument(path)
    n0 = 1_000_000
    t = time.perf_counter()
    pyoracle.run(inst, n0, [1], threads=threads)
    r0 = n0 / (time.perf_counter() - t)
    n = int(min(max(r0 * target_s, 1e6), 2e9))
    t = time.perf_counter()
    pyoracle.run(inst, n, [2], threads=threads)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{n} histories of the same LET-like instrument, one chunk, {dt:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path shards by "
              "history range across GPUs, which this build does not launch", file=sys.stderr)
        sys.exit(2)
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
    path = os.path.join(ROOT, "ramp_amd", "instruments", "let_like.json")
    N = int(a.histories)
    inst = ramp_amd.Instrument(path, hist_mode=a.hist_mode, chunk=N)
    seed = 20261203
    for w in range(a.warmup):
        inst.execute(N, seeds=[(seed + w) % 4097])
    sync()
    seg_ms, out0, ev, qin, qout = 0.0, 0, 0, 0, 0
    t0 = time.perf_counter()
    for k in range(a.steps):
        inst.execute(N, seeds=[(seed + 100 + k) % 4097])
        st = inst.last_stats
        seg_ms += st["ms_source_segment"]
        out0 += st["state_out_first"]
        ev += st["n_events"]
        qin += st["state_in"]
        qout += st["state_out"]
    sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    K = max(a.steps, 1)
    seg_avg_ms = seg_ms / K
    # algorithmic HBM bytes of the source segment per launch: the 44-byte
    # SoA record (x,y,z,vx,vy,vz,w,t,ctr,key,status) of every survivor it
    # appends to the queue (it reads no state: histories are generated)
    bytes_seg0 = 44.0 * out0 / K
    hbm_gbs = bytes_seg0 / (seg_avg_ms * 1e-3) / 1e9 if seg_avg_ms > 0 else 0.0
    roof = {"bound": "valu", "kernel": "k_segment<false,true> (source segment)",
            "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": bytes_seg0},
            "traffic": None, "launch_ms": seg_avg_ms}
    if VALU_INSTS_PER_HISTORY:
        ach = VALU_INSTS_PER_HISTORY * 64 * N / (seg_avg_ms * 1e-3) / 1e12
        roof.update({"achieved": ach, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                     "frac": ach / VALU_PEAK_TOPS})
    else:
        roof.update({"achieved": None, "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": None})
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
        "config": {"workload": "let_like_1e8", "histories_per_step": N, "chunk": N,
                   "hist_mode": a.hist_mode, "instrument": "ramp_amd/instruments/let_like.json",
                   "detector_bins": 768 * 128 * 1000, "events_per_step": ev / K,
                   "queue_in_per_step": qin / K, "queue_out_per_step": qout / K},
        "roofline": roof,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(path, a.cpu_seconds)
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    inst.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
