"""Build libramp.so (all HIP code, gfx950) in-tree with hipcc.

Flags: -ffp-contract=off on host and device (no implicit fused
multiply-adds, so host-precomputed constants, kernels and the CPU oracle
agree bit for bit); packed FP32 instructions disabled for device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "ramp_gpu.hip")
OUT = os.path.join(HERE, "libramp.so")
ARCH = os.environ.get("RAMP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# headers a run-time compiled plugin kernel includes (hipRTC has no file
# system view of the package; they are embedded in the library)
PLUGIN_HEADERS = ["plugin_kernel.h", "ramp_plugin.h", "dev_queue.h", "dev_components.h", "dev_plan.h",
                  "dev_mesh.h", "bvh_types.h", "dev_adjoint.h", "dev_samples.h", "dev_mcpl.h", "rmath.h", "rvec.h",
                  "rrng.h", "rconst.h", "rslots.h", "rdefs.h"]
# hipRTC provides the fixed-width integer types in its own namespace and
# no <stddef.h>/<stdint.h>; these stand in for the two system headers
RTC_STDDEF = "#pragma once\n"
RTC_STDINT = ("#pragma once\n" + "".join(f"using __hip_internal::{t};"
              for t in ("uint8_t", "uint16_t", "uint32_t", "uint64_t", "int8_t", "int16_t", "int32_t",
                        "int64_t")) + "\n")


def write_plugin_headers():
    """csrc/plugin_headers.inc: the embedded header table (generated)."""
    csrc = os.path.join(HERE, "csrc")
    entries = [(h, open(os.path.join(csrc, h)).read()) for h in PLUGIN_HEADERS]
    entries.append(("../../include/ramp.h", open(os.path.join(HERE, "..", "include", "ramp.h")).read()))
    entries += [("stddef.h", RTC_STDDEF.replace("\\n", "\n")), ("stdint.h", RTC_STDINT.replace("\\n", "\n"))]
    out = ["/* generated by ramp_amd/build.py: headers for hipRTC plugin programs */"]
    out.append(f"static const int kPluginHeaderCount = {len(entries)};")
    out.append("static const char *const kPluginHeaderNames[] = {" +
               ", ".join('"' + n + '"' for n, _ in entries) + "};")
    out.append("static const char *const kPluginHeaderTexts[] = {")
    for n, t in entries:
        assert ")rampdelim\"" not in t
        out.append('R"rampdelim(' + t + ')rampdelim",')
    out.append("};")
    path = os.path.join(csrc, "plugin_headers.inc")
    text = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != text:  # keep its time stamp when unchanged
        with open(path, "w") as f:
            f.write(text)
    return path


def _units():
    """(source, object, defines) of every translation unit: the eighteen
    feature-set units of the large kernels and the host/C-ABI unit."""
    csrc = os.path.join(HERE, "csrc")
    obj = os.path.join(HERE, "_build")
    units = []
    # roughly the largest units first
    for v, part in ((5, 0), (5, 2), (2, 2), (2, 0), (4, 0), (5, 1), (1, 0), (4, 1), (4, 2), (1, 2), (1, 1),
                    (3, 0), (3, 2), (2, 1), (3, 1), (0, 1), (0, 0), (0, 2)):
        units.append((os.path.join(csrc, "kern_v.hip"), os.path.join(obj, f"kern_v{v}_{part}.o"),
                      [f"-DRAMP_KV={v}", f"-DRAMP_KPART={part}"]))
    units.append((SRC, os.path.join(obj, "ramp_gpu.o"), []))
    return obj, units


def _includes(src, seen=None):
    """src and every file it includes with #include "...", recursively."""
    import re
    seen = [] if seen is None else seen
    src = os.path.normpath(src)
    if src in seen or not os.path.exists(src):
        return seen
    seen.append(src)
    for m in re.finditer(r'^\s*#\s*include\s+"([^"]+)"', open(src).read(), re.M):
        _includes(os.path.join(os.path.dirname(src), m.group(1)), seen)
    return seen


def _stale(obj_path, deps):
    return not os.path.exists(obj_path) or os.path.getmtime(obj_path) < max(os.path.getmtime(d) for d in deps)


def build(verbose=False, extra=(), jobs=None, incremental=False):
    """Compile the units in parallel (at most `jobs`, default the CPU count
    capped at 8) and link libramp.so.  With incremental=True a unit whose
    object is newer than every source and header is not recompiled."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    write_plugin_headers()
    obj, units = _units()
    os.makedirs(obj, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    jobs = jobs or min(8, os.cpu_count() or 1)
    pending = list(units)
    if incremental and not extra:
        pending = [u for u in units if _stale(u[1], _includes(u[0]) + [os.path.abspath(__file__)])]
    running = []
    failed = []
    while pending or running:
        while pending and len(running) < jobs:
            src, o, defs = pending.pop(0)
            cmd = [hipcc, *cflags, *extra, *defs, "-c", src, "-o", o]
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((o, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        o, pr = running.pop(0)
        out, _ = pr.communicate()
        if pr.returncode != 0:
            failed.append((o, out))
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        for o, out in failed:
            sys.stderr.write(f"---- {o}\n{out}")
        raise RuntimeError("hipcc failed building libramp.so")
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[u[1] for u in units], "-o", tmp,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libramp.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
