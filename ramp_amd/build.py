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
                  "dev_mesh.h", "dev_adjoint.h", "dev_samples.h", "dev_mcpl.h", "rmath.h", "rvec.h",
                  "rrng.h", "rdf.h", "rconst.h", "rslots.h", "rdefs.h"]
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
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path


def build(verbose=False, extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    write_plugin_headers()
    cmd = [hipcc, *FLAGS, *extra, SRC, "-o", tmp, "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libramp.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
