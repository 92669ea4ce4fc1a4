"""VALU instructions of each arithmetic primitive on gfx950: the static
count of `v_` instructions in the primitive's probe kernel
(primitive_cost.hip) less the count of the probe's loads and store (the
base kernel of the same arity).  A primitive with branches or a loop
(exp's and tanh's range splits, fmod's long division) counts each
instruction of its code once.  Compiled here with hipcc; nothing is run.

    python tests/native/primitive_cost.py   -> JSON {primitive: VALU instructions}
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "primitive_cost.hip")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S"]
BINARY = {"pow", "atan2", "fmod", "div", "threefry"}


def valu_counts(asm):
    """{kernel name: number of VALU instructions in its body}."""
    out, cur = {}, None
    for line in asm.split("\n"):
        m = re.match(r"^(pc_\w+):", line)
        if m:
            cur = m.group(1)
            out[cur] = 0
            continue
        if cur and line.startswith(".Lfunc_end"):
            cur = None
            continue
        if cur and line.strip().startswith("v_"):
            out[cur] += 1
    return out


def costs():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory() as tmp:
        s = os.path.join(tmp, "pc.s")
        subprocess.run([hipcc, *FLAGS, SRC, "-o", s], check=True, capture_output=True)
        n = valu_counts(open(s).read())
    base1, base2 = n["pc_base1"], n["pc_base2"] - 1   # base2's add is not a load or store
    return {k[3:]: v - (base2 if k[3:] in BINARY else base1)
            for k, v in n.items() if not k.startswith("pc_base")}


if __name__ == "__main__":
    json.dump(costs(), sys.stdout)
    print()
