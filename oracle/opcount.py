"""Algorithmic operation counts per history on the oracle (TEST
INFRASTRUCTURE): an instrument run single-threaded through
liboracle_count.so (oracle/Makefile), which counts Threefry evaluations,
FP32 divisions, square roots, calls of each builtin transcendental and
guide bounce-loop iterations on the per-history path (parameter-only
expressions and unread values excluded; ramp_oracle.c, ORACLE_COUNT).

The default instrument is the LET-like beamline cut after its first
chopper, which is the source segment of the GPU run (MCont, four guides,
the chopper).  bench.py carries the result (ALG_PER_HISTORY) for its
algorithmic roofline; tests/test_opcount.py checks that the two agree.

    python oracle/opcount.py [--instrument file.json] [n_histories] [seed]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
from collections import OrderedDict

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LET = os.path.join(ROOT, "ramp_amd", "instruments", "let_like.json")
LET_SOURCE_SEGMENT = 6   # entries of the LET-like beamline in its source segment
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def counting_lib():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_count.so"], check=True)
    return os.path.join(HERE, "liboracle_count.so")


def prefix(path, k, tmpdir):
    """The instrument file cut to the first k entries of its first block
    (the source counts as one), written to tmpdir; returns (path, names)."""
    import ramp_amd
    text = ramp_amd.Instrument._substitute_params(None, open(path).read())
    desc = json.loads(text, object_pairs_hook=OrderedDict)
    bname, block = next(iter(desc.items()))
    names = [n for n in block if n not in ("linear", "multi")][:k]
    cut = OrderedDict((n, v) for n, v in block.items() if n in ("linear", "multi") or n in names)
    out = os.path.join(tmpdir, f"prefix{k}.json")
    with open(out, "w") as f:
        json.dump({bname: cut}, f)
    return out, names


def count(path, n_hist, seed):
    """{name: count per history} of n_hist histories of the instrument file."""
    os.environ["RAMP_ORACLE_LIB"] = counting_lib()
    import ramp_amd
    from oracle import pyoracle
    assert pyoracle.LIB == os.environ["RAMP_ORACLE_LIB"], "pyoracle was loaded with another build"
    inst = ramp_amd.Instrument(path)
    pyoracle.opcount(reset=True)
    pyoracle.run(inst, n_hist, [seed], threads=1)
    return {k: v / n_hist for k, v in pyoracle.opcount(reset=True).items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instrument", default=None)
    ap.add_argument("n_hist", nargs="?", type=float, default=65536)
    ap.add_argument("seed", nargs="?", type=int, default=7)
    a = ap.parse_args()
    n = int(a.n_hist)
    with tempfile.TemporaryDirectory() as tmp:
        if a.instrument:
            path, names = a.instrument, None
        else:
            path, names = prefix(LET, LET_SOURCE_SEGMENT, tmp)
        per = count(path, n, a.seed)
    print(json.dumps({"histories": n, "seed": a.seed, "components": names, "per_history": per}))


if __name__ == "__main__":
    main()
