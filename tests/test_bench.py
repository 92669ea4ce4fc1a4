"""bench.py's host side on the CPU: the JSON contract's refusal of more than
one GPU, and the instrument prefixes behind `--prefix` (the first k entries
of the first block, the source counted, with $expr$ substitution done)."""
import json
import os
import subprocess
import sys

from tests.parity_util import ROOT


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_instrument_prefix(tmp_path):
    sys.path.insert(0, ROOT)
    import bench
    import ramp_amd
    path = os.path.join(ROOT, "ramp_amd", "instruments", "let_like.json")
    out, names = bench.instrument_prefix(path, {}, 6, str(tmp_path))
    assert names == ["source", "guide01", "guide02", "guide03", "guide04", "chopper_7.83"]
    desc = json.load(open(out))
    assert list(desc) == ["let_like"] and list(desc["let_like"])[1:] == names
    inst = ramp_amd.Instrument(out)
    assert [c.name for c in inst.kernel_refs] == names[1:]
    # the components keep their positions and numbering (source = 1)
    full = ramp_amd.Instrument(path)
    for a, b in zip(inst.kernel_refs, full.kernel_refs):
        assert a.name == b.name and a.idx == b.idx and (a.pos == b.pos).all()
    _, one = bench.instrument_prefix(path, {}, 1, str(tmp_path))
    assert one == ["source"]
