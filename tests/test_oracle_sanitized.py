"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile: liboracle_san.so; host code only), run in a child process
over the parity instruments, the chunked path, the adjoint and the pins'
small instruments.  Any out-of-bounds access, use after free or undefined
behaviour aborts the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, ROOT)
import numpy as np
import ramp_amd
from oracle import pyoracle
from tests.parity_util import instrument_cases, mcpl_cases
cases = instrument_cases()
cases.update(mcpl_cases())
for name, (path, kw) in sorted(cases.items()):
    inst = ramp_amd.Instrument(path, **kw)
    n = 3000 if name not in ("imaging", "mesh_multi") else 600
    r = pyoracle.run(inst, n, [7, 8], chunk=(n + 1) // 2, keep_state=True, threads=2)
    print(name, "ok", flush=True)
sys.path.insert(0, ROOT + "/ramp_amd/instruments")
import powder_diff
from ramp_amd.diff import DiffInstrument
d, F2 = powder_diff.lines()
di = DiffInstrument.__new__(DiffInstrument)
from tests.test_oracle import _pd_build
di, th0 = _pd_build(4000, 2.0)
dl = np.random.default_rng(0).normal(size=2048 * 256).astype(np.float32)
g = pyoracle.adjoint(di.build(th0), 4000, [3], 0, dl, 44)
assert np.isfinite(g).all()
print("adjoint ok", flush=True)
'''


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_san.so"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan):
        pytest.skip("gcc has no AddressSanitizer runtime")
    env = dict(os.environ)
    env.update(RAMP_ORACLE_LIB=os.path.join(ROOT, "oracle", "liboracle_san.so"), LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + CHILD], env=env, capture_output=True,
                       text=True, timeout=900, cwd=str(tmp_path))
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "adjoint ok" in r.stdout
