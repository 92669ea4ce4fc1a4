"""Write tests/golden/mcpl_fixtures.npz from the reference's two MCPL files.

examples/MCPLtestback.mcpl (written by McStas 2.6's mcpl_out, read by the
reference's SMCPLIn in examples/MCPL_testback.json) and
examples/MCPLtest.mcpl (written by the reference's SMCPLOut in
examples/MCPL_test.json).  Kept: the header fields and the particle
records as numbers, and the reference file's 56 header bytes.  Not kept:
the McStas file's embedded instrument text (its blobs), of which only the
keys and lengths are recorded.

The records are taken straight from the bytes at the header's data offset
(little-endian float32, 11 per particle, as both headers declare), not
through ramp_amd.mcpl, so the reader can be tested against them.

Run here, where the reference is present:  python tests/golden/make_mcpl_fixtures.py
"""
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/examples"


def header_fields(b):
    n = struct.unpack_from("<Q", b, 8)[0]
    arr = struct.unpack_from("<IIIIIiII", b, 16)
    off = 48
    strings = []
    for _ in range(1 + arr[0] + 2 * arr[1]):
        ln = struct.unpack_from("<I", b, off)[0]
        strings.append(b[off + 4:off + 4 + ln])
        off += 4 + ln
    return n, arr, strings, off


def main():
    out = {}
    b = open(os.path.join(REF, "MCPLtestback.mcpl"), "rb").read()
    n, arr, strings, off = header_fields(b)
    out["back_n"] = np.uint64(n)
    out["back_opts"] = np.array(arr, dtype=np.int64)  # ncomments nblobs userflags pol single pdg size uweight
    out["back_data_offset"] = np.int64(off)
    out["back_source"] = np.frombuffer(strings[0], np.uint8)
    out["back_comment"] = np.frombuffer(strings[1], np.uint8)
    out["back_blob_keys"] = np.frombuffer(b"\0".join(strings[2:2 + arr[1]]), np.uint8)
    out["back_blob_lengths"] = np.array([len(s) for s in strings[2 + arr[1]:]], dtype=np.int64)
    out["back_records"] = np.frombuffer(b, "<f4", count=11 * n, offset=off).reshape(n, 11).copy()
    b = open(os.path.join(REF, "MCPLtest.mcpl"), "rb").read()
    n, arr, strings, off = header_fields(b)
    assert off == 56
    out["ramp_header"] = np.frombuffer(b[:56], np.uint8).copy()
    out["ramp_n"] = np.uint64(n)
    out["ramp_records"] = np.frombuffer(b, "<f4", count=11 * 4000, offset=off).reshape(4000, 11).copy()
    np.savez_compressed(os.path.join(HERE, "mcpl_fixtures.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
