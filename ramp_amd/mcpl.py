"""MCPL particle files (format version 3): the host side of SMCPLIn and
SMCPLOut.

The reference reads files through the `mcpl` Python package
(scat/mcpl_in.py:7, 60-66) and writes them itself (scat/mcpl_out.py:7-117).
Here the reader parses the published format - header, then fixed-size
little- or big-endian particle records - into a canonical array of 11
values per particle, without arithmetic:

    [px, py, pz, x, y, z, u1, u2, e, t, w]
     polarisation, position (cm), packed direction (u1, u2) and kinetic
     energy e (MeV) whose sign bit is the sign of the dropped direction
     component, time (ms), weight

(float32 for single-precision files, float64 otherwise).  The unpacking to
neutron states runs on the GPU, in the SMCPLIn step.  The writer produces
the reference's files: its 56-byte header (write_hdr, mcpl_out.py:7-34)
followed by the records packed on the GPU.
"""
import gzip
import struct

import numpy as np

MAGIC = b"MCPL"
RECORD_FIELDS = 11
PDG_NEUTRON = 2112


class MCPLFormatError(ValueError):
    pass


def _open_bytes(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:2] == b"\x1f\x8b":
        data = gzip.decompress(data)
    return data


def parse_header(data):
    """Header of an MCPL file (bytes).  Returns a dict with the particle
    count, the option flags, the strings and blobs, the record size and
    the offset of the first record."""
    if len(data) < 48 or data[:4] != MAGIC:
        raise MCPLFormatError("not an MCPL file")
    version = data[4:7]
    if version != b"003":
        raise MCPLFormatError(f"MCPL format version {version.decode(errors='replace')} is not supported "
                              "(only version 3)")
    endian = chr(data[7])
    if endian not in "LB":
        raise MCPLFormatError(f"bad endianness marker {endian!r}")
    e = "<" if endian == "L" else ">"
    off = 8
    (n,) = struct.unpack_from(e + "Q", data, off)
    off += 8
    ncomments, nblobs, userflags, pol, single, updg, psize, uweight = struct.unpack_from(
        e + "IIIIIiII", data, off)
    off += 32
    universal_weight = None
    if uweight:
        (universal_weight,) = struct.unpack_from(e + "d", data, off)
        off += 8

    def buf():
        nonlocal off
        (ln,) = struct.unpack_from(e + "I", data, off)
        off += 4
        if off + ln > len(data):
            raise MCPLFormatError("truncated header")
        b = data[off:off + ln]
        off += ln
        return b

    source = buf().decode("latin-1")
    comments = [buf().decode("latin-1") for _ in range(ncomments)]
    keys = [buf().decode("latin-1") for _ in range(nblobs)]
    blobs = [buf() for _ in range(nblobs)]
    fsize = 4 if single else 8
    expect = fsize * (3 * bool(pol) + 3 + 3 + 1 + (0 if uweight else 1)) + \
        (0 if updg else 4) + (4 if userflags else 0)
    if psize != expect:
        raise MCPLFormatError(f"particle size {psize} does not match the options ({expect})")
    return {"n_particles": n, "endian": endian, "polarisation": bool(pol), "single_precision": bool(single),
            "universal_pdgcode": updg, "universal_weight": universal_weight, "userflags": bool(userflags),
            "particle_size": psize, "source_name": source, "comments": comments,
            "blob_keys": keys, "blobs": blobs, "data_offset": off}


def records_from_bytes(raw, hdr, n=None):
    """Canonical (n, 11) records from the raw particle bytes of a file with
    header `hdr` (values copied, never computed)."""
    e = "<" if hdr["endian"] == "L" else ">"
    ft = e + ("f4" if hdr["single_precision"] else "f8")
    fields = []
    if hdr["polarisation"]:
        fields += [("px", ft), ("py", ft), ("pz", ft)]
    fields += [("x", ft), ("y", ft), ("z", ft), ("u1", ft), ("u2", ft), ("e", ft), ("t", ft)]
    if hdr["universal_weight"] is None:
        fields.append(("w", ft))
    if not hdr["universal_pdgcode"]:
        fields.append(("pdg", e + "i4"))
    if hdr["userflags"]:
        fields.append(("uf", e + "u4"))
    dt = np.dtype(fields)
    assert dt.itemsize == hdr["particle_size"]
    count = len(raw) // dt.itemsize if n is None else int(n)
    if count * dt.itemsize > len(raw):
        raise MCPLFormatError(f"file holds {len(raw) // dt.itemsize} of {count} particles")
    p = np.frombuffer(raw, dtype=dt, count=count)
    out_t = np.float32 if hdr["single_precision"] else np.float64
    rec = np.zeros((count, RECORD_FIELDS), dtype=out_t)
    for k, f in enumerate(("px", "py", "pz", "x", "y", "z", "u1", "u2", "e", "t")):
        if f in dt.names:
            rec[:, k] = p[f]
    rec[:, 10] = p["w"] if "w" in dt.names else hdr["universal_weight"]
    return rec


def read_mcpl(path):
    """(header, records) of an MCPL file (optionally gzipped)."""
    data = _open_bytes(path)
    hdr = parse_header(data)
    rec = records_from_bytes(data[hdr["data_offset"]:], hdr, hdr["n_particles"])
    return hdr, rec


def reference_header(count, source_name="RAMP"):
    """write_hdr (mcpl_out.py:7-34): version 3, little endian, no comments
    or blobs, polarisation, single precision, universal PDG code 2112,
    44-byte records, per-particle weights."""
    name = source_name.encode("ascii")
    return (b"MCPL003L" + struct.pack("<Q", int(count)) +
            struct.pack("<IIIIIIII", 0, 0, 0, 1, 1, PDG_NEUTRON, 44, 0) +
            struct.pack("<I", len(name)) + name)


def write_mcpl(path, records, source_name="RAMP"):
    """The reference's file: its header and the float32 records (n, 11)."""
    rec = np.ascontiguousarray(records, dtype="<f4").reshape(-1, RECORD_FIELDS)
    with open(path, "wb") as f:
        f.write(reference_header(rec.shape[0], source_name))
        f.write(rec.tobytes())


def write_mcpl_full(path, records, source_name="RAMP", comments=(), blobs=None,
                    universal_weight=None, pdgcodes=None, userflags=None, single_precision=True,
                    polarisation=True, endian="L"):
    """A general MCPL format-3 file (every header option), from canonical
    records.  Used to exercise the reader on layouts the reference never
    writes."""
    e = "<" if endian == "L" else ">"
    rec = np.asarray(records, dtype=np.float64).reshape(-1, RECORD_FIELDS)
    blobs = dict(blobs or {})
    ft = e + ("f4" if single_precision else "f8")
    fields = []
    if polarisation:
        fields += [("px", ft), ("py", ft), ("pz", ft)]
    fields += [("x", ft), ("y", ft), ("z", ft), ("u1", ft), ("u2", ft), ("e", ft), ("t", ft)]
    if universal_weight is None:
        fields.append(("w", ft))
    if pdgcodes is not None:
        fields.append(("pdg", e + "i4"))
    if userflags is not None:
        fields.append(("uf", e + "u4"))
    dt = np.dtype(fields)
    p = np.zeros(rec.shape[0], dtype=dt)
    for k, f in enumerate(("px", "py", "pz", "x", "y", "z", "u1", "u2", "e", "t", "w")):
        if f in dt.names:
            p[f] = rec[:, k]
    if pdgcodes is not None:
        p["pdg"] = pdgcodes
    if userflags is not None:
        p["uf"] = userflags

    def buf(b):
        b = b.encode("latin-1") if isinstance(b, str) else bytes(b)
        return struct.pack(e + "I", len(b)) + b

    hdr = b"MCPL003" + endian.encode() + struct.pack(e + "Q", rec.shape[0])
    hdr += struct.pack(e + "IIIIIiII", len(comments), len(blobs), 1 if userflags is not None else 0,
                       1 if polarisation else 0, 1 if single_precision else 0,
                       0 if pdgcodes is not None else PDG_NEUTRON, dt.itemsize,
                       1 if universal_weight is not None else 0)
    if universal_weight is not None:
        hdr += struct.pack(e + "d", universal_weight)
    hdr += buf(source_name)
    for c in comments:
        hdr += buf(c)
    for k in blobs:
        hdr += buf(k)
    for k in blobs:
        hdr += buf(blobs[k])
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(p.tobytes())
