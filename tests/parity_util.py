"""Helpers shared by the GPU parity tests: run an instrument on the GPU
(libramp.so) and on the CPU oracle with the same seeds, and compare."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


_MESH_DIR = None


def mesh_files():
    """Small noisy icospheres written once per process to a temp dir."""
    global _MESH_DIR
    import tempfile
    from ramp_amd import meshes
    from ramp_amd.components import write_stl_binary
    if _MESH_DIR is None:
        _MESH_DIR = tempfile.mkdtemp(prefix="ramp_mesh_")
        write_stl_binary(os.path.join(_MESH_DIR, "ico12.stl"),
                         meshes.icosphere(12, radius=0.015, sigma=0.001, seed=3))
        write_stl_binary(os.path.join(_MESH_DIR, "inner.stl"),
                         meshes.icosphere(6, radius=0.008, sigma=0.0005, seed=4))
        write_stl_binary(os.path.join(_MESH_DIR, "outer.stl"),
                         meshes.icosphere(8, radius=0.012, sigma=0.0005, seed=5))
    return {k: os.path.join(_MESH_DIR, k + ".stl") for k in ("ico12", "inner", "outer")}


def instrument_cases():
    ang = float(np.arcsin((9.045 / np.sqrt(5.0)) / (2 * 3.355)))
    m = mesh_files()
    return {
        "imaging": (os.path.join(ROOT, "ramp_amd", "instruments", "imaging.json"),
                    {"mesh": repr(m["ico12"])}),
        "mesh_multi": (os.path.join(ROOT, "tests", "instruments", "mesh_multi.json"),
                       {"inner": repr(m["inner"]), "outer": repr(m["outer"])}),
        "minimal": (os.path.join(ROOT, "ramp_amd", "instruments", "minimal.json"), {}),
        "let_like": (os.path.join(ROOT, "ramp_amd", "instruments", "let_like.json"), {}),
        "guide_chopper": (os.path.join(ROOT, "tests", "instruments", "guide_chopper.json"),
                          {"v_foc": 841.3}),
        "guide_variants": (os.path.join(ROOT, "tests", "instruments", "guide_variants.json"), {}),
        "powder": (os.path.join(ROOT, "tests", "instruments", "powder.json"), {"Ei": 25.0}),
        "tas": (os.path.join(ROOT, "tests", "instruments", "tas.json"), {"Mono_angle": ang}),
        "powder_diff": (os.path.join(ROOT, "ramp_amd", "instruments", "powder_diff.json"),
                        _powder_diff_vars()),
        "polarisation": (os.path.join(ROOT, "ramp_amd", "instruments", "polarisation.json"),
                         {"py": 1.0, "ax": 0.6, "ay": 0.8, "az": 0.0}),
        "isis_isotropic": (os.path.join(ROOT, "tests", "instruments", "isis_isotropic.json"),
                           {"spec": repr(table_files()["spectrum"]), "sqw": repr(table_files()["sqw"])}),
    }


def isis_spectrum_file(path, n_e=40, n_t=60):
    """A synthetic moderator spectrum in the McStas/MCNP tally layout the
    MISIS loader reads (isis_mod.py:64-166): a TimeOffset comment, then per
    energy bin (MeV) a time column (units of 1e-8 s) of
    'value flux (error)' lines and a total line."""
    e_edges = np.logspace(-9, -7, n_e + 1)           # 1 .. 100 meV
    t = np.logspace(0, 5, n_t)                       # 1e-8 .. 1e-3 s
    lines = ["1tally 5 nps 1000000", "c synthetic moderator spectrum", "c TimeOffset 100"]
    for k in range(n_e):
        e_mid = 0.5 * (e_edges[k] + e_edges[k + 1]) * 1e9
        lines += [" surface 9999", "segment:  0", "user bin total",
                  f"energy bin: {e_edges[k]:.6e} to {e_edges[k + 1]:.6e}", "cosine:   1", " time "]
        tot = 0.0
        for ti in t:
            # Maxwellian-like in energy, Ikeda-Carpenter-like in time
            f = 1e-9 * np.sqrt(e_mid) * np.exp(-e_mid / 25.0) * (ti / 100.0) * np.exp(-ti / (100.0 + 20 * e_mid))
            tot += f
            lines.append(f"{ti:.6e} {f:.6e} ({0.01 * f:.6e}) ")
        lines.append(f"  total      {tot:.6e} ({0.01 * tot:.6e})")
    lines.append("1end")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def sqw_file(path, rho=0.07, sigma_coh=2.0, sigma_abs=0.3):
    """A synthetic S(q, w) table in the layout of SIsotropic's loader
    (isotropic.py:171-208): '# key value' header lines, a q line, a w line,
    then one row of S per q.  The loader consumes the first row when it
    stops scanning and drops the last q, so the table it keeps is rows
    1..nq-1 against q_0..q_{nq-2}."""
    q = np.linspace(0.2, 6.0, 30)
    w = np.linspace(-2.0, 2.0, 41)
    S = np.exp(-0.5 * (w[None, :] / (0.2 + 0.05 * q[:, None])) ** 2) * np.exp(-0.05 * q[:, None] ** 2)
    with open(path, "w") as fh:
        fh.write(f"# V_rho {rho}\n# sigma_coh {sigma_coh}\n# sigma_inc 0.5\n# sigma_abs {sigma_abs}\n")
        fh.write(" ".join(f"{x:.6f}" for x in q) + "\n")
        fh.write(" ".join(f"{x:.6f}" for x in w) + "\n")
        for row in S:
            fh.write(" ".join(f"{x:.6e}" for x in row) + "\n")


def table_files():
    """Synthetic spectrum and S(q, w) files, written once per process."""
    global _TAB_DIR
    import tempfile
    if _TAB_DIR is None:
        _TAB_DIR = tempfile.mkdtemp(prefix="ramp_tables_")
        isis_spectrum_file(os.path.join(_TAB_DIR, "spectrum.mcstas"))
        sqw_file(os.path.join(_TAB_DIR, "inc.sqw"))
    return {"spectrum": os.path.join(_TAB_DIR, "spectrum.mcstas"), "sqw": os.path.join(_TAB_DIR, "inc.sqw")}


_TAB_DIR = None


_MCPL_DIR = None


def mcpl_fixtures():
    return np.load(os.path.join(ROOT, "tests", "golden", "mcpl_fixtures.npz"))


def mcpl_files():
    """MCPL files written once per process from the fixtures: 'back', the
    McStas-written particles of the reference's MCPLtestback.mcpl, in a
    file with the same header structure (source name, one comment, two
    blobs of the recorded lengths with filler bytes)."""
    global _MCPL_DIR
    import tempfile
    from ramp_amd.mcpl import write_mcpl_full
    if _MCPL_DIR is None:
        _MCPL_DIR = tempfile.mkdtemp(prefix="ramp_mcpl_")
        fx = mcpl_fixtures()
        keys = bytes(fx["back_blob_keys"]).split(b"\0")
        blobs = {k.decode(): b"#" * int(n) for k, n in zip(keys, fx["back_blob_lengths"])}
        write_mcpl_full(os.path.join(_MCPL_DIR, "back.mcpl"), fx["back_records"],
                        source_name=bytes(fx["back_source"]).decode(),
                        comments=[bytes(fx["back_comment"]).decode()], blobs=blobs)
    return {"back": os.path.join(_MCPL_DIR, "back.mcpl"), "dir": _MCPL_DIR}


def mcpl_cases():
    f = mcpl_files()
    return {
        "mcpl_in": (os.path.join(ROOT, "ramp_amd", "instruments", "mcpl_in.json"), {"mcpl_file": repr(f["back"])}),
        "mcpl_out": (os.path.join(ROOT, "ramp_amd", "instruments", "mcpl_out.json"),
                     {"mcpl_file": repr(os.path.join(f["dir"], "out_ref.mcpl")), "standard": 0}),
        "mcpl_out_std": (os.path.join(ROOT, "ramp_amd", "instruments", "mcpl_out.json"),
                         {"mcpl_file": repr(os.path.join(f["dir"], "out_std.mcpl")), "standard": 1}),
    }


def _powder_diff_vars():
    sys.path.insert(0, os.path.join(ROOT, "ramp_amd", "instruments"))
    import powder_diff
    d, F2 = powder_diff.lines()
    return powder_diff.variables(powder_diff.theta0(), F2)


def run_gpu(path, kw, n, seeds, chunk=None, hist_mode="deterministic", keep_state=True):
    import ramp_amd
    inst = ramp_amd.Instrument(path, hist_mode=hist_mode, chunk=chunk, **kw)
    inst.execute(n, keep_state=keep_state, seeds=seeds)
    st = inst.state(inst.last_chunk if n % inst.last_chunk == 0 else n % inst.last_chunk) \
        if keep_state else None
    hists = [inst.hist64(c.name) for c in inst._comp_list if getattr(c.scat, "n_bins", 0)]
    stats = inst.stats()
    inst.close()
    return st, hists, stats


def run_oracle(path, kw, n, seeds, chunk=0):
    import ramp_amd
    from oracle import pyoracle
    return pyoracle.run(ramp_amd.Instrument(path, **kw), n, seeds, chunk=chunk, keep_state=True)


def compare_states(g, o):
    """Counts of differences between GPU and oracle final states."""
    g16, g8, gi = g
    o16, o8, oi = o
    out = {}
    out["dead"] = int((g16[:, 15] != o16[:, 15]).sum())
    out["iidx"] = int((gi != oi).sum())
    out["ctr"] = int((g16[:, 11] != o16[:, 11]).sum())
    bits = (g16.view(np.uint32) != o16.view(np.uint32)) & ~(np.isnan(g16) & np.isnan(o16))
    out["bitdiff_values"] = int(bits.sum())
    out["bitdiff_rows"] = int(bits.any(axis=1).sum())
    with np.errstate(invalid="ignore", over="ignore"):
        scale = np.maximum(np.nanmax(np.abs(o16), axis=0), 1e-30)
        # equal infinities and NaN pairs count as equal
        close = np.isclose(g16, o16, rtol=1e-5, atol=0) | (np.abs(g16 - o16) <= 1e-6 * scale) | \
            (np.isnan(g16) & np.isnan(o16)) | (g16 == o16)
    out["beyond_1e-5"] = int((~close).sum())
    out["beyond_rows"] = int((~close).any(axis=1).sum())
    return out
