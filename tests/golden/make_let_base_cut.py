"""Cut a small fixture from the reference's moderator spectrum file
examples/Let_Base.mcstas (McStas/MCNP tally layout read by MISIS,
moderator/isis_mod.py:64-166): its seven header lines (tally, comments,
TimeOffset) and the energy-bin blocks 43..52 (1.8957e-9 to 6.8129e-9 MeV,
i.e. about 1.9 to 6.8 meV), each block unchanged (surface, segment, user bin,
energy bin, cosine and time lines, 352 time rows, total line), then the
closing line.  Written gzip-compressed to tests/golden/let_base_cut.mcstas.gz.

Run here, where the reference tree exists:
    python tests/golden/make_let_base_cut.py
"""
import gzip
import os

SRC = "/root/reference/examples/Let_Base.mcstas"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "let_base_cut.mcstas.gz")
FIRST, LAST = 43, 52   # 1-based energy-bin blocks kept


def main():
    lines = open(SRC).read().split("\n")
    starts = [i for i, l in enumerate(lines) if l.startswith(" surface")]
    header = lines[:starts[0]]
    blocks = []
    for k, s in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else next(i for i in range(s, len(lines)) if lines[i].startswith("1end"))
        blocks.append(lines[s:e])
    out = header + sum(blocks[FIRST - 1:LAST], []) + ["1end", ""]
    with gzip.open(OUT, "wt") as f:
        f.write("\n".join(out))
    print(f"{OUT}: {len(out)} lines, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
