"""Write the golden-vector fixtures of tests/golden/.

threefry_golden.txt: output of oracle/_ref/threefry_golden, the driver
compiled (oracle/Makefile, target `ref`) against the reference's own
Random123 headers (mcramp/include/Random123/threefry.h, u01fixedpt.h).
Columns: 4 counter words, 4 key words, 4 output words of threefry4x32
(20 rounds), float bits of u01fixedpt_closed_closed_32_24(word 0).

Run here, where the reference is present:  python tests/golden/make_golden.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "threefry_golden")],
                         capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "threefry_golden.txt"), "w") as f:
        f.write("# c0 c1 c2 c3 k0 k1 k2 k3 r0 r1 r2 r3 u01bits  (reference Random123, 20 rounds)\n")
        f.write(out)
    print("wrote", len(out.splitlines()), "vectors")


if __name__ == "__main__":
    main()
