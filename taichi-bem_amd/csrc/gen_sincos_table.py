"""Writes bem_sincos_table.h: {sin, cos}(2 pi j / 2048), j < 2048, each
rounded to nearest from a 200-bit evaluation (mpmath).  The header is
committed; run this only to regenerate it:  python3 gen_sincos_table.py
"""
import os

import mpmath as mp

N = 2048
mp.mp.prec = 200


def main():
    rows = []
    for j in range(N):
        a = 2 * mp.pi * j / N
        rows.append("{%r, %r}" % (float(mp.sin(a)), float(mp.cos(a))))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bem_sincos_table.h")
    with open(out, "w") as f:
        f.write("// Generated by gen_sincos_table.py: {sin, cos}(2 pi j / %d), j < %d,\n" % (N, N))
        f.write("// correctly rounded to double (200-bit evaluation).\n")
        f.write("#pragma once\n\nnamespace bem {\n\n")
        f.write("static __constant__ const double2 kSinCos2048[%d] = {\n" % N)
        for k in range(0, N, 2):
            f.write("    " + ", ".join(rows[k:k + 2]) + ",\n")
        f.write("};\n\n}  // namespace bem\n")


if __name__ == "__main__":
    main()
