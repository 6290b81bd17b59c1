"""Write prh_sincos_table.h: sin and cos of i/128, i = 0..111, as double-doubles.

Each entry is (hi, lo) with hi the double nearest the exact value and lo the
double nearest the remainder, evaluated with mpmath at 300 bits.  These are
the table nodes of the dbl-64 sin/cos algorithm restated in prh_libm.h
(glibc sysdeps/ieee754/dbl-64/s_sin.c, IBM Accurate Mathematical Library),
whose table holds the same values.  Run from this directory:

    python gen_sincos_table.py
"""

import mpmath

NODES = 112  # |x| < 0.855469 + 1/256 rounds to at most node 110


def main() -> None:
    mpmath.mp.prec = 300
    lines = [
        "// Generated by gen_sincos_table.py: {sin, sin remainder, cos, cos remainder} of i/128.",
        "#pragma once",
        "",
        "namespace prh {",
        "",
        f"constexpr int SINCOS_NODES = {NODES};",
        "",
        f"constexpr double SINCOS_TABLE[{NODES}][4] = {{",
    ]
    for i in range(NODES):
        x = mpmath.mpf(i) / 128
        row = []
        for fn in (mpmath.sin, mpmath.cos):
            exact = fn(x)
            hi = float(exact)
            lo = float(exact - mpmath.mpf(hi))
            row += [hi, lo]
        lines.append("    {" + ", ".join(v.hex() for v in row) + "},")
    lines += ["};", "", "}  // namespace prh", ""]
    with open("prh_sincos_table.h", "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    main()
