#!/usr/bin/env python3
"""Generate the two lookup tables of the photon step (mcml_tables.h).

mcml_sc_table[j] = {cos(2 pi j / 256), sin(2 pi j / 256)}, j = 0..255:
    the azimuth table of mcml_sincos_u32 (angle addition with a short polynomial for the
    remaining angle below 2 pi / 256).

mcml_lg_table[j] = {c_j, L_j}, j = 0..127, with
    c_j = binary32(1 / (1 + j / 128))   (reciprocal of the start of mantissa cell j)
    L_j = binary32(-ln(c_j))            (logarithm of the rounded reciprocal)
    the table of mcml_neglog_u32, after Tang, "Table-driven implementation of the
    logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990: for a
    mantissa m in cell j, r = m c_j - 1 is small (0 <= r < 1/128 up to rounding) and
    ln m = L_j + ln(1 + r).  Cell 0 has c_0 = 1 and L_0 = 0, so a mantissa of exactly 1
    gives ln m = 0 exactly.

Every entry is the double-precision value rounded to nearest binary32, so the table is a
function of the formulas alone; tests/test_detmath.py recomputes it and compares bits.

Usage: python gen_tables.py > mcml_tables.h
"""
import math
import struct

SC_N = 256
LG_N = 128


def f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def lit(x):
    """A binary32 value as an exact C hexadecimal float literal."""
    v = f32(x)
    return "0.0f" if v == 0.0 else float.hex(v) + "f"


def sc_entries():
    return [(f32(math.cos(2 * math.pi * j / SC_N)), f32(math.sin(2 * math.pi * j / SC_N))) for j in range(SC_N)]


def lg_entries():
    out = []
    for j in range(LG_N):
        c = f32(1.0 / (1.0 + j / LG_N))
        out.append((c, f32(-math.log(c))))
    return out


def main():
    print("/* Generated by gen_tables.py; do not edit.  cos/sin of 2 pi j / %d, and the" % SC_N)
    print(" * reciprocal / logarithm pairs of the %d mantissa cells of [1, 2). */" % LG_N)
    print("#ifndef MCML_TABLES_H")
    print("#define MCML_TABLES_H")
    print('#include "mcml_detmath.h" /* mcml_pair */')
    print("#if defined(__HIPCC__)")
    print("#define MCML_TABLE __device__ static const")
    print("#else")
    print("#define MCML_TABLE static const")
    print("#endif")
    print("#define MCML_SC_N %d" % SC_N)
    print("#define MCML_LG_N %d" % LG_N)
    print("MCML_TABLE mcml_pair mcml_sc_table[MCML_SC_N] = {")
    for c, s in sc_entries():
        print("    {%s, %s}," % (lit(c), lit(s)))
    print("};")
    print("MCML_TABLE mcml_pair mcml_lg_table[MCML_LG_N] = {")
    for c, L in lg_entries():
        print("    {%s, %s}," % (lit(c), lit(L)))
    print("};")
    print("#endif")


if __name__ == "__main__":
    main()
