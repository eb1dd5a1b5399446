#!/usr/bin/env python3
"""Generates wmh_log_table.h: the 512-entry range-reduction table of wmh_logf.

The table is derived from exact decimal arithmetic (Python's `decimal` module at
60 significant digits), so it does not depend on any libm.  Entry i covers a
sub-interval of the reduced mantissa m in [0.75, 1.5):

  i in [0, 256):   m in [0.75 + i*2^-10, 0.75 + (i+1)*2^-10)
  i in [256, 512): m in [1 + (i-256)*2^-9, 1 + (i-255)*2^-9)

and stores the centre c (exactly representable), RN(1/c), and ln(c) split into
a float pair hi + lo.  The two sub-intervals touching 1 use c = 1 exactly so
that inputs near 1 keep full relative accuracy (ln c = 0, r = m - 1 exactly).

Run:  python3 gen_log_table.py > wmh_log_table.h
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 60


def f32(x):
    """Round a Decimal/float to the nearest float32 (ties to even) via float64.

    Going through float64 first is a double rounding; the values here are
    either exactly representable (c) or checked below against the exact
    decimal value so a double-rounding error would be caught."""
    return struct.unpack("<f", struct.pack("<f", float(x)))[0]


def f32_exact_rn(d):
    """Round a Decimal to the nearest float32 exactly (ties to even)."""
    if d == 0:
        return 0.0
    cand = f32(d)
    # check neighbours: pick the float32 closest to d
    bits = struct.unpack("<I", struct.pack("<f", cand))[0]
    best = cand
    best_err = abs(Decimal(cand) - d)
    for nb in (bits - 1, bits + 1):
        v = struct.unpack("<f", struct.pack("<I", nb & 0xFFFFFFFF))[0]
        err = abs(Decimal(v) - d)
        if err < best_err:
            best, best_err = v, err
    return best


def hexf(v):
    return float(v).hex() + "f"


def main():
    rows = []
    for i in range(512):
        if i < 256:
            lo = Decimal("0.75") + Decimal(i) / Decimal(1024)
            w = Decimal(1) / Decimal(1024)
        else:
            lo = Decimal(1) + Decimal(i - 256) / Decimal(512)
            w = Decimal(1) / Decimal(512)
        if i == 255 or i == 256:
            c = Decimal(1)
        else:
            c = lo + w / 2
        cf = f32(c)
        assert Decimal(cf) == c, (i, c)
        inv = f32_exact_rn(Decimal(1) / c)
        ln = c.ln()
        hi = f32_exact_rn(ln)
        lo_part = f32_exact_rn(ln - Decimal(hi))
        rows.append((cf, inv, hi, lo_part))
    print("// Generated by gen_log_table.py -- do not edit by hand.")
    print("// {c, RN(1/c), RN(ln c), RN(ln c - RN(ln c))} for the 512 sub-intervals")
    print("// of the reduced mantissa used by wmh_logf.")
    print("#pragma once")
    print("WMH_TABLE_QUALIFIER const float wmh_log_table[512][4] = {")
    for cf, inv, hi, lo_part in rows:
        print("  {%s, %s, %s, %s}," % (hexf(cf), hexf(inv), hexf(hi), hexf(lo_part)))
    print("};")


if __name__ == "__main__":
    main()
