"""Writes sw_math_tables.h, the tables of sw_math.h, from 60-digit decimal
arithmetic (run once; the header is committed):

    python csrc/gen_sw_math_tables.py

- SWM_EXP_TAB[256]: 2^(k/128) ~= H_k (1 + T_k), k = 0..127, H_k the double
  nearest 2^(k/128) and T_k the double nearest 2^(k/128) / H_k - 1, stored as
  bits(T_k), bits(H_k) - k 2^45 (the layout of the table-driven exp of
  Tang's method as in ARM's optimized-routines, whose exp glibc ships);
- SWM_LOG_HI / SWM_LOG_LO[363]: log(1 + j / 512), j = -150..212, as the
  double nearest it and the double nearest the remainder (a double-double);
- SWM_LN2_HI / SWM_LN2_LO: ln 2 split so that k * SWM_LN2_HI is exact for
  |k| < 2^11.
"""
import os
import struct
from decimal import Decimal, getcontext

getcontext().prec = 60
HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main():
    out = ["// generated by gen_sw_math_tables.py -- do not edit", "#pragma once", ""]
    tab = []
    for k in range(128):
        v = Decimal(2) ** (Decimal(k) / Decimal(128))
        H = float(v)
        T = float((v - Decimal(H)) / Decimal(H))
        tab += [bits(T), (bits(H) - (k << 45)) & 0xFFFFFFFFFFFFFFFF]
    out.append("SWM_TAB uint64_t SWM_EXP_TAB[256] = {")
    for i in range(0, 256, 4):
        out.append("    " + ", ".join("0x%016xULL" % t for t in tab[i:i + 4]) + ",")
    out.append("};")
    his, los = [], []
    for j in range(-150, 213):
        L = (Decimal(1) + Decimal(j) / Decimal(512)).ln()
        hi = float(L)
        los.append(float(L - Decimal(hi)))
        his.append(hi)
    for name, arr in (("SWM_LOG_HI", his), ("SWM_LOG_LO", los)):
        out.append(f"SWM_TAB double {name}[363] = {{")
        for i in range(0, len(arr), 4):
            out.append("    " + ", ".join(x.hex() for x in arr[i:i + 4]) + ",")
        out.append("};")
    ln2 = Decimal(2).ln()
    hi = float(ln2)
    hb = bits(hi) & ~((1 << 11) - 1)  # 11 low bits dropped: a 42-bit head
    hi = struct.unpack("<d", struct.pack("<Q", hb))[0]
    lo = float(ln2 - Decimal(hi))
    out += [f"#define SWM_LN2_HI {hi.hex()}", f"#define SWM_LN2_LO {lo.hex()}", ""]
    with open(os.path.join(HERE, "sw_math_tables.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
