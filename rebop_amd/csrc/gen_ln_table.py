"""Generates the reduction table of rebop_ln64 (rebop_rng.h) with mpmath.

For x = 2^k m with m in [0.70703125, 1.4140625) (bit pattern 0x3FE6A... to
0x3FF6A...), bucket i = (bits(m) - 0x3FE6A00000000000) >> 45, i = 0..127,
covers 2^45 consecutive doubles.  Entry i holds invc = 1/c rounded to double
for the bucket's midpoint c (1.0 exactly for the two buckets that touch 1.0,
so that ln of values near 1 is r + O(r^2) without cancellation) and
-ln(invc) split into a double and the double nearest its remainder.

Run: python rebop_amd/csrc/gen_ln_table.py  (prints the C table)
"""
import struct

import mpmath as mp

mp.mp.prec = 200


def b2f(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def main() -> None:
    base = 0x3FE6A00000000000
    out = []
    for i in range(128):
        lo, hi = b2f(base + i * (1 << 45)), b2f(base + (i + 1) * (1 << 45))
        invc = 1.0 if (lo == 1.0 or hi == 1.0) else float(2 / (mp.mpf(lo) + mp.mpf(hi)))
        logc = -mp.log(mp.mpf(invc))
        lh = float(logc)
        ll = float(logc - mp.mpf(lh))
        out.append("    {%s, %s, %s, 0.0}," % (invc.hex(), lh.hex(), ll.hex()))
    print("\n".join(out))


if __name__ == "__main__":
    main()
