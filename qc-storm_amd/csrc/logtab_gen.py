"""Write qcs_logtab.inc: the logarithm table of the one-emitter fitting loss.

For a positive float x with bit pattern b, g = (int)(b - K) >> 16 with K = 0x3f368000 numbers
the interval of x: binade e = g >> 7 of the reduction x = m 2^e, m in [0.712890625, 1.42578125),
and one of 128 equal bit-pattern steps of m.  K's low 16 bits (0x8000) put 1.0 in the middle of
interval 73 of binade 0, [0.998046875, 1.00390625), whose pair is (1, 0): log x = log1p(x - 1)
there, with x - 1 exact, so there is no cancellation next to 1.  For e in [-10, 21] the table
holds, at index g + 1280, the pair (i, t): i a float close to the reciprocal of the interval's
centre, and t = -log(i) rounded to float.  Among the 8001 floats nearest that reciprocal, i is
the one whose -log(i) (evaluated in double) lies closest to a float, relative to that float's
ulp (accurate tables; every entry ends within 1e-3 ulp of -log(i)).

Then log x = t + log1p(r), r = x i - 1, |r| < 2^-8 + 4e-4.  K's low 16 bits being 0x8000 and the
step 2^16 make the byte offset of a pair one add and one shift of the upper half-word.

Run: python3 logtab_gen.py > qcs_logtab.inc  (the committed file is this script's output).
"""
import sys

import numpy as np

K = 0x3F368000
EMIN, EMAX = -10, 21
SH = 16
PER = 128          # intervals per binade (2^(23 - SH))


def _f(bits):
    return float(np.array([bits], np.uint32).view(np.float32)[0])


def _bits(v):
    return int(np.array([v], np.float32).view(np.uint32)[0])


def table():
    out = []
    for e in range(EMIN, EMAX + 1):
        for j in range(PER):
            g = e * PER + j
            lo_b = (K + (g << SH)) & 0xFFFFFFFF
            lo, hi = _f(lo_b), _f(lo_b + (1 << SH) - 1)
            if lo < 1.0 < hi:
                assert e == 0 and j == 73
                out.append((_bits(1.0), 0))
                continue
            b0 = _bits(np.float32(2.0 / (lo + hi)))
            cand = np.arange(b0 - 4000, b0 + 4001, dtype=np.int64).astype(np.uint32)
            inv = cand.view(np.float32).astype(np.float64)
            t = -np.log(inv)
            tf = t.astype(np.float32)
            err = np.abs(t - tf.astype(np.float64)) / np.spacing(np.abs(tf)).astype(np.float64)
            k = int(np.argmin(err))
            out.append((int(cand[k]), _bits(tf[k])))
    return out


def main():
    rows = table()
    print("/* qcs_logtab.inc -- written by logtab_gen.py: %d (1/c, -log(1/c)) pairs as float bit patterns,"
          % len(rows))
    print(" * binades %d..%d of x = m 2^e, m in [0.712890625, 1.42578125), %d intervals each */" % (EMIN, EMAX, PER))
    words = [w for pair in rows for w in pair]
    for i in range(0, len(words), 8):
        print("    " + ", ".join("0x%08xu" % w for w in words[i:i + 8]) + ",")


if __name__ == "__main__":
    sys.exit(main())
