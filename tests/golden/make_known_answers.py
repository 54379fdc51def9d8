"""Writes tests/golden/known_answers.json: closed-form values for the torus numerics, the gadget
decomposition and the gate constants, computed here from their mathematical definitions with
exact rational arithmetic (independently of the oracle's C code, which reproduces the reference's
integer formulas).  The reference itself commits no golden vectors; the values it pins directly
are included and cited:
  offset 2164391936 for l=3, Bgbit=7      kernel/src/include/constants.h:24
  MU = modSwitchToTorus32(1, 8)           kernel/src/tfhe_blindRotate_FFT/tfhe_blindRotate_FFT_tb.cpp:267
  OR(499, 388) over 16 bits = 503         test_bench/client/alice.c:33-44, verify.c:31

Run:  python tests/golden/make_known_answers.py
"""
import json
import os
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
TWO32 = 1 << 32


def to_signed(x):
    x %= TWO32
    return x - TWO32 if x >= 1 << 31 else x


def modswitch_from(phase, msize):
    """Nearest element of Z_msize to the torus value phase / 2^32 (ties toward +), reduced mod msize."""
    t = Fraction(phase % TWO32, TWO32) * msize
    k = (t + Fraction(1, 2)).numerator // (t + Fraction(1, 2)).denominator  # floor(t + 1/2)
    return k % msize


def modswitch_to(mu, msize):
    """floor(mu * 2^32 / msize) on the torus (exact when msize divides 2^32)."""
    return to_signed((mu * TWO32) // msize)


def approx_phase(phase, msize):
    """Torus value of the nearest multiple of 1/msize."""
    return to_signed(modswitch_from(phase, msize) * TWO32 // msize)


def dtot32(d):
    """Fractional part of d (sign of d kept) scaled to 2^32, as the reference's truncation does."""
    f = Fraction(d) - int(Fraction(d))  # int() truncates toward zero
    return to_signed(int(f * TWO32))


def decomp_offset(l, bgbit):
    half = 1 << (bgbit - 1)
    return sum(half * (1 << (32 - (p + 1) * bgbit)) for p in range(l)) % TWO32


def main():
    edges = [0, 1, -1, 2**31 - 1, -2**31, 2**20, 2**20 - 1, -2**20, 2**21, 2**21 + 1, 3 * 2**20, 123456789,
             -987654321, 2**30, -2**30, 2**31 - 2**20, -2**31 + 2**20 - 1]
    out = {
        "modswitch_from": [[x, m, modswitch_from(x, m)] for m in (2048, 4096, 8, 4) for x in edges],
        "modswitch_to": [[mu, m, modswitch_to(mu, m)] for m in (8, 4, 2048) for mu in (1, -1, 3, 0, 7)],
        "approx_phase": [[x, m, approx_phase(x, m)] for m in (8, 4) for x in edges],
        "dtot32": [[d, dtot32(d)] for d in (0.0, 0.25, -0.25, 0.5, 0.125, -0.125, 1.75, -3.0625, 2.0 ** -15,
                                            -(2.0 ** -25))],
        "decomp_offset": [[3, 7, decomp_offset(3, 7)], [3, 8, decomp_offset(3, 8)], [2, 10, decomp_offset(2, 10)]],
        "mu": modswitch_to(1, 8),
        "gate_constants": {"1/8": modswitch_to(1, 8), "-1/8": modswitch_to(-1, 8), "1/4": modswitch_to(1, 4),
                           "-1/4": modswitch_to(-1, 4)},
        "known_answer_or": {"a": 499, "b": 388, "bits": 16, "result": 499 | 388},
        "pinned_by_reference": {"offset_l3_bg7": 2164391936, "mu": 536870912},
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
