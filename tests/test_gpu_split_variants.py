"""GPU: the split-FP32 external product in its four measured forms -- key halves split at 24 bits
(the default: hi half 8 bits) or at 16 bits (two 16-bit halves), each with plain FP32 FMA
accumulation or compensated accumulation (TwoProduct + TwoSum) -- against the exact int64
negacyclic product of the oracle.  The 24-bit forms keep the stated 2^16 bound (2^-16 of the torus)
and gates decrypt; the 16-bit forms are bounded only loosely: their hi product (about 2^25.7) no
longer fits FP32's 24-bit significand, so it cannot round to its exact integer, and each unit of its
error is worth 2^16."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_1024
N = P.N


def T():
    import tfhe_amd
    return tfhe_amd


def variant_key(ks, bits):
    T().set_split_bits(bits)
    try:
        return T().DeviceKey.from_keyset(ks)
    finally:
        T().set_split_bits(0)


def product_error(dk, keys, acc, rows):
    got = dk.extern_product(acc, rows, prec=1)
    e = [np.abs((got[g].astype(np.int64) - O.extern_product_exact(P, acc[g], keys.bk[rows[g]]).astype(np.int64)
                 + 2**31) % 2**32 - 2**31).ravel() for g in range(acc.shape[0])]
    e = np.concatenate(e).astype(np.float64)
    return e.max(), float(np.sqrt((e ** 2).mean()))


def test_split_variants_error_bounds(gpu_env):
    ks, _, keys = gpu_env
    rng = np.random.default_rng(4040)
    acc = rng.integers(-2**31, 2**31, (6, 2, N), dtype=np.int64).astype(np.int32)
    rows = np.array([0, 9, 77, 300, 512, 629], np.int32)
    res = {}
    for bits in (24, 16):
        dk = variant_key(ks, bits)
        try:
            assert dk.split_bits == bits
            for mac in (0, 1):
                T().set_split_mac(mac)
                try:
                    print(f"[split variants] split {bits} mac {mac}", flush=True)
                    res[(bits, mac)] = product_error(dk, keys, acc, rows)
                finally:
                    T().set_split_mac(0)
        finally:
            dk.free()
    for k, (mx, rms) in res.items():
        print(f"split {k[0]} {'compensated' if k[1] else 'plain'}: external product error max {mx:.0f} rms {rms:.1f}")
    for mac in (0, 1):
        assert res[(24, mac)][0] < 2**16
        assert res[(16, mac)][0] < 2**24
    # compensation removes the MAC's own rounding error; the transforms' error remains
    assert res[(24, 1)][1] < 1.2 * res[(24, 0)][1]


def test_compensated_split_gates_and_hi_margin(gpu_env):
    """The compensated form at the default split: NAND and MUX gates on a launch of 300 (one
    bootstrap per wavefront) decrypt correctly, and the hi-half product before rounding stays within
    0.25 of its integer over 32,768 coefficients."""
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(4041)
    bits = rng.integers(0, 2, (3, 300)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    T().set_split_mac(1)
    try:
        print("[split variants] compensated NAND/MUX x300", flush=True)
        out = dk.gate("NAND", ca, cb, prec=1)
        assert np.array_equal(ks.decrypt(out), 1 - (bits[0] & bits[1]))
        out = dk.gate("MUX", ca, cb, cc, prec=1)
        assert np.array_equal(ks.decrypt(out), np.where(bits[0] == 1, bits[1], bits[2]))
        acc = rng.integers(-2**31, 2**31, (16, 2, N), dtype=np.int64).astype(np.int32)
        rows = rng.integers(0, P.n, 16).astype(np.int32)
        _, hi = dk.extern_product_split_hi(acc, rows)
        d = np.abs(hi - np.rint(hi))
        print(f"compensated split 24: hi distance to an integer max {d.max():.4f}")
        assert d.max() < 0.25
    finally:
        T().set_split_mac(0)
