"""CPU: the checks that pin the oracle catch plausible mistakes.  Each case plants one mistake in
the oracle (a wrong sign, a dropped term, an index off by one, a transposed operand, a wrong
twiddle sign), re-runs the check that should catch it, and requires that check to fail."""
import numpy as np
import pytest

import oracle as O

P = O.PARAMS_1024
N = P.N
TWO32 = 2**32


def sdiff(a, b):
    return ((np.asarray(a, np.int64) - np.asarray(b, np.int64) + 2**31) % TWO32) - 2**31


def rotation_ok():
    n = 16
    rng = np.random.default_rng(0)
    src = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    for a in range(2 * n):
        mono = np.zeros(n, dtype=object)
        mono[a % n] = 1 if a < n else -1
        full = np.convolve(mono, src.astype(object))
        want = full[:n].copy()
        want[: n - 1] -= full[n:]
        want = np.array([((int(v) - int(s) + 2**31) % TWO32) - 2**31 for v, s in zip(want, src)], np.int32)
        if not np.array_equal(O.poly_mul_xai_minus_one(a, src), want):
            return False
    return True


def decomposition_ok():
    rng = np.random.default_rng(1)
    x = rng.integers(-2**31, 2**31, 512, dtype=np.int64).astype(np.int32)
    d = O.decompose(x, 3, 7).astype(np.int64)
    if d.min() < -64 or d.max() >= 64:
        return False
    r = (x.astype(np.int64) - (d[0] * 2**25 + d[1] * 2**18 + d[2] * 2**11)) % TWO32
    return bool(np.array_equal(r, x.astype(np.int64) % 2**11))


def extract_ok(K):
    rng = np.random.default_rng(2)
    acc = rng.integers(-2**31, 2**31, (2, N), dtype=np.int64).astype(np.int32)
    return O.lwe_phase(O.extract(P, acc)[None], K.tlwe_key)[0] == O.tlwe_phase(P, acc, K.tlwe_key)[0]


def noiseless_ksk(K, rng):
    """Keyswitch key with zero noise: ks[i][j][h] = (a, a.s + h s'_i 2^(32-2(j+1)))."""
    kn, t, base, n1 = K.ksk.shape
    a = rng.integers(-2**31, 2**31, (kn, t, base, n1 - 1), dtype=np.int64)
    msg = (K.tlwe_key.astype(np.int64)[:, None, None] * np.arange(base)[None, None, :]
           * (1 << (32 - 2 * (np.arange(t)[None, :, None] + 1))))
    b = (a * K.lwe_key.astype(np.int64)).sum(axis=-1) + msg
    out = np.concatenate([a, b[..., None]], axis=-1)
    out[:, :, 0, :] = 0
    return (((out + 2**31) % TWO32) - 2**31).astype(np.int32)


def keyswitch_exact_ok(K):
    """With a noiseless key and input coefficients on the 2^-16 grid (no rounding), the keyswitch
    preserves the phase exactly."""
    rng = np.random.default_rng(3)
    K2 = O.Keys(P, K.lwe_key, K.tlwe_key, noiseless_ksk(K, rng), K.bk)
    for _ in range(3):
        a = (rng.integers(-2**15, 2**15, N).astype(np.int64) << 16)
        m = int(rng.integers(-2**31, 2**31))
        b = (m + int((a * K.tlwe_key).sum())) % TWO32
        u = np.append((((a + 2**31) % TWO32) - 2**31).astype(np.int32), np.int32(((b + 2**31) % TWO32) - 2**31))
        out = O.keyswitch(K2, u)
        if sdiff(O.lwe_phase(out[None], K.lwe_key), m)[0] != 0:
            return False
    return True


def transform_ok():
    rng = np.random.default_rng(4)
    p = rng.integers(-64, 64, 64).astype(np.int32)
    j = np.arange(64)
    k = np.arange(32)
    ref = (p[None, :] * np.exp(1j * np.pi * ((j[None, :] * (4 * k + 1)[:, None]) % 128) / 64)).sum(axis=1)
    return np.abs(O.fft_forward(p) - ref).max() < 1e-9


def extern_product_ok(K):
    rng = np.random.default_rng(5)
    msg = (rng.integers(0, 8, N) * (1 << 29)).astype(np.int64)
    a = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
    b = ((msg + O.negacyclic_mul_naive(K.tlwe_key, a).astype(np.int64) + 2**31) % TWO32 - 2**31).astype(np.int32)
    i = int(np.nonzero(K.lwe_key)[0][0])
    out = O.extern_product_exact(P, np.stack([a, b]), K.bk[i])
    e = sdiff(O.tlwe_phase(P, out, K.tlwe_key), msg) / TWO32
    return np.abs(e).max() < 2.0 ** -10


CASES = {
    1: ("rotation sign at the wrap", lambda K: rotation_ok()),
    2: ("decomposition offset dropped", lambda K: decomposition_ok()),
    3: ("extraction index off by one", extract_ok),
    4: ("keyswitch last level dropped", keyswitch_exact_ok),
    5: ("transform twiddle sign", lambda K: transform_ok()),
    6: ("external product operands transposed", extern_product_ok),
}


@pytest.mark.parametrize("fault", sorted(CASES))
def test_planted_mistake_is_caught(orc_keys_1024, fault):
    name, check = CASES[fault]
    assert check(orc_keys_1024), f"check fails without a planted mistake: {name}"
    O.set_fault(fault)
    try:
        caught = not check(orc_keys_1024)
    finally:
        O.set_fault(0)
    assert caught, f"planted mistake not caught: {name}"
