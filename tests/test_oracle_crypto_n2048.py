"""CPU: the oracle's crypto pipeline at the N=2048 set (n=750, l=3, Bgbit=8, ks t=10) against the
identities that define it, so that the GPU parity tests at this set rest on a pinned oracle:
TGSW row phases, the FFT external product against the exact int64 one, the external product's
phase, the blind rotation's noiseless result, keyswitch phase, gate truth tables, and one full
bootstrap through both products."""
import numpy as np
import pytest

import oracle as O

P = O.PARAMS_2048
N = P.N
TWO32 = 2**32


def sdiff(a, b):
    return ((np.asarray(a, np.int64) - np.asarray(b, np.int64) + 2**31) % TWO32) - 2**31


def wrap(x):
    return ((np.asarray(x, np.int64) + 2**31) % TWO32) - 2**31


@pytest.fixture(scope="module")
def keys2048():
    k = O.keygen(P, (0,))
    yield k
    k.free()


def test_parameters_and_offset():
    """The set's closed forms: decomposition offset 128 (2^24 + 2^16 + 2^8) = 0x80808000, kpl = 6."""
    assert (P.n, P.N, P.l, P.Bgbit, P.ks_t, P.ks_basebit) == (750, 2048, 3, 8, 10, 2)
    assert P.kpl == 6
    offset = sum((1 << (P.Bgbit - 1)) << (32 - (p + 1) * P.Bgbit) for p in range(P.l))
    assert offset == 0x80808000
    x = np.array([0, 1, -1, 2**31 - 1, -2**31, 0x12345678], np.int64).astype(np.int32)
    d = O.decompose(x, P.l, P.Bgbit)
    assert d.min() >= -128 and d.max() <= 127
    # sum_p d_p 2^(32 - (p+1) Bgbit) = x - r (mod 2^32) with 0 <= r < 2^(32 - l Bgbit): the offset's
    # low 8 bits are zero, so r is x's low byte
    rec = sum(d[p].astype(np.int64) << (32 - (p + 1) * P.Bgbit) for p in range(P.l))
    r = (x.astype(np.int64) - rec) % TWO32
    assert np.array_equal(r, x.astype(np.int64) & 0xFF)


def test_bootstrapping_key_rows_decrypt(keys2048):
    K = keys2048
    s = K.tlwe_key
    for i in (0, 374, 749):
        for r in range(P.kpl):
            bloc, lvl = divmod(r, P.l)
            ph = O.tlwe_phase(P, K.bk[i, r], s)
            h = 1 << (32 - (lvl + 1) * P.Bgbit)
            want = np.zeros(N, np.int64)
            if bloc == 1:
                want[0] = int(K.lwe_key[i]) * h
            else:
                want = -int(K.lwe_key[i]) * h * s.astype(np.int64)
            e = sdiff(ph, want) / TWO32
            assert np.abs(e).max() < 7 * P.bk_alpha, (i, r)


def test_extern_product_fft_equals_exact(keys2048):
    K = keys2048
    rng = np.random.default_rng(20)
    for rs in (0, 374, 749):
        acc = rng.integers(-2**31, 2**31, (2, N), dtype=np.int64).astype(np.int32)
        ex, raw = O.extern_product_exact(P, acc, K.bk[rs], want_raw=True)
        ff, rawf = O.extern_product_fft(K, acc, rs, want_raw=True)
        assert np.array_equal(ex, ff)
        assert np.abs(rawf - raw.astype(np.float64)).max() < 0.25


def test_extern_product_phase_identity(keys2048):
    """bk_i [.] c has phase s_i * phase(c) up to the key noise and the decomposition error."""
    K = keys2048
    rng = np.random.default_rng(21)
    for i in (1, 2, 3, 4):
        msg = (rng.integers(0, 8, N) * (1 << 29)).astype(np.int64)
        a = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
        b = np.asarray(wrap(msg + O.negacyclic_mul_naive(K.tlwe_key, a).astype(np.int64)), np.int32)
        out = O.extern_product_exact(P, np.stack([a, b]), K.bk[i])
        e = sdiff(O.tlwe_phase(P, out, K.tlwe_key), msg * int(K.lwe_key[i])) / TWO32
        # key noise: 6 rows x 2048 coefficients x digits of rms 74 x stdev 2^-30 -> stdev 2^-17,
        # plus the decomposition rounding (< 2^-25 per coefficient, weighted by the 0/1 key)
        assert np.abs(e).max() < 2.0 ** -12, i


def test_blind_rotate_noiseless_result(keys2048):
    K = keys2048
    rng = np.random.default_rng(22)
    mu = O.modswitch_to(1, 8)
    acc0 = O.test_vector(P, 1234, mu)
    bara = rng.integers(0, 2 * N, P.n).astype(np.int32)
    acc = O.blind_rotate(K, acc0, bara)
    rot = int((bara.astype(np.int64) * K.lwe_key).sum() % (2 * N))
    ideal = O.poly_mul_xai(rot, acc0[1])
    e = sdiff(O.tlwe_phase(P, acc, K.tlwe_key), ideal) / TWO32
    assert np.abs(e).max() < 2.0 ** -8
    assert np.sqrt((e ** 2).mean()) < 2.0 ** -10


def test_keyswitch_preserves_phase(keys2048):
    K = keys2048
    rng = np.random.default_rng(23)
    for _ in range(6):
        m = int(rng.integers(0, 8)) << 29
        a = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
        b = int(wrap(m + int((a.astype(np.int64) * K.tlwe_key).sum())))
        out = O.keyswitch(K, np.append(a, np.int32(b)))
        e = sdiff(O.lwe_phase(out[None], K.lwe_key), m)[0] / TWO32
        # about 3tN/4 selected key samples of stdev 2^-17 (stdev 2^-10) plus the 2^-(2t+1)
        # rounding of each input coefficient; bound = 8 stdev
        assert abs(e) < 2.0 ** -7


@pytest.mark.parametrize("gate", ["NAND", "MUX"])
def test_gate_truth_tables(keys2048, gate):
    K = keys2048
    combos = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in ((0, 1) if gate == "MUX" else (0,))]
    for a, b, c in combos:
        ca, cb, cc = (O.boots_encrypt(P, [v], K.lwe_key)[0] for v in (a, b, c))
        out = O.gate(K, gate, ca, cb, cc)
        assert O.boots_decrypt(out[None], K.lwe_key)[0] == O.TRUTH[gate](a, b, c), (a, b, c)


def test_exact_and_fft_bootstrap_agree(keys2048):
    """One full bootstrap through the exact int64 product and through the FP64 transform."""
    K = keys2048
    ca, cb = O.boots_encrypt(P, [1, 1], K.lwe_key)
    x = O.gate_combine(P, "NAND", ca, cb)
    mu = O.modswitch_to(1, 8)
    assert np.array_equal(O.bootstrap_woks(K, x, mu), O.bootstrap_woks(K, x, mu, exact=True))
