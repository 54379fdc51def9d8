"""CPU: the oracle's crypto pipeline against the identities that define it -- TLWE/TGSW
decryption, the external product's phase, the blind rotation's noiseless result, extraction,
keyswitch, gate truth tables, and the reference's known answer OR(499, 388) = 503."""
import numpy as np
import pytest

import oracle as O

P = O.PARAMS_1024
N = P.N
TWO32 = 2**32


def sdiff(a, b):
    return ((np.asarray(a, np.int64) - np.asarray(b, np.int64) + 2**31) % TWO32) - 2**31


def test_lwe_encrypt_decrypt(orc_keys_1024):
    K = orc_keys_1024
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2, 200)
    ct = O.boots_encrypt(P, bits, K.lwe_key)
    assert np.array_equal(O.boots_decrypt(ct, K.lwe_key), bits)
    mu = O.modswitch_to(1, 8)
    err = sdiff(O.lwe_phase(ct, K.lwe_key), np.where(bits == 1, mu, -mu)) / TWO32
    assert np.abs(err).max() < 6 * P.lwe_alpha


def test_bootstrapping_key_rows_decrypt(orc_keys_1024):
    """Each TGSW row r = bloc*l + lvl of bk_i has phase s_i * h_lvl on polynomial bloc (message
    times the gadget, upstream tGswAddMuIntH) plus noise of stdev bk_alpha."""
    K = orc_keys_1024
    s = K.tlwe_key
    for i in (0, 17, 629):
        for r in range(P.kpl):
            bloc, lvl = divmod(r, P.l)
            row = K.bk[i, r]
            ph = O.tlwe_phase(P, row, s)
            want = np.zeros(N, np.int64)
            h = 1 << (32 - (lvl + 1) * P.Bgbit)
            if bloc == 1:
                want[0] = int(K.lwe_key[i]) * h
            else:
                # message on the mask polynomial: b - a.s carries -s * (mu h) = -mu h * s(X)
                want = -int(K.lwe_key[i]) * h * s.astype(np.int64)
            e = sdiff(ph, want) / TWO32
            assert np.abs(e).max() < 7 * P.bk_alpha, (i, r)


def test_extern_product_fft_equals_exact(orc_keys_1024):
    K = orc_keys_1024
    rng = np.random.default_rng(2)
    for rs in (0, 100, 629):
        acc = rng.integers(-2**31, 2**31, (2, N), dtype=np.int64).astype(np.int32)
        ex, raw = O.extern_product_exact(P, acc, K.bk[rs], want_raw=True)
        ff, rawf = O.extern_product_fft(K, acc, rs, want_raw=True)
        assert np.array_equal(ex, ff)
        assert np.abs(rawf - raw.astype(np.float64)).max() < 0.25


def test_extern_product_phase_identity(orc_keys_1024):
    """bk_i [.] c has phase s_i * phase(c) up to the key noise and the decomposition error."""
    K = orc_keys_1024
    rng = np.random.default_rng(3)
    for i in (1, 2, 3, 4):
        msg = (rng.integers(0, 8, N) * (1 << 29)).astype(np.int64)
        a = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
        b = np.asarray(wrap(msg + O.negacyclic_mul_naive(K.tlwe_key, a).astype(np.int64)), np.int32)
        c = np.stack([a, b])
        out = O.extern_product_exact(P, c, K.bk[i])
        ph = O.tlwe_phase(P, out, K.tlwe_key)
        want = msg * int(K.lwe_key[i])
        e = sdiff(ph, want) / TWO32
        # key noise: 6 rows x N coefficients x digits of rms 37 x stdev 2^-25 -> stdev 2^-13.5,
        # plus the decomposition rounding (< 2^-22 per coefficient); bound = 11 stdev
        assert np.abs(e).max() < 2.0 ** -10, i


def wrap(x):
    return ((np.asarray(x, np.int64) + 2**31) % TWO32) - 2**31


def test_blind_rotate_noiseless_result(orc_keys_1024):
    K = orc_keys_1024
    rng = np.random.default_rng(4)
    mu = O.modswitch_to(1, 8)
    acc0 = O.test_vector(P, 321, mu)
    bara = rng.integers(0, 2 * N, P.n).astype(np.int32)
    acc = O.blind_rotate(K, acc0, bara)
    rot = int((bara.astype(np.int64) * K.lwe_key).sum() % (2 * N))
    ideal = O.poly_mul_xai(rot, acc0[1])
    e = sdiff(O.tlwe_phase(P, acc, K.tlwe_key), ideal) / TWO32
    assert np.abs(e).max() < 2.0 ** -5
    assert np.sqrt((e ** 2).mean()) < 2.0 ** -7


def test_extract_phase_identity(orc_keys_1024):
    """The extracted LWE sample's phase is coefficient 0 of the TLWE phase."""
    K = orc_keys_1024
    rng = np.random.default_rng(5)
    acc = rng.integers(-2**31, 2**31, (2, N), dtype=np.int64).astype(np.int32)
    u = O.extract(P, acc)
    ph_lwe = O.lwe_phase(u[None], K.tlwe_key)[0]
    ph_tlwe = O.tlwe_phase(P, acc, K.tlwe_key)[0]
    assert ph_lwe == ph_tlwe


def test_keyswitch_preserves_phase(orc_keys_1024):
    K = orc_keys_1024
    rng = np.random.default_rng(6)
    for _ in range(8):
        m = int(rng.integers(0, 8)) << 29
        a = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
        b = int(wrap(m + int((a.astype(np.int64) * K.tlwe_key).sum())))
        u = np.append(a, np.int32(b))
        out = O.keyswitch(K, u)
        e = sdiff(O.lwe_phase(out[None], K.lwe_key), m)[0] / TWO32
        # about 3tN/4 selected key samples of stdev 2^-15 (stdev 2^-8.7) plus the 2^-(2t+1)
        # rounding of each input coefficient; bound = 8 stdev
        assert abs(e) < 2.0 ** -5.7


@pytest.mark.parametrize("gate", list(O.GATES))
def test_gate_truth_tables(orc_keys_1024, gate):
    K = orc_keys_1024
    combos = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in ((0, 1) if gate == "MUX" else (0,))]
    for a, b, c in combos:
        ca, cb, cc = (O.boots_encrypt(P, [v], K.lwe_key)[0] for v in (a, b, c))
        out = O.gate(K, gate, ca, cb, cc)
        assert O.boots_decrypt(out[None], K.lwe_key)[0] == O.TRUTH[gate](a, b, c), (a, b, c)


def test_known_answer_or_499_388(orc_keys_1024):
    K = orc_keys_1024
    a = [(499 >> i) & 1 for i in range(16)]
    b = [(388 >> i) & 1 for i in range(16)]
    ca, cb = O.boots_encrypt(P, a, K.lwe_key), O.boots_encrypt(P, b, K.lwe_key)
    out, _ = O.gate_batch(K, "OR", ca, cb)
    bits = O.boots_decrypt(out, K.lwe_key)
    assert sum(int(v) << i for i, v in enumerate(bits)) == 503


def test_exact_and_fft_bootstrap_agree(orc_keys_1024):
    """One full bootstrap through the exact int64 product and through the FP64 transform."""
    K = orc_keys_1024
    ca, cb = O.boots_encrypt(P, [1, 0], K.lwe_key)
    x = O.gate_combine(P, "NAND", ca, cb)
    assert np.array_equal(O.bootstrap_woks(K, x, O.modswitch_to(1, 8)),
                          O.bootstrap_woks(K, x, O.modswitch_to(1, 8), exact=True))
