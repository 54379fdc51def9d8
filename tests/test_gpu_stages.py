"""GPU parity of the stage kernels against the oracle, through the C ABI.

Bars: bit-exact for the integer stages (rotation, decomposition, extraction, keyswitch) and for
the FP64 external product (its FFT error stays far below 1/2, so rounding recovers the exact
int64 product); the split-FP32 external product is held to a stated torus error bound.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_1024
N, M = P.N, P.N // 2
FP = dict(n=P.n, N=P.N, k=P.k, l=P.l, Bgbit=P.Bgbit, ks_t=P.ks_t, ks_basebit=P.ks_basebit)


def T():
    import tfhe_amd
    return tfhe_amd


def closed_form_transform(p):
    """F_k = p(w^(4k+1)), w = e^{i pi/N}: direct evaluation at the roots of X^N + 1."""
    p = np.asarray(p, dtype=np.float64)
    n = len(p)
    j = np.arange(n)
    k = np.arange(n // 2)
    e = (j[None, :] * (4 * k + 1)[:, None]) % (2 * n)
    return (p[None, :] * np.exp(1j * np.pi * e / n)).sum(axis=1)


def rand_torus(rng, shape):
    return rng.integers(-2**31, 2**31, size=shape, dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("prec,rtol", [(0, 1e-12), (1, 4e-7)])
def test_fft_forward_matches_closed_form(prec, rtol, rng):
    polys = np.concatenate([rand_torus(rng, (3, N)), rng.integers(-64, 64, (3, N)).astype(np.int32),
                            np.eye(N, dtype=np.int32)[[0, 1, N // 2, N - 1]]])
    got = T().fft_forward(polys, prec=prec)
    for i in range(polys.shape[0]):
        ref = closed_form_transform(polys[i])
        err = np.abs(got[i] - ref).max()
        assert err <= rtol * max(1.0, np.abs(ref).max()) * 8, (i, err)


def test_fft_forward_digits_matrix_core_matches_closed_form(rng):
    """The split path's forward transform of gadget digits, first pass on the matrix cores (two FP16
    pieces of the twisted 8-point DFT matrix, FP32 accumulation), against direct evaluation at the
    roots of X^N + 1.  Inputs: random digits of both ring sizes' ranges, the extreme digits, unit
    impulses at the ends of each half, and the largest inputs the entry point accepts; the bound is
    the FP32 transform's (same as test_fft_forward_matches_closed_form at prec 1)."""
    polys = np.concatenate([rng.integers(-64, 64, (4, N)), rng.integers(-128, 128, (2, N)),
                            np.full((1, N), -64), np.full((1, N), 63), rng.choice([-2047, 2047], (1, N)),
                            np.eye(N, dtype=np.int64)[[0, 1, 63, 64, M - 1, M, M + 1, N - 1]]]).astype(np.int32)
    got = T().fft_forward_digits(polys)
    ref_fp32 = T().fft_forward(polys, prec=1)
    for i in range(polys.shape[0]):
        ref = closed_form_transform(polys[i])
        scale = max(1.0, np.abs(ref).max())
        err = np.abs(got[i] - ref).max()
        assert err <= 4e-7 * scale * 8, (i, err)
        # and no worse than the all-vector-ALU FP32 transform by more than a small factor
        assert err <= 4 * max(np.abs(ref_fp32[i] - ref).max(), 1e-7 * scale), (i, err)


def test_fft_forward_digits_rejects_out_of_range_inputs():
    polys = np.zeros((2, N), np.int32)
    polys[1, 5] = 2048
    with pytest.raises(T().TfheError, match="inputs must lie"):
        T().fft_forward_digits(polys)


@pytest.mark.parametrize("prec,tol", [(0, 1e-3), (1, None)])
def test_fft_roundtrip(prec, tol, rng):
    polys = rng.integers(-64, 64, (4, N)).astype(np.int32)
    F = T().fft_forward(polys, prec=prec)
    back = T().fft_inverse(F, prec=prec)
    err = np.abs(back - polys).max()
    assert err < (tol if tol is not None else 1e-3 * 64), err


def test_rotate_decompose_bit_exact(rng):
    acc = rand_torus(rng, (12, 2, N))
    bara = np.array([0, 1, 2, N - 1, N, N + 1, 2 * N - 1, 7, 1000, 1500, 2000, 513], np.int32)
    rot, dig = T().rotate_decompose(FP, acc, bara)
    for g in range(len(bara)):
        for q in range(2):
            want = O.poly_mul_xai_minus_one(int(bara[g]), acc[g, q])
            assert np.array_equal(rot[g, q], want), (g, q)
            wd = O.decompose(want, P.l, P.Bgbit)
            assert np.array_equal(dig[g, q * P.l:(q + 1) * P.l], wd), (g, q)


def test_rotate_decompose_every_rotation(rng):
    """Every rotation amount a in [0, 2N) on one accumulator: the blind-rotation loop's byte-offset
    rotation and signed-extract digits against the oracle (the kernel also cross-checks them
    against the plain forms and poisons any disagreement)."""
    acc1 = rand_torus(rng, (1, 2, N))
    acc = np.repeat(acc1, 2 * N, axis=0)
    bara = np.arange(2 * N, dtype=np.int32)
    rot, dig = T().rotate_decompose(FP, acc, bara)
    for a in range(2 * N):
        for q in range(2):
            want = O.poly_mul_xai_minus_one(a, acc1[0, q])
            assert np.array_equal(rot[a, q], want), (a, q)
            if a % 64 == 0:
                wd = O.decompose(want, P.l, P.Bgbit)
                assert np.array_equal(dig[a, q * P.l:(q + 1) * P.l], wd), (a, q)


def test_extract_bit_exact(rng):
    acc = rand_torus(rng, (5, 2, N))
    u = T().extract(FP, acc)
    for g in range(acc.shape[0]):
        assert np.array_equal(u[g], O.extract(P, acc[g]))


def test_extern_product_fp64_exact(gpu_env, rng):
    _, dk, keys = gpu_env
    acc = rand_torus(rng, (6, 2, N))
    rows = np.array([0, 1, 5, 100, 628, 629], np.int32)
    got, raw = dk.extern_product(acc, rows, prec=0, want_raw=True)
    worst = 0.0
    for g in range(acc.shape[0]):
        want, exact = O.extern_product_exact(P, acc[g], keys.bk[rows[g]], want_raw=True)
        assert np.array_equal(got[g], want), g
        worst = max(worst, float(np.abs(raw[g] - exact.astype(np.float64)).max()))
    print(f"FP64 external product: max |pre-rounding - exact int64| = {worst:.3e} (integer units)")
    assert worst < 0.25


def test_extern_product_split32_error(gpu_env, rng):
    _, dk, keys = gpu_env
    acc = rand_torus(rng, (6, 2, N))
    rows = np.array([0, 3, 17, 300, 500, 629], np.int32)
    got = dk.extern_product(acc, rows, prec=1)
    errs = []
    for g in range(acc.shape[0]):
        want = O.extern_product_exact(P, acc[g], keys.bk[rows[g]])
        d = (got[g].astype(np.int64) - want.astype(np.int64) + 2**31) % 2**32 - 2**31
        errs.append(np.abs(d))
    e = np.concatenate([x.ravel() for x in errs])
    print(f"split-FP32 external product torus error: max {e.max()} rms {np.sqrt((e.astype(float)**2).mean()):.1f}"
          f" (units of 2^-32)")
    # stated bound: 2^16 units = 2^-16 of the torus per external product
    assert e.max() < 2**16


@pytest.mark.parametrize("cnt", [1, 40, 600])
def test_keyswitch_bit_exact(gpu_env, rng, cnt):
    """Up to 512 ciphertexts the sliced kernel (input indices over workgroups, atomic sums) runs,
    above it the tiled kernel; both bit-exact with the oracle (40 samples checked, ends included)."""
    _, dk, keys = gpu_env
    u = rand_torus(rng, (cnt, N + 1))
    got = dk.keyswitch(u)
    idx = sorted({0, cnt - 1} | set(rng.choice(cnt, min(cnt, 38), replace=False).tolist()))
    for g in idx:
        assert np.array_equal(got[g], O.keyswitch(keys, u[g])), g


def test_blind_rotate_fp64_matches_oracle(gpu_env, rng):
    _, dk, keys = gpu_env
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in (0, 5, 1777)])
    bara = rng.integers(0, 2 * N, (3, P.n)).astype(np.int32)
    bara[1, ::3] = 0
    got = dk.blind_rotate(acc, bara, prec=0)
    for g in range(acc.shape[0]):
        want = O.blind_rotate(keys, acc[g], bara[g])
        assert np.array_equal(got[g], want), g


def test_blind_rotate_fp64_matches_exact_int64(gpu_env, rng):
    _, dk, keys = gpu_env
    acc = O.test_vector(P, 321, T().MU)[None]
    bara = rng.integers(0, 2 * N, (1, P.n)).astype(np.int32)
    got = dk.blind_rotate(acc, bara, prec=0)
    want = O.blind_rotate(keys, acc[0], bara[0], exact=True)
    assert np.array_equal(got[0], want)


@pytest.mark.parametrize("cnt", [2, 300])
def test_blind_rotate_split32_phase_error(gpu_env, rng, cnt):
    """After an inexact step the mask words diverge (different digits select different key-row
    combinations), so both paths are measured by the TLWE phase b - a.s against the noiseless
    rotation X^(sum bara_i s_i) * testvector, which is what decryption sees.  Two accumulators run
    the latency forms; 300 (more than one per CU) run the batched kernels, of which 24 spread over
    the launch (both ends included) are checked."""
    _, dk, keys = gpu_env
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in [99, 1234] + list(rng.integers(0, 2 * N, cnt - 2))])
    bara = rng.integers(0, 2 * N, (cnt, P.n)).astype(np.int32)
    res = {p: dk.blind_rotate(acc, bara, prec=p) for p in (0, 1)}
    rms = {}
    worst = {}
    checked = sorted({0, 1, cnt - 1} | set(np.linspace(0, cnt - 1, 24).astype(int).tolist()))
    for p in (0, 1):
        errs = []
        for g in checked:
            rot = int((bara[g].astype(np.int64) * keys.lwe_key).sum() % (2 * N))
            ideal = O.poly_mul_xai(rot, acc[g, 1])
            ph = O.tlwe_phase(P, res[p][g], keys.tlwe_key)
            errs.append((ph.astype(np.int64) - ideal.astype(np.int64) + 2**31) % 2**32 - 2**31)
        e = np.abs(np.concatenate(errs)).astype(np.float64)
        rms[p], worst[p] = np.sqrt((e ** 2).mean()), e.max()
    print(f"blind rotation phase noise over {len(checked)} accumulators (units of 2^-32): FP64 rms {rms[0]:.3e} "
          f"max {worst[0]:.3e}; split-FP32 rms {rms[1]:.3e} max {worst[1]:.3e}")
    # stated bounds: split-FP32 adds at most as much noise again as the key's own noise, and no
    # coefficient strays beyond 2^-6 of the torus (a quarter of the 1/16 gate decision margin)
    assert rms[1] <= 2.0 * rms[0] + 2.0**18
    assert worst[1] < 2**26


def test_bootstrap_woks_fp64_bit_exact(gpu_env, rng):
    ks, dk, keys = gpu_env
    bits = rng.integers(0, 2, (2, 5)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    x = np.stack([O.gate_combine(P, "NAND", ca[i], cb[i]) for i in range(5)])
    u, acc = dk.bootstrap_woks(x, prec=0, want_acc=True)
    for g in range(5):
        assert np.array_equal(u[g], O.bootstrap_woks(keys, x[g], T().MU)), g
        assert np.array_equal(u[g], O.extract(P, acc[g])), g


def hi_half(bk, bits):
    """The hi halves the device splits each key coefficient into: K = hi * 2^bits + lo with
    lo in [-2^(bits-1), 2^(bits-1))."""
    return ((bk.astype(np.int64) + (1 << (bits - 1))) >> bits).astype(np.int32)


def test_split_hi_rounding_margin(gpu_env):
    """The split path is exact only while the product of the digits with the hi key halves rounds to
    its exact integer value (a slip costs 2^24 units at once).  Over 131,072 coefficients (64
    external products) the FP32 value before rounding stays within 0.25 of an integer; on 16 of
    them it is compared with the exact int64 product of the same digits and hi halves (the oracle's
    exact external product with the hi-half key); the rounded product is the one the plain split
    entry point returns; and the default split point at N=1024 is 24 bits."""
    _, dk, keys = gpu_env
    assert dk.split_bits == 24
    rng = np.random.default_rng(424)
    cnt = 64
    acc = rand_torus(rng, (cnt, 2, N))
    rows = rng.integers(0, P.n, cnt).astype(np.int32)
    got, hi = dk.extern_product_split_hi(acc, rows)
    assert np.array_equal(got, dk.extern_product(acc, rows, prec=1))
    dist = np.abs(hi - np.rint(hi))
    worst_exact = 0.0
    for g in range(16):
        _, exact = O.extern_product_exact(P, acc[g], hi_half(keys.bk[rows[g]], 24), want_raw=True)
        worst_exact = max(worst_exact, float(np.abs(hi[g] - exact.astype(np.float64)).max()))
    print(f"N=1024 split 24: hi product distance to an integer max {dist.max():.4f} rms {np.sqrt((dist**2).mean()):.4f}"
          f" over {dist.size} coefficients; max |hi - exact| = {worst_exact:.4f}; |hi| max {np.abs(hi).max():.3e}",
          flush=True)
    assert dist.max() < 0.25
    assert worst_exact < 0.25


@pytest.mark.parametrize("variant", [0, 1])
def test_split_digits_fp16_bit_exact(rng, variant):
    """The split kernels' conversion of gadget digits to FP16 pairs (byte permutes, an exponent
    trick and one packed FP16 add) against the oracle's signed decomposition, digit by digit:
    random coefficients and the ends of each digit's range (every 7-bit field at -64 and 63, the
    coefficients whose offset addition carries through every field).  variant 0 is the batched
    kernel's three-level form, 1 the latency form's one-level form."""
    x = np.concatenate([rand_torus(rng, 4096),
                        np.array([0, -1, 2**31 - 1, -2**31, 0x7E000000, -0x7E000000, 0x3F7F0000, -0x40810000],
                                 np.int64).astype(np.int32)])
    x = np.concatenate([x, x[::-1]]).reshape(-1, 2)
    got = T().split_digits(x, P.Bgbit, variant)
    for h in range(2):
        want = O.decompose(x[:, h], P.l, P.Bgbit)      # [l][count]
        assert np.array_equal(got[:, :, h].T, want.astype(np.float32)), h


def _extreme_accumulators(P_, n_rand, rng):
    """Accumulators for the exactness sweeps: random torus values, every coefficient with all three
    digits at -2^(Bg-1) or at 2^(Bg-1) - 1 (the largest digit magnitudes the products see), the two
    alternating, and the int32 ends."""
    Np = P_.N
    unit = sum(1 << (32 - (p + 1) * P_.Bgbit) for p in range(P_.l))
    lo = np.int64(-(1 << (P_.Bgbit - 1)) * unit)
    hi = np.int64(((1 << (P_.Bgbit - 1)) - 1) * unit)
    wrap = lambda v: ((np.asarray(v, np.int64) + 2**31) % 2**32 - 2**31).astype(np.int32)
    alt = np.where(np.arange(Np) % 2 == 0, lo, hi)
    special = [np.full((2, Np), lo), np.full((2, Np), hi), np.stack([alt, alt[::-1]]),
               np.full((2, Np), 2**31 - 1), np.full((2, Np), -2**31)]
    return np.concatenate([rand_torus(rng, (n_rand, 2, Np)), wrap(np.stack(special))])


def test_extern_product_fp64_exact_sweep(gpu_env):
    """FP64 external products bit-exact with the exact int64 product over 48 products: rows spread
    over the key, random and extreme-digit accumulators; the largest distance of the pre-rounding
    value from the exact integer is bounded well below 1/2."""
    _, dk, keys = gpu_env
    rng = np.random.default_rng(1048)
    acc = _extreme_accumulators(P, 43, rng)
    rows = np.linspace(0, P.n - 1, acc.shape[0]).astype(np.int32)
    got, raw = dk.extern_product(acc, rows, prec=0, want_raw=True)
    worst = 0.0
    for g in range(acc.shape[0]):
        want, exact = O.extern_product_exact(P, acc[g], keys.bk[rows[g]], want_raw=True)
        assert np.array_equal(got[g], want), (g, int(rows[g]))
        worst = max(worst, float(np.abs(raw[g] - exact.astype(np.float64)).max()))
    print(f"FP64 external product sweep ({acc.shape[0]} products): max |pre-rounding - exact| = {worst:.3e}")
    assert worst < 0.125
