"""CPU: bootstrapping keys in the upstream FFT-domain order (LagrangeHalfC values, torus units).

The convention is pinned to the reference's own conversions, restated in the oracle by direct
summation: int -> values (tGswFFTExternMulToTLwe.cpp:97-130) and values -> torus
(TorusPolynomial_fft.cpp:4-43, truncating).  The library's conversion must agree with the first
and recover the key exactly where the second truncates."""
import numpy as np
import pytest

import oracle as O
import tfhe_amd as T

TWO32 = 2.0**32


def flat(N, n=1, l=1):
    return dict(n=n, N=N, k=1, l=l, Bgbit=7, ks_t=8, ks_basebit=2)


def ref_values(t):
    """upstream values of a torus polynomial: the reference's int convention, in torus units"""
    return O.lagrange_half_reference(t) / TWO32


def wrap_diff(a, b):
    return ((a.astype(np.int64) - b.astype(np.int64) + 2**31) % 2**32) - 2**31


@pytest.mark.parametrize("N", [16, 1024])
def test_torus_unit_scale_pinned_by_reference(N):
    """values of t in torus units, sent back through TorusPolynomial_fft, give t (up to its
    truncation): a missing 1/2, 1/N or 2^32 factor, or a conjugated set, would fail"""
    rng = np.random.default_rng(N)
    for _ in range(3):
        t = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
        back = O.torus_from_lagrange_reference(ref_values(t))
        assert np.abs(wrap_diff(back, t)).max() <= 1
    # a conjugated value set does not map back
    t = rng.integers(-2**31, 2**31, N, dtype=np.int64).astype(np.int32)
    wrong = O.torus_from_lagrange_reference(np.conj(ref_values(t)))
    assert np.abs(wrap_diff(wrong, t)).max() > 1000


@pytest.mark.parametrize("N", [16, 1024])
def test_library_values_match_reference_convention(N):
    rng = np.random.default_rng(N + 1)
    fp = flat(N)
    bk = rng.integers(-2**31, 2**31, (1, 2, 2, N), dtype=np.int64).astype(np.int32)
    got = T.bk_to_lagrange(fp, bk)
    assert got.shape == (1, 2, 2, N // 2)
    for r in range(2):
        for q in range(2):
            want = ref_values(bk[0, r, q])
            assert np.abs(got[0, r, q] - want).max() <= 1e-12 * N


@pytest.mark.parametrize("N", [16, 1024, 2048])
def test_from_reference_values_is_exact(N):
    rng = np.random.default_rng(N + 2)
    fp = flat(N)
    bk = rng.integers(-2**31, 2**31, (1, 2, 2, N), dtype=np.int64).astype(np.int32)
    bk[0, 0, 0, :4] = [-2**31, 2**31 - 1, 0, -1]          # wrap edges
    lag = np.stack([ref_values(bk[0, r, q]) for r in range(2) for q in range(2)]).reshape(1, 2, 2, N // 2)
    assert np.array_equal(T.bk_from_lagrange(fp, lag), bk)


def test_full_key_round_trip_n1024():
    ks = T.KeySet(1024, seed=(5,))
    try:
        lag = T.bk_to_lagrange(ks.fp, ks.bk)
        assert lag.shape == ks.bk.shape[:-1] + (512,)
        assert np.array_equal(T.bk_from_lagrange(ks.fp, lag), ks.bk)
    finally:
        ks.free()


def test_not_a_torus_transform_is_refused():
    N = 1024
    rng = np.random.default_rng(9)
    fp = flat(N)
    bk = rng.integers(-2**31, 2**31, (1, 2, 2, N), dtype=np.int64).astype(np.int32)
    lag = T.bk_to_lagrange(fp, bk)
    # half an integer unit on coefficient 0 of polynomial 3: delta * (2/N) * 2^32 = 1/2
    bad = lag.copy()
    bad[0, 1, 1, 0] += 0.5 * N / 2 / TWO32
    with pytest.raises(T.TfheError, match="polynomial 3"):
        T.bk_from_lagrange(fp, bad)
    bad = lag.copy()
    bad[0, 0, 1, 7] = np.nan
    with pytest.raises(T.TfheError, match="polynomial 1"):
        T.bk_from_lagrange(fp, bad)
    # a tenth of a unit is still read as the exact key
    ok = lag.copy()
    ok[0, 1, 1, 0] += 0.1 * N / 2 / TWO32
    assert np.array_equal(T.bk_from_lagrange(fp, ok), bk)


def test_bad_arguments():
    with pytest.raises(T.TfheError):
        T.bk_from_lagrange(flat(1000), np.zeros((1, 2, 2, 500), np.complex128))   # N not a power of two
