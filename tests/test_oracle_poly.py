"""CPU: the oracle's polynomial arithmetic against independent restatements -- negacyclic products
by numpy convolution and folding, rotations as products with the monomial X^a, the transform
against direct evaluation at the roots of X^N + 1, and the reference's 2N-point convention."""
import numpy as np
import pytest

import oracle as O


def negacyclic_numpy(a, b):
    """(a * b) mod (X^N + 1) over the integers, via a full convolution and one fold."""
    a = np.asarray(a, dtype=object)
    b = np.asarray(b, dtype=object)
    n = len(a)
    full = np.convolve(a, b)
    res = full[:n].copy()
    res[: n - 1] -= full[n:]
    return res


def monomial(a, n):
    """X^a mod (X^N + 1) for a in [0, 2N) as an integer vector."""
    m = np.zeros(n, dtype=np.int64)
    m[a % n] = 1 if a < n else -1
    return m


def wrap32(x):
    return np.array([((int(v) + 2**31) % 2**32) - 2**31 for v in x], dtype=np.int32)


@pytest.mark.parametrize("n", [16, 64])
def test_rotation_all_exponents(n):
    rng = np.random.default_rng(n)
    src = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    for a in range(2 * n):
        want = wrap32(negacyclic_numpy(monomial(a, n), src))
        assert np.array_equal(O.poly_mul_xai(a, src), want), a
        want1 = wrap32(np.asarray(want, dtype=np.int64) - src.astype(np.int64))
        assert np.array_equal(O.poly_mul_xai_minus_one(a, src), want1), a


def test_rotation_n1024_samples():
    rng = np.random.default_rng(5)
    n = 1024
    src = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    for a in (0, 1, 511, 1023, 1024, 1025, 2047, 1500):
        want = wrap32(negacyclic_numpy(monomial(a, n), src) - src.astype(object))
        assert np.array_equal(O.poly_mul_xai_minus_one(a, src), want), a


@pytest.mark.parametrize("n", [16, 1024])
def test_negacyclic_int64_product(n):
    rng = np.random.default_rng(n + 1)
    d = rng.integers(-64, 64, n).astype(np.int32)
    t = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    want = negacyclic_numpy(d, t)
    got = O.negacyclic_mul_i64(d, t)
    assert all(int(g) == int(w) for g, w in zip(got, want))
    assert np.array_equal(O.negacyclic_mul_naive(d, t), wrap32(want))


def direct_eval(p):
    n = len(p)
    j = np.arange(n)
    k = np.arange(n // 2)
    e = (j[None, :] * (4 * k + 1)[:, None]) % (2 * n)
    return (np.asarray(p, np.float64)[None, :] * np.exp(1j * np.pi * e / n)).sum(axis=1)


@pytest.mark.parametrize("n", [16, 256, 1024, 2048])
def test_transform_is_evaluation_at_roots(n):
    rng = np.random.default_rng(n + 2)
    for p in (rng.integers(-64, 64, n), rng.integers(-2**31, 2**31, n, dtype=np.int64), np.eye(n, dtype=np.int64)[3]):
        p = p.astype(np.int32)
        F = O.fft_forward(p)
        ref = direct_eval(p)
        assert np.abs(F - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * n


@pytest.mark.parametrize("n", [16, 1024])
def test_transform_product_and_inverse(n):
    rng = np.random.default_rng(n + 3)
    a = rng.integers(-64, 64, n).astype(np.int32)
    b = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    back = O.fft_inverse_raw(O.fft_forward(a), n)
    assert np.abs(back - a).max() < 1e-6
    prod = O.fft_inverse_raw(O.fft_forward(a) * O.fft_forward(b), n)
    exact = negacyclic_numpy(a, b)
    assert np.abs(prod - np.array(exact, dtype=np.float64)).max() < 0.25


@pytest.mark.parametrize("n", [16, 1024])
def test_reference_lagrange_half_convention(n):
    """The reference's transform (tGswFFTExternMulToTLwe.cpp:97-130) gives p(w^(2i+1)), i < N/2.
    The folded transform gives p(w^(4k+1)); exponent 2i+1 = 4k+1 maps directly, 2i+1 = 4k+3 is the
    conjugate of the folded value at 2N - (2i+1)."""
    rng = np.random.default_rng(n + 4)
    p = rng.integers(-2**20, 2**20, n).astype(np.int32)
    ref = O.lagrange_half_reference(p)
    F = O.fft_forward(p)
    for i in range(n // 2):
        e = 2 * i + 1
        if e % 4 == 1:
            mapped = F[(e - 1) // 4]
        else:
            mapped = np.conj(F[(2 * n - e - 1) // 4])
        assert abs(mapped - ref[i]) <= 1e-9 * max(1.0, abs(ref[i])) * n, i
