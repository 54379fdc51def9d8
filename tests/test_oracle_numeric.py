"""CPU: the oracle's torus numerics against the closed-form golden values (tests/golden) and brute
force, and the decomposition against its reconstruction invariant."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def test_modswitch_from_golden():
    for x, m, want in GOLD["modswitch_from"]:
        assert O.modswitch_from(x, m) == want, (x, m)


def test_modswitch_to_golden():
    for mu, m, want in GOLD["modswitch_to"]:
        assert O.modswitch_to(mu, m) == want, (mu, m)


def test_approx_phase_golden():
    for x, m, want in GOLD["approx_phase"]:
        assert O.approx_phase(x, m) == want, (x, m)


def test_dtot32_golden():
    for d, want in GOLD["dtot32"]:
        assert O.dtot32(d) == want, d


def test_constants_pinned_by_reference():
    assert O.modswitch_to(1, 8) == GOLD["mu"] == GOLD["pinned_by_reference"]["mu"]
    assert O.decomp_offset(3, 7) == GOLD["pinned_by_reference"]["offset_l3_bg7"] == 0x81020000
    for l, bg, want in GOLD["decomp_offset"]:
        assert O.decomp_offset(l, bg) == want
    assert O.decomp_offset(3, 8) == 0x80808000


def test_modswitch_brute_force_every_bucket_edge():
    """Every multiple of 2^21 (2N = 2048) and its neighbours: the bucket edges of the rounding."""
    for k in range(0, 2048, 7):
        base = (k << 21) - (1 << 20)  # exact half point between buckets k-1 and k
        for delta, want in ((-1, (k - 1) % 2048), (0, k % 2048), (1, k % 2048)):
            x = ((base + delta + 2**31) % 2**32) - 2**31
            assert O.modswitch_from(x, 2048) == want, (k, delta)


def test_modswitch_round_trip():
    rng = np.random.default_rng(0)
    for m in (2048, 4096):
        for k in rng.integers(0, m, 64):
            x = O.modswitch_to(int(k), m)
            assert O.modswitch_from(x, m) == k
            assert O.approx_phase(x, m) == x


def test_gaussian_statistics():
    O.set_seed([7])
    sigma = 2.0 ** -15
    xs = np.array([O.gaussian32(0, sigma) for _ in range(20000)], dtype=np.float64) / 2**32
    assert abs(xs.mean()) < 4 * sigma / np.sqrt(len(xs))
    assert abs(xs.std() / sigma - 1) < 0.03


@pytest.mark.parametrize("l,bg", [(3, 7), (3, 8), (2, 10)])
def test_decomposition_reconstruction(l, bg):
    """sum_p d_p 2^(32-(p+1)Bg) = x - r (mod 2^32) with 0 <= r < 2^(32 - l Bg) (the offset carries no
    rounding term, so the low bits are truncated) and d_p in [-Bg/2, Bg/2)."""
    rng = np.random.default_rng(1)
    x = rng.integers(-2**31, 2**31, 4096, dtype=np.int64).astype(np.int32)
    x[:6] = [0, -1, 2**31 - 1, -2**31, 1 << (32 - l * bg - 1), -(1 << (32 - l * bg - 1))]
    d = O.decompose(x, l, bg).astype(np.int64)
    half = 1 << (bg - 1)
    assert d.min() >= -half and d.max() < half
    rec = sum(d[p] * (1 << (32 - (p + 1) * bg)) for p in range(l))
    r = (x.astype(np.int64) - rec) % 2**32
    assert r.min() >= 0 and r.max() < 1 << (32 - l * bg)
    # the remainder is exactly the truncated low part of x
    assert np.array_equal(r, x.astype(np.int64) % (1 << (32 - l * bg)))


def test_decomposition_brute_force_small():
    """Brute force over every value of the low bits for a fixed high part (l=3, Bg=2^7)."""
    vals = np.arange(-(1 << 12), 1 << 12, dtype=np.int64) * 1021 + 12345
    x = (((vals + 2**31) % 2**32) - 2**31).astype(np.int32)
    d = O.decompose(x, 3, 7).astype(np.int64)
    rec = d[0] * 2**25 + d[1] * 2**18 + d[2] * 2**11
    r = (x.astype(np.int64) - rec) % 2**32
    assert np.array_equal(r, x.astype(np.int64) % 2**11)
