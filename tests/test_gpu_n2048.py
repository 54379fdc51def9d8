"""GPU parity at the N=2048 parameter set (n=750, l=3, Bgbit=8, ks t=10): the transform against its
closed form, the integer stages and the FP64 external product and gates bit-exact against the
oracle, split-FP32 within stated error bounds, and gate truth tables on both paths."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_2048
N = P.N
FP = dict(n=P.n, N=P.N, k=P.k, l=P.l, Bgbit=P.Bgbit, ks_t=P.ks_t, ks_basebit=P.ks_basebit)


def T():
    import tfhe_amd
    return tfhe_amd


def rand_torus(rng, shape):
    return rng.integers(-2**31, 2**31, size=shape, dtype=np.int64).astype(np.int32)


@pytest.fixture(scope="module")
def env2048():
    T().init(0)
    print("[n2048] keygen + device key", flush=True)
    ks = T().KeySet(2048, seed=(0,))
    dk = T().DeviceKey.from_keyset(ks)
    keys = O.Keys(P, ks.lwe_key, ks.tlwe_key, ks.ksk, ks.bk)
    yield ks, dk, keys
    dk.free()
    keys.free()
    ks.free()


@pytest.mark.parametrize("prec,rtol", [(0, 1e-12), (1, 4e-7)])
def test_fft_forward_n2048_matches_closed_form(prec, rtol):
    rng = np.random.default_rng(2048)
    polys = np.concatenate([rand_torus(rng, (2, N)), rng.integers(-128, 128, (2, N)).astype(np.int32),
                            np.eye(N, dtype=np.int32)[[0, 1, N // 2, N - 1]]])
    got = T().fft_forward(polys, prec=prec)
    j = np.arange(N)
    for i in range(polys.shape[0]):
        k = np.arange(N // 2)
        e = (j[None, :] * (4 * k + 1)[:, None]) % (2 * N)
        ref = (polys[i].astype(np.float64)[None, :] * np.exp(1j * np.pi * e / N)).sum(axis=1)
        err = np.abs(got[i] - ref).max()
        assert err <= rtol * max(1.0, np.abs(ref).max()) * 8, (i, err)


def test_rotate_decompose_n2048_bit_exact():
    rng = np.random.default_rng(2049)
    acc = rand_torus(rng, (8, 2, N))
    bara = np.array([0, 1, N - 1, N, N + 1, 2 * N - 1, 3000, 1234], np.int32)
    rot, dig = T().rotate_decompose(FP, acc, bara)
    for g in range(len(bara)):
        for q in range(2):
            want = O.poly_mul_xai_minus_one(int(bara[g]), acc[g, q])
            assert np.array_equal(rot[g, q], want), (g, q)
            assert np.array_equal(dig[g, q * P.l:(q + 1) * P.l], O.decompose(want, P.l, P.Bgbit)), (g, q)


def test_extern_product_n2048_fp64_exact(env2048):
    _, dk, keys = env2048
    rng = np.random.default_rng(2050)
    acc = rand_torus(rng, (3, 2, N))
    rows = np.array([0, 374, 749], np.int32)
    got, raw = dk.extern_product(acc, rows, prec=0, want_raw=True)
    worst = 0.0
    for g in range(acc.shape[0]):
        want, exact = O.extern_product_exact(P, acc[g], keys.bk[rows[g]], want_raw=True)
        assert np.array_equal(got[g], want), g
        worst = max(worst, float(np.abs(raw[g] - exact.astype(np.float64)).max()))
    print(f"N=2048 FP64 external product: max |pre-rounding - exact int64| = {worst:.3e}")
    assert worst < 0.25


def test_extern_product_n2048_split32_error(env2048):
    _, dk, keys = env2048
    rng = np.random.default_rng(2051)
    acc = rand_torus(rng, (3, 2, N))
    rows = np.array([1, 400, 748], np.int32)
    got = dk.extern_product(acc, rows, prec=1)
    e = np.concatenate([np.abs((got[g].astype(np.int64) - O.extern_product_exact(P, acc[g], keys.bk[rows[g]])
                                .astype(np.int64) + 2**31) % 2**32 - 2**31).ravel() for g in range(3)])
    print(f"N=2048 split-FP32 external product torus error: max {e.max()} "
          f"rms {np.sqrt((e.astype(float) ** 2).mean()):.1f} (units of 2^-32)")
    # stated bound: 2^18 units = 2^-14 of the torus per external product
    assert e.max() < 2**18


@pytest.mark.parametrize("cnt", [20, 600])
def test_keyswitch_n2048_bit_exact(env2048, cnt):
    """The sliced kernel (up to 512 ciphertexts) and the tiled one, bit-exact with the oracle on
    20 samples, ends included."""
    _, dk, keys = env2048
    rng = np.random.default_rng(2052)
    u = rand_torus(rng, (cnt, N + 1))
    got = dk.keyswitch(u)
    idx = sorted({0, cnt - 1} | set(rng.choice(cnt, min(cnt, 18), replace=False).tolist()))
    for g in idx:
        assert np.array_equal(got[g], O.keyswitch(keys, u[g])), g


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("gate", ["NAND", "XOR", "MUX", "AND"])
def test_gate_n2048_truth_table(env2048, gate, prec):
    ks, dk, _ = env2048
    bits = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)] * 3, np.int32).T
    print(f"[n2048] {gate} prec={prec} count={bits.shape[1]}", flush=True)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = dk.gate(gate, ca, cb, cc if gate == "MUX" else None, prec=prec)
    want = np.array([O.TRUTH[gate](int(a), int(b), int(c)) for a, b, c in bits.T])
    assert np.array_equal(ks.decrypt(out), want)


@pytest.mark.parametrize("gate", ["NAND", "MUX"])
def test_gate_n2048_fp64_bit_exact_vs_oracle(env2048, gate):
    ks, dk, keys = env2048
    rng = np.random.default_rng(2054)
    bits = rng.integers(0, 2, (3, 2)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = dk.gate(gate, ca, cb, cc if gate == "MUX" else None, prec=0)
    for i in range(2):
        assert np.array_equal(out[i], O.gate(keys, gate, ca[i], cb[i], cc[i])), i
