"""GPU: launch form 6 -- the split-FP32 bootstrap with four bootstraps per wavefront (16 lanes each)
and the pointwise products on the matrix cores.  Its transform against the closed form and as a
round trip; blind rotations (counts that do not fill the last wavefront included) within the split
path's phase bounds next to the other split form; gates decrypt to their truth tables.  The form
reads a key layout that only keys created while it is selected carry."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_1024
N = P.N


def T():
    import tfhe_amd
    return tfhe_amd


def closed_form(p):
    j = np.arange(N // 2)
    k = np.arange(N // 2)
    z = (p[:N // 2] + 1j * p[N // 2:]) * np.exp(1j * np.pi * j / N)
    return (z[None, :] * np.exp(2j * np.pi * np.outer(k, j) / (N // 2))).sum(axis=1)


def test_q4_transform_closed_form_and_round_trip():
    rng = np.random.default_rng(707)
    polys = np.concatenate([rng.integers(-64, 64, (5, N)), rng.integers(-2**31, 2**31, (2, N), dtype=np.int64),
                            np.eye(N, dtype=np.int64)[[0, 1, 511, 512, 1023]]]).astype(np.int32)
    print(f"[q4] transforms of {polys.shape[0]} polynomials", flush=True)
    F = T().fft_q4_forward(polys)
    for i in range(polys.shape[0]):
        ref = closed_form(polys[i].astype(np.float64))
        err = np.abs(F[i] - ref).max()
        assert err <= 4e-7 * max(1.0, np.abs(ref).max()) * 8, (i, err)
    back = T().fft_q4_inverse(F)
    small = polys[:5].astype(np.float64)
    assert np.abs(back[:5] - small).max() < 1e-3


@pytest.fixture(scope="module")
def q4_env(gpu_env):
    """The session key set with a device key created while form 6 is selected (so that it carries
    the form's key layout)."""
    ks, _, keys = gpu_env
    T().set_launch_form(6)
    try:
        dk = T().DeviceKey.from_keyset(ks)
    finally:
        T().set_launch_form(0)
    yield ks, dk, keys
    dk.free()


def rotate(dk, acc, bara, form):
    T().set_launch_form(form)
    try:
        return dk.blind_rotate(acc, bara, prec=1)
    finally:
        T().set_launch_form(0)


@pytest.mark.parametrize("cnt", [5, 300])
def test_q4_blind_rotation_phase(q4_env, cnt):
    _, dk, keys = q4_env
    rng = np.random.default_rng(708 + cnt)
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P.n)).astype(np.int32)
    bara[1, ::7] = 0
    print(f"[q4] blind rotation of {cnt} accumulators", flush=True)
    res = {f: rotate(dk, acc, bara, f) for f in (2, 6)}
    stats = {}
    for f, out in res.items():
        errs = []
        for g in sorted({0, 1, cnt - 1} | set(np.linspace(0, cnt - 1, 24).astype(int).tolist())):
            rot = int((bara[g].astype(np.int64) * keys.lwe_key).sum() % (2 * N))
            ideal = O.poly_mul_xai(rot, acc[g, 1])
            ph = O.tlwe_phase(P, out[g], keys.tlwe_key)
            errs.append((ph.astype(np.int64) - ideal.astype(np.int64) + 2**31) % 2**32 - 2**31)
        e = np.abs(np.concatenate(errs)).astype(np.float64)
        stats[f] = (np.sqrt((e ** 2).mean()), e.max())
        print(f"form {f}: phase error rms {stats[f][0]:.3e} max {stats[f][1]:.3e}")
    assert stats[6][1] < 2**26
    assert stats[6][0] < 2.0 * stats[2][0] + 2**18


@pytest.mark.parametrize("gate,cnt", [("NAND", 300), ("MUX", 150), ("XOR", 3)])
def test_q4_gates(q4_env, gate, cnt):
    ks, dk, _ = q4_env
    rng = np.random.default_rng(709 + cnt)
    bits = rng.integers(0, 2, (3, cnt)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    T().set_launch_form(6)
    try:
        print(f"[q4] {gate} x{cnt}", flush=True)
        out = dk.gate(gate, ca, cb, cc if gate == "MUX" else None, prec=1)
    finally:
        T().set_launch_form(0)
    want = np.array([O.TRUTH[gate](int(a), int(b), int(c)) for a, b, c in bits.T])
    assert np.array_equal(ks.decrypt(out), want)


def test_q4_bootstrap_woks_returns_accumulator(q4_env):
    """A bootstrap without keyswitch that also asks for its accumulators (mode 0 with an accumulator
    output): under form 6 every returned sample is the extraction of the returned accumulator, and
    the samples decrypt to NAND of the inputs (7 bootstraps: the last wavefront is partly idle)."""
    ks, dk, keys = q4_env
    rng = np.random.default_rng(711)
    cnt = 7
    bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    x = np.stack([O.gate_combine(P, "NAND", ca[i], cb[i]) for i in range(cnt)])
    T().set_launch_form(6)
    try:
        print(f"[q4] bootstrap without keyswitch x{cnt}, accumulators returned", flush=True)
        u, acc = dk.bootstrap_woks(x, prec=T().PREC_SPLIT32, want_acc=True)
    finally:
        T().set_launch_form(0)
    for g in range(cnt):
        assert np.array_equal(u[g], O.extract(P, acc[g])), g
    want = 1 - (bits[0] & bits[1])
    assert np.array_equal(O.boots_decrypt(u, np.asarray(keys.tlwe_key).reshape(-1)), want)
