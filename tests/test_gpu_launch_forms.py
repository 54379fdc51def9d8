"""GPU: every launch form the circuit autotuner chooses between, forced through
tfhe_gpu_set_launch_form.  FP64 blind rotations are bit-identical across forms (and with the oracle);
split-FP32 ones differ only in FP32 summation order and stay within the stated phase bounds.  A
levelised circuit run with autotuning reports the form each level took."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P1, P2 = O.PARAMS_1024, O.PARAMS_2048


def T():
    import tfhe_amd
    return tfhe_amd


@pytest.fixture(scope="module")
def env2048():
    T().init(0)
    ks = T().KeySet(2048, seed=(0,))
    dk = T().DeviceKey.from_keyset(ks)
    keys = O.Keys(P2, ks.lwe_key, ks.tlwe_key, ks.ksk, ks.bk)
    yield ks, dk, keys
    dk.free()
    keys.free()
    ks.free()


def rotate_forms(dk, acc, bara, prec, forms):
    out = {}
    try:
        for f in forms:
            T().set_launch_form(f)
            print(f"[forms] prec={prec} form={f} count={acc.shape[0]}", flush=True)
            out[f] = dk.blind_rotate(acc, bara, prec=prec)
    finally:
        T().set_launch_form(0)
    return out


def phase_errors(P, keys, acc, bara, res, idx):
    N = P.N
    errs = []
    for g in idx:
        rot = int((bara[g].astype(np.int64) * keys.lwe_key).sum() % (2 * N))
        ideal = O.poly_mul_xai(rot, acc[g, 1])
        ph = O.tlwe_phase(P, res[g], keys.tlwe_key)
        errs.append((ph.astype(np.int64) - ideal.astype(np.int64) + 2**31) % 2**32 - 2**31)
    return np.abs(np.concatenate(errs)).astype(np.float64)


def test_forms_n1024(gpu_env):
    _, dk, keys = gpu_env
    rng = np.random.default_rng(5151)
    cnt = 40
    N = P1.N
    acc = np.stack([O.test_vector(P1, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P1.n)).astype(np.int32)
    f64 = rotate_forms(dk, acc, bara, 0, [1, 2, 3])
    for f in (2, 3):
        assert np.array_equal(f64[f], f64[1]), f
    for g in (0, cnt - 1):
        assert np.array_equal(f64[1][g], O.blind_rotate(keys, acc[g], bara[g])), g
    s32 = rotate_forms(dk, acc, bara, 1, [2, 3])
    for f, res in s32.items():
        e = phase_errors(P1, keys, acc, bara, res, range(cnt))
        print(f"N=1024 split form {f}: phase error rms {np.sqrt((e**2).mean()):.3e} max {e.max():.3e}")
        assert e.max() < 2**26


def test_forms_n2048(env2048):
    _, dk, keys = env2048
    rng = np.random.default_rng(5252)
    cnt = 40
    N = P2.N
    acc = np.stack([O.test_vector(P2, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P2.n)).astype(np.int32)
    f64 = rotate_forms(dk, acc, bara, 0, [2, 3, 4, 5])
    for f in (3, 4, 5):
        assert np.array_equal(f64[f], f64[2]), f
    for g in (0, cnt - 1):
        assert np.array_equal(f64[2][g], O.blind_rotate(keys, acc[g], bara[g])), g
    s32 = rotate_forms(dk, acc, bara, 1, [2, 3, 4, 5])
    for f, res in s32.items():
        e = phase_errors(P2, keys, acc, bara, res, range(cnt))
        print(f"N=2048 split form {f}: phase error rms {np.sqrt((e**2).mean()):.3e} max {e.max():.3e}")
        assert np.sqrt((e ** 2).mean()) < 2**25
        assert e.max() < 2**27


def test_autotuned_levels_report_their_forms(gpu_env):
    """An 8-bit adder over a batch of 400 instances (levels of 3200 and 1200 bootstraps), run with
    autotuning on both precisions: every level reports a form among its candidates, and the sums are
    right."""
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(5353)
    B = 400
    c = T().ripple_carry_adder(8)
    x, y = rng.integers(0, 256, B), rng.integers(0, 256, B)
    bits = np.array([[(int(v) >> i) & 1 for v in vals] for vals in (x, y) for i in range(8)], np.int32)
    ct = np.stack([ks.encrypt(bits[i]) for i in range(16)])
    try:
        for prec, allowed in ((0, {1, 2}), (1, {2})):
            print(f"[forms] autotuned adder prec={prec}", flush=True)
            out = c.run(dk, ct, prec=prec, autotune=True)
            dec = np.stack([ks.decrypt(out[i]) for i in range(9)])
            assert [int(sum(int(dec[i][j]) << i for i in range(9))) for j in range(B)] == [int(a + b) for a, b in zip(x, y)]
            plan = c.level_plan()
            print(f"prec {prec} level plan (bootstraps per instance, form): {plan}")
            for nb, form in plan:
                jobs = nb * B
                assert form in (allowed | ({3} if jobs <= 1024 else set())), (nb, form)
    finally:
        c.free()
