"""GPU: the launch forms of the FP64 bootstrap (key rows shared by a workgroup of 8 bootstraps, one
bootstrap per workgroup, and the latency form that spreads one bootstrap over six wavefronts)
against the oracle and against each other.  Batch sizes that are
not multiples of 8 leave idle wavefronts in the last workgroup; zero rotations make some
bootstraps of a workgroup skip steps that the others run."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_1024
N = P.N


def T():
    import tfhe_amd
    return tfhe_amd


@pytest.fixture
def forms():
    yield
    T().set_fp64_shared_rows("auto")


@pytest.mark.parametrize("shared", [True, False, "auto"])
def test_blind_rotate_fp64_forms_match_oracle(gpu_env, forms, shared):
    """Key-sharing workgroups (True), one bootstrap per workgroup (False), and the default choice
    for a launch this small: the latency form, six wavefronts per bootstrap."""
    _, dk, keys = gpu_env
    T().set_fp64_shared_rows(shared)
    rng = np.random.default_rng(77)
    cnt = 11
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P.n)).astype(np.int32)
    bara[2, ::2] = 0          # half of one bootstrap's steps are skipped
    bara[9, :] = 0            # one bootstrap skips every step: its accumulator is unchanged
    bara[:, -1] = 0           # the last step is skipped by everyone
    print(f"[fp64 forms] shared={shared} blind rotate count={cnt}", flush=True)
    got = dk.blind_rotate(acc, bara, prec=0)
    assert np.array_equal(got[9], acc[9])
    for g in (0, 2, 7, 8, 10):
        assert np.array_equal(got[g], O.blind_rotate(keys, acc[g], bara[g])), g


@pytest.mark.parametrize("cnt", [8 * 5 + 3, 256, 257, 1100])
def test_fp64_forms_agree_on_gates(gpu_env, forms, cnt):
    """Both forms forced, and the default choice by size (latency form up to one bootstrap per CU,
    one per workgroup up to 4 per CU, key sharing above), give identical outputs; 256 and 257 NANDs
    sit on either side of the latency form's limit on a 256-CU device."""
    ks, dk, keys = gpu_env
    rng = np.random.default_rng(78)
    bits = rng.integers(0, 2, (3, cnt)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = {}
    for shared in (True, False, "auto"):
        T().set_fp64_shared_rows(shared)
        print(f"[fp64 forms] shared={shared} NAND/MUX count={cnt}", flush=True)
        out[shared] = (dk.gate("NAND", ca, cb, prec=0), dk.gate("MUX", ca, cb, cc, prec=0))
    for k in range(2):
        assert np.array_equal(out[True][k], out[False][k])
        assert np.array_equal(out["auto"][k], out[False][k])
    assert np.array_equal(ks.decrypt(out[True][0]), 1 - (bits[0] & bits[1]))
    assert np.array_equal(ks.decrypt(out[True][1]), np.where(bits[0] == 1, bits[1], bits[2]))
    for i in (0, cnt - 1):
        assert np.array_equal(out[True][0][i], O.gate(keys, "NAND", ca[i], cb[i])), i


def test_fp64_shared_rows_flag_rejects_bad_values():
    import tfhe_amd
    for bad in (-1, 3):
        with pytest.raises(tfhe_amd.TfheError):
            tfhe_amd._ck(tfhe_amd.lib().tfhe_gpu_set_fp64_shared_rows(bad), "set_fp64_shared_rows")
