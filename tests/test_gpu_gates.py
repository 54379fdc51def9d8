"""End-to-end gate bootstrapping on the GPU: truth tables under decryption, the reference's
known answer (OR of 499 and 388 over 16 bits = 503, test_bench/client/alice.c:33-44 and
verify.c:31), bit-exact FP64 parity with the oracle on identical inputs, and batch edge cases."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P = O.PARAMS_1024


def T():
    import tfhe_amd
    return tfhe_amd


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("gate", list(O.GATES))
def test_gate_truth_table(gpu_env, gate, prec, rng):
    ks, dk, keys = gpu_env
    cnt = 96
    bits = rng.integers(0, 2, (3, cnt)).astype(np.int32)
    bits[:2, :8] = np.array([[0, 0, 1, 1, 0, 0, 1, 1], [0, 1, 0, 1, 0, 1, 0, 1]])
    bits[2, :8] = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = dk.gate(gate, ca, cb, cc if gate == "MUX" else None, prec=prec)
    got = ks.decrypt(out)
    want = np.array([O.TRUTH[gate](int(a), int(b), int(c)) for a, b, c in bits.T])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("gate", ["NAND", "XOR", "MUX"])
def test_gate_fp64_bit_exact_vs_oracle(gpu_env, gate, rng):
    ks, dk, keys = gpu_env
    bits = rng.integers(0, 2, (3, 3)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = dk.gate(gate, ca, cb, cc if gate == "MUX" else None, prec=0)
    for i in range(3):
        want = O.gate(keys, gate, ca[i], cb[i], cc[i])
        assert np.array_equal(out[i], want), i


@pytest.mark.parametrize("prec", [0, 1])
def test_known_answer_or_499_388(gpu_env, prec):
    ks, dk, _ = gpu_env
    a = np.array([(499 >> i) & 1 for i in range(16)], np.int32)
    b = np.array([(388 >> i) & 1 for i in range(16)], np.int32)
    out = dk.gate("OR", ks.encrypt(a), ks.encrypt(b), prec=prec)
    bits = ks.decrypt(out)
    assert int(sum(int(v) << i for i, v in enumerate(bits))) == (499 | 388) == 503


def test_empty_and_ragged_batches(gpu_env, rng):
    ks, dk, _ = gpu_env
    z = np.zeros((0, P.n + 1), np.int32)
    assert dk.gate("NAND", z, z).shape == (0, P.n + 1)
    for cnt in (1, 31, 33, 65):
        bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
        out = dk.gate("AND", ks.encrypt(bits[0]), ks.encrypt(bits[1]))
        assert np.array_equal(ks.decrypt(out), bits[0] & bits[1]), cnt


@pytest.mark.parametrize("cnt", [256, 257])
def test_split_latency_form_boundary(gpu_env, cnt, rng):
    """Split FP32 NANDs on either side of the latency form's limit (one bootstrap per CU on a 256-CU
    device: the latency form, then the one-wavefront kernel) all decrypt correctly."""
    ks, dk, _ = gpu_env
    bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    out = dk.gate("NAND", ca, cb, prec=1)
    assert np.array_equal(ks.decrypt(out), 1 - (bits[0] & bits[1]))


def test_full_batch_16384_both_paths(gpu_env, rng):
    """The headline workload's size (16384 NANDs in one launch, device-resident ciphertexts):
    every output decrypts to the plaintext NAND on both paths, the host-array entry point returns
    the device path's FP64 outputs bit for bit, and 48 gates spread over the batch (both ends
    included) are bit-exact with the oracle on the FP64 path."""
    ks, dk, keys = gpu_env
    cnt = 16384
    bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    want = 1 - (bits[0] & bits[1])
    d_a, d_b = T().DeviceBuffer.from_array(ca), T().DeviceBuffer.from_array(cb)
    d_out = T().DeviceBuffer(ca.nbytes)
    try:
        outs = {}
        for prec in (0, 1):
            dk.gate_dev("NAND", d_out.ptr, d_a.ptr, d_b.ptr, None, cnt, prec=prec)
            outs[prec] = d_out.to_array(ca.shape)
            assert np.array_equal(ks.decrypt(outs[prec]), want), prec
    finally:
        for b in (d_a, d_b, d_out):
            b.free()
    assert np.array_equal(dk.gate("NAND", ca, cb, prec=0), outs[0])
    idx = np.unique(np.concatenate([[0, 1, cnt - 2, cnt - 1], rng.choice(cnt, 44, replace=False)]))
    ref, _ = O.gate_batch(keys, "NAND", ca[idx], cb[idx])
    assert np.array_equal(outs[0][idx], ref)


def test_host_array_mux_batch_matches_device_path(gpu_env, rng):
    """A large, ragged host-array MUX batch (8269 gates, split-FP32): outputs decrypt to the
    plaintext MUX and are bit-identical to the device-resident entry point's."""
    ks, dk, _ = gpu_env
    cnt = 2 * 4096 + 77
    bits = rng.integers(0, 2, (3, cnt)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    out = dk.gate("MUX", ca, cb, cc, prec=1)
    assert np.array_equal(ks.decrypt(out), np.where(bits[0] == 1, bits[1], bits[2]))
    bufs = [T().DeviceBuffer.from_array(x) for x in (ca, cb, cc)]
    d_out = T().DeviceBuffer(ca.nbytes)
    try:
        dk.gate_dev("MUX", d_out.ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, cnt, prec=1)
        assert np.array_equal(d_out.to_array(ca.shape), out)
    finally:
        for b in bufs + [d_out]:
            b.free()


def test_chained_gates_stay_correct(gpu_env, rng):
    """Outputs feed the next level (bootstrapped noise, not fresh noise)."""
    ks, dk, _ = gpu_env
    cnt = 128
    bits = rng.integers(0, 2, cnt).astype(np.int32)
    ct = ks.encrypt(bits)
    for level in range(6):
        perm = rng.permutation(cnt)
        out = dk.gate("NAND", ct, ct[perm], prec=level % 2)
        bits = 1 - (bits & bits[perm])
        ct = out
        assert np.array_equal(ks.decrypt(ct), bits), level


@pytest.mark.parametrize("prec", [1, 0])
def test_million_chained_gates_no_decryption_failure(gpu_env, prec):
    """Zero decryption failures in 10^6 gates on each path: 62 levels of 16384 NANDs (1,015,808
    gates), each level's inputs the previous level's outputs under a random pairing, so every gate
    sees bootstrapped and keyswitched noise.  Decryption on the device; the expected bits follow the
    decrypted ones, so a failure is counted once."""
    ks, dk, _ = gpu_env
    batch, levels = 16384, 62
    rng = np.random.default_rng(101 + prec)
    bits = rng.integers(0, 2, batch).astype(np.int32)
    cur = T().DeviceBuffer.from_array(ks.encrypt(bits))
    other = T().DeviceBuffer(cur.nbytes)
    out = T().DeviceBuffer(cur.nbytes)
    fails = 0
    try:
        for level in range(levels):
            perm = rng.permutation(batch)
            pair = np.ascontiguousarray(cur.to_array((batch, P.n + 1))[perm])
            assert T().lib().tfhe_gpu_memcpy_h2d(other.ptr, pair.ctypes.data_as(T()._vp), pair.nbytes) == 0
            dk.gate_dev("NAND", out.ptr, cur.ptr, other.ptr, None, batch, prec=prec)
            got = T().decrypt_dev(out, ks.lwe_key, batch)
            want = 1 - (bits & bits[perm])
            fails += int((got != want).sum())
            bits = got
            cur, out = out, cur
    finally:
        for b in (cur, other, out):
            b.free()
    assert batch * levels >= 10**6
    assert fails == 0, (prec, fails)


def test_upstream_api_boots_gates(gpu_env):
    """bootsNAND/bootsMUX/bootsCONSTANT through the upstream-compatible entry points."""
    ks, _, _ = gpu_env
    L = T().lib()
    arr = L.new_gate_bootstrapping_ciphertext_array(4, ks.params_ptr)
    try:
        L.bootsSymEncrypt(ctypes.byref(arr[0]), 1, ks.ptr)
        L.bootsSymEncrypt(ctypes.byref(arr[1]), 0, ks.ptr)
        L.bootsCONSTANT(ctypes.byref(arr[2]), 1, ks.cloud_ptr)
        # MUX(sel=1, b=0, c=1) = 0, written over its own selector input
        L.bootsMUX(ctypes.byref(arr[3]), ctypes.byref(arr[0]), ctypes.byref(arr[1]), ctypes.byref(arr[2]),
                   ks.cloud_ptr)
        assert L.bootsSymDecrypt(ctypes.byref(arr[3]), ks.ptr) == 0
        L.bootsNAND.argtypes = [ctypes.c_void_p] * 4
        L.bootsNAND(ctypes.addressof(arr[0]), ctypes.addressof(arr[0]), ctypes.addressof(arr[2]), ks.cloud_ptr)
        assert L.bootsSymDecrypt(ctypes.byref(arr[0]), ks.ptr) == 0
    finally:
        L.delete_gate_bootstrapping_ciphertext_array(4, arr)
