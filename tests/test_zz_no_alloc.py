"""GPU: the gate entry point that allocates nothing (fixed device scratch) against the ordinary
device-resident gate entry point, called through the C ABI: bit-identical outputs on two calls in a
row, the second smaller, on both precisions at N=1024 and on FP64 at N=2048; a count beyond the
fixed scratch is refused without a launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def T():
    import tfhe_amd
    return tfhe_amd


def run_pair(ks, dk, gate, count, prec, rng):
    L = T().lib()
    n1 = ks.fp.n + 1
    bits = rng.integers(0, 2, (3, count)).astype(np.int32)
    ins = [T().DeviceBuffer.from_array(ks.encrypt(bits[i])) for i in range(3)]
    o_ref = T().DeviceBuffer(count * n1 * 4)
    o_new = T().DeviceBuffer(count * n1 * 4)
    try:
        code = T().GATES[gate]
        c = ins[2].ptr if gate == "MUX" else None
        assert L.tfhe_gpu_gate_batch_dev(dk.ptr, prec, code, o_ref.ptr, ins[0].ptr, ins[1].ptr, c, count) == 0
        rc = L.tfhe_gpu_gate_batch_dev_noalloc(dk.ptr, prec, code, o_new.ptr, ins[0].ptr, ins[1].ptr, c, count)
        assert rc == 0, L.tfhe_gpu_last_error()
        T().sync()
        ref = o_ref.to_array((count, n1))
        new = o_new.to_array((count, n1))
        assert np.array_equal(new, ref), (gate, count, prec)
        want = {"NAND": 1 - (bits[0] & bits[1]), "XOR": bits[0] ^ bits[1],
                "MUX": np.where(bits[0] == 1, bits[1], bits[2])}[gate]
        assert np.array_equal(ks.decrypt(new), want)
    finally:
        for b in ins + [o_ref, o_new]:
            b.free()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_gate_batch_dev_n1024(gpu_env, prec):
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(77 + prec)
    print(f"[noalloc] N=1024 prec={prec}: 4092 NAND then 700 MUX then 3 XOR", flush=True)
    run_pair(ks, dk, "NAND", 4092, prec, rng)
    run_pair(ks, dk, "MUX", 700, prec, rng)
    run_pair(ks, dk, "XOR", 3, prec, rng)


def test_noalloc_matches_gate_batch_dev_n2048():
    T().init(0)
    ks = T().KeySet(2048, seed=(0,))
    dk = T().DeviceKey.from_keyset(ks)
    try:
        rng = np.random.default_rng(79)
        print("[noalloc] N=2048 FP64: 1023 MUX then 300 NAND", flush=True)
        run_pair(ks, dk, "MUX", 1023, 0, rng)
        run_pair(ks, dk, "NAND", 300, 0, rng)
    finally:
        dk.free()
        ks.free()


def test_noalloc_refuses_counts_beyond_scratch(gpu_env):
    ks, dk, _ = gpu_env
    L = T().lib()
    n1 = ks.fp.n + 1
    buf = T().DeviceBuffer(4093 * n1 * 4)
    try:
        rc = L.tfhe_gpu_gate_batch_dev_noalloc(dk.ptr, 0, T().GATES["NAND"], buf.ptr, buf.ptr, buf.ptr, None, 4093)
        assert rc == -1
        assert b"fixed scratch" in L.tfhe_gpu_last_error()
        rc = L.tfhe_gpu_gate_batch_dev_noalloc(dk.ptr, 0, T().GATES["MUX"], buf.ptr, buf.ptr, buf.ptr, buf.ptr, 2047)
        assert rc == -1
    finally:
        buf.free()
