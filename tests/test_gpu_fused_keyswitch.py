"""GPU: the keyswitch fused into the split-FP32 bootstrap kernel (an evaluated alternative to the
separate keyswitch kernel).  With both paths forced onto the same bootstrap kernel (one bootstrap per
wavefront), the fused outputs are bit-identical to bootstrap + separate keyswitch, and decrypt to the
gate's truth table; MUX keeps the separate keyswitch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def T():
    import tfhe_amd
    return tfhe_amd


@pytest.mark.parametrize("gate,cnt", [("NAND", 300), ("XOR", 7), ("AND", 1)])
def test_fused_keyswitch_bit_identical(gpu_env, gate, cnt):
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(6060 + cnt)
    bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    T().set_launch_form(2)
    try:
        print(f"[fused ks] {gate} x{cnt}", flush=True)
        ref = dk.gate(gate, ca, cb, prec=1)
        T().set_fused_keyswitch(True)
        try:
            got = dk.gate(gate, ca, cb, prec=1)
            mux = dk.gate("MUX", ca, cb, ca, prec=1)
        finally:
            T().set_fused_keyswitch(False)
    finally:
        T().set_launch_form(0)
    assert np.array_equal(got, ref)
    want = {"NAND": 1 - (bits[0] & bits[1]), "XOR": bits[0] ^ bits[1], "AND": bits[0] & bits[1]}[gate]
    assert np.array_equal(ks.decrypt(got), want)
    assert np.array_equal(ks.decrypt(mux), np.where(bits[0] == 1, bits[1], bits[0]))
