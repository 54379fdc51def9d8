"""GPU: circuits evaluated level by level (one bootstrap launch per level) -- ripple-carry adders
against plaintext addition on both paths, FP64 outputs bit-exact against the oracle evaluating the
same netlist gate by gate, HIP-graph replays against eager runs, autotuned launch forms, NOT /
COPY / constants, and the single-level entry point against the gate batch."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = O.PARAMS_1024


def T():
    import tfhe_amd
    return tfhe_amd


class Recorder:
    """Builds a product circuit and keeps the netlist for the oracle."""

    def __init__(self, n_inputs):
        self.c = T().Circuit(n_inputs)
        self.ops = [("IN", i) for i in range(n_inputs)]
        self.outs = []

    def add(self, op, a, b=-1, c=-1):
        w = self.c.add(op, a, b, c)
        assert w == len(self.ops)
        self.ops.append((op, a, b, c))
        return w

    def constant(self, v):
        w = self.c.constant(v)
        assert w == len(self.ops)
        self.ops.append(("CONST", v))
        return w

    def output(self, w):
        self.c.output(w)
        self.outs.append(w)


def adder(bits):
    r = Recorder(2 * bits)
    p = [r.add("XOR", i, bits + i) for i in range(bits)]
    carry = r.constant(0)
    for i in range(bits):
        r.output(r.add("XOR", p[i], carry))
        carry = r.add("MUX", p[i], carry, i)
    r.output(carry)
    return r


def oracle_eval(keys, r, inputs):
    """Gate-by-gate oracle evaluation of one instance: inputs [n_inputs][n+1]."""
    vals = []
    for op in r.ops:
        if op[0] == "IN":
            vals.append(inputs[op[1]])
        elif op[0] == "CONST":
            t = np.zeros(P.n + 1, np.int32)
            t[P.n] = T().MU if op[1] else -T().MU
            vals.append(t)
        elif op[0] == "NOT":
            vals.append((-vals[op[1]].astype(np.int64)).astype(np.int32))
        elif op[0] == "COPY":
            vals.append(vals[op[1]].copy())
        else:
            name, a, b, c = op
            vals.append(O.gate(keys, name, vals[a], vals[b], vals[c] if name == "MUX" else None))
    return np.stack([vals[w] for w in r.outs])


def plain_eval(r, bits):
    """Plaintext evaluation: bits [n_inputs][batch] -> [n_outputs][batch]."""
    vals = []
    for op in r.ops:
        if op[0] == "IN":
            vals.append(bits[op[1]])
        elif op[0] == "CONST":
            vals.append(np.full(bits.shape[1], op[1], np.int32))
        elif op[0] == "NOT":
            vals.append(1 - vals[op[1]])
        elif op[0] == "COPY":
            vals.append(vals[op[1]])
        else:
            name, a, b, c = op
            vals.append(np.array([O.TRUTH[name](int(x), int(y), int(z)) for x, y, z in
                                  zip(vals[a], vals[b], vals[c] if name == "MUX" else vals[a])], np.int32))
    return np.stack([vals[w] for w in r.outs])


def encrypt_inputs(ks, bits):
    return np.stack([ks.encrypt(bits[i]) for i in range(bits.shape[0])])


def decrypt_outputs(ks, ct):
    return np.stack([ks.decrypt(ct[i]) for i in range(ct.shape[0])])


def operand_bits(vals, bits):
    return np.array([[(int(v) >> i) & 1 for v in vals] for i in range(bits)], np.int32)


def from_bits(rows):
    return [int(sum(int(rows[i][j]) << i for i in range(len(rows)))) for j in range(len(rows[0]))]


def test_adder4_fp64_bit_exact_vs_oracle(gpu_env):
    ks, dk, keys = gpu_env
    rng = np.random.default_rng(31)
    r = adder(4)
    x, y = rng.integers(0, 16, 3), rng.integers(0, 16, 3)
    bits = np.concatenate([operand_bits(x, 4), operand_bits(y, 4)])
    ct = encrypt_inputs(ks, bits)
    print("[circuits] 4-bit adder FP64 batch 3", flush=True)
    out = r.c.run(dk, ct, prec=0)
    got = from_bits(decrypt_outputs(ks, out))
    assert got == [int(a + b) for a, b in zip(x, y)]
    want = oracle_eval(keys, r, ct[:, 0])
    assert np.array_equal(out[:, 0], want)
    r.c.free()


@pytest.mark.parametrize("prec", [0, 1])
def test_adder32_matches_plain_addition(gpu_env, prec):
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(32 + prec)
    B = 48
    x = rng.integers(0, 2**32, B, dtype=np.uint64)
    y = rng.integers(0, 2**32, B, dtype=np.uint64)
    x[:3] = [0, 2**32 - 1, 2**31]
    y[:3] = [0, 1, 2**31]
    r = adder(32)
    bits = np.concatenate([operand_bits(x, 32), operand_bits(y, 32)])
    ct = encrypt_inputs(ks, bits)
    print(f"[circuits] 32-bit adder prec={prec} batch {B}", flush=True)
    out = r.c.run(dk, ct, prec=prec, graph=True, autotune=True)
    dec = decrypt_outputs(ks, out)
    s = from_bits(dec[:32])
    assert s == [int((a + b) % 2**32) for a, b in zip(x, y)]
    assert list(dec[32]) == [int((a + b) >> 32) for a, b in zip(x, y)]
    r.c.free()


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_match_eager(gpu_env, prec):
    """A captured graph is replayed three times on fresh inputs copied into the bound input buffer,
    with eager library work between replays; every replay must equal an eager run."""
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(40 + prec)
    r = adder(8)
    B = 16
    n1 = P.n + 1
    d_in = T().DeviceBuffer(16 * B * n1 * 4)
    d_out = T().DeviceBuffer(9 * B * n1 * 4)
    try:
        for rep in range(3):
            x, y = rng.integers(0, 256, B), rng.integers(0, 256, B)
            ct = encrypt_inputs(ks, np.concatenate([operand_bits(x, 8), operand_bits(y, 8)]))
            eager = r.c.run(dk, ct, prec=prec)              # separate buffers, no graph
            T().lib().tfhe_gpu_memcpy_h2d(d_in.ptr, ct.ctypes.data_as(T()._vp), ct.nbytes)
            dk.gate("NAND", ct[0], ct[1], prec=prec)       # unrelated eager work between replays
            print(f"[circuits] graph replay {rep} prec={prec}", flush=True)
            r.c.run_dev(dk, d_in, d_out, B, prec=prec, graph=True)
            T().sync()
            got = d_out.to_array((9, B, n1))
            assert np.array_equal(got, eager), rep
            assert from_bits(decrypt_outputs(ks, got)) == [int(a + b) for a, b in zip(x, y)]
    finally:
        d_in.free()
        d_out.free()
        r.c.free()


def test_not_copy_constants(gpu_env):
    ks, dk, keys = gpu_env
    rng = np.random.default_rng(50)
    r = Recorder(3)
    one = r.constant(1)
    zero = r.constant(0)
    na = r.add("NOT", 0)
    g1 = r.add("NAND", na, 1)              # NOT folded into the gate coefficients
    g2 = r.add("MUX", r.add("NOT", 2), g1, one)
    g3 = r.add("XOR", r.add("COPY", g2), zero)
    for w in (g1, g2, g3, na, one, r.add("NOT", zero), r.add("COPY", 1)):
        r.output(w)
    B = 8
    bits = rng.integers(0, 2, (3, B)).astype(np.int32)
    ct = encrypt_inputs(ks, bits)
    print("[circuits] NOT/COPY/constants FP64", flush=True)
    out = r.c.run(dk, ct, prec=0)
    assert np.array_equal(decrypt_outputs(ks, out), plain_eval(r, bits))
    assert np.array_equal(out[:, 0], oracle_eval(keys, r, ct[:, 0]))
    assert r.c.info()["levels"] == 3
    r.c.free()


def test_circuit_level_entry_point_matches_gate_batch(gpu_env):
    ks, dk, _ = gpu_env
    rng = np.random.default_rng(60)
    B = 10
    bits = rng.integers(0, 2, (3, B)).astype(np.int32)
    ct = encrypt_inputs(ks, bits)
    bufs = [T().DeviceBuffer.from_array(ct[i]) for i in range(3)]
    outs = [T().DeviceBuffer(ct[0].nbytes) for _ in range(3)]
    ops = ["NAND", "MUX", "XOR"]
    try:
        print("[circuits] circuit_level NAND/MUX/XOR FP64", flush=True)
        T().circuit_level(dk, ops, [bufs[0], bufs[0], bufs[1]], [bufs[1], bufs[1], bufs[2]],
                          [None, bufs[2], None], outs, B, prec=0)
        got = [o.to_array(ct[0].shape) for o in outs]
        want = [dk.gate("NAND", ct[0], ct[1], prec=0), dk.gate("MUX", ct[0], ct[1], ct[2], prec=0),
                dk.gate("XOR", ct[1], ct[2], prec=0)]
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    finally:
        for b in bufs + outs:
            b.free()


def test_graph_rebinds_after_key_replaced():
    """A circuit bound (and graph-captured) with one device key, then run with GRAPH on a key
    created after the first was freed -- possibly at the same host address -- must rebind and
    recapture: the outputs decrypt under the new key set."""
    r = adder(8)
    B = 8
    n1 = P.n + 1
    rng = np.random.default_rng(70)
    d_in = T().DeviceBuffer(16 * B * n1 * 4)
    d_out = T().DeviceBuffer(9 * B * n1 * 4)
    try:
        for seed in (101, 202):
            ks = T().KeySet(1024, seed=(seed,))
            dk = T().DeviceKey.from_keyset(ks)
            try:
                x, y = rng.integers(0, 256, B), rng.integers(0, 256, B)
                ct = encrypt_inputs(ks, np.concatenate([operand_bits(x, 8), operand_bits(y, 8)]))
                T().lib().tfhe_gpu_memcpy_h2d(d_in.ptr, ct.ctypes.data_as(T()._vp), ct.nbytes)
                print(f"[circuits] graph run on a fresh key (seed {seed})", flush=True)
                r.c.run_dev(dk, d_in, d_out, B, prec=0, graph=True)
                T().sync()
                got = decrypt_outputs(ks, d_out.to_array((9, B, n1)))
                assert from_bits(got) == [int(a + b) for a, b in zip(x, y)], seed
            finally:
                dk.free()
                ks.free()
    finally:
        d_in.free()
        d_out.free()
        r.c.free()


@pytest.mark.parametrize("prec", [0, 1])
def test_adders_4096_full_workload(gpu_env, prec):
    """The 4096-adder workload itself: 4096 independent 32-bit ripple-carry adders evaluated as one
    circuit over a batch of 4096 (33 bootstrapped levels: 32 x 4096 XOR, then MUX + XOR levels of
    12,288 bootstraps), graph-replayed with autotuned launch forms.  Every sum and carry matches
    plain addition (edge operands included); on FP64 the 33 outputs of the first and the last
    instance are bit-exact against the oracle evaluating the same netlist gate by gate (the
    reference's circuit shape: gates called in host loops, test_bench/server/cloud.c:8-37)."""
    ks, dk, keys = gpu_env
    B = 4096
    n1 = P.n + 1
    rng = np.random.default_rng(3 + prec)
    x = rng.integers(0, 2**32, B, dtype=np.uint64)
    y = rng.integers(0, 2**32, B, dtype=np.uint64)
    x[:3] = [0, 2**32 - 1, 2**31]
    y[:3] = [0, 1, 2**31]
    x[-1], y[-1] = 2**32 - 1, 2**32 - 1
    r = adder(32)
    bits = np.concatenate([operand_bits(x, 32), operand_bits(y, 32)])
    ct = encrypt_inputs(ks, bits)
    d_in = T().DeviceBuffer.from_array(ct)
    d_out = T().DeviceBuffer(33 * B * n1 * 4)
    try:
        print(f"[circuits] 4096 32-bit adders prec={prec}", flush=True)
        r.c.run_dev(dk, d_in, d_out, B, prec=prec, graph=True, autotune=True)
        T().sync()
        out = d_out.to_array((33, B, n1))
        dec = np.stack([T().decrypt_dev(d_out.ptr.value + i * B * n1 * 4, ks.lwe_key, B) for i in range(33)])
        assert from_bits(dec[:32]) == [int((a + b) % 2**32) for a, b in zip(x, y)]
        assert list(dec[32]) == [int((a + b) >> 32) for a, b in zip(x, y)]
        assert np.array_equal(decrypt_outputs(ks, out[:, [0, B - 1]]), dec[:, [0, B - 1]])
        if prec == 0:
            for g in (0, B - 1):
                print(f"[circuits] oracle gate-by-gate evaluation of instance {g}", flush=True)
                assert np.array_equal(out[:, g], oracle_eval(keys, r, ct[:, g])), g
    finally:
        d_in.free()
        d_out.free()
        r.c.free()
