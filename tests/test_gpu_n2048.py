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


def closed_form_transform(p):
    """F_k = p(w^(4k+1)), w = e^{i pi/N}: direct evaluation at the roots of X^N + 1."""
    j = np.arange(N)
    k = np.arange(N // 2)
    e = (j[None, :] * (4 * k + 1)[:, None]) % (2 * N)
    return (np.asarray(p, np.float64)[None, :] * np.exp(1j * np.pi * e / N)).sum(axis=1)


def test_fft_forward_digits_n2048_matrix_core_matches_closed_form():
    """The split path's forward transform of gadget digits as the two-wavefront bootstrap kernel
    runs it (first pass on the matrix cores: two FP16 pieces of the twisted 8-point DFT over the two
    folded inputs, FP32 accumulation) against direct evaluation at the roots of X^N + 1: random
    digits of the Bgbit = 8 range, the extreme digits, impulses at the ends of each quarter (each
    lands on another matrix-core operand column), and the largest accepted inputs."""
    rng = np.random.default_rng(2050)
    polys = np.concatenate([rng.integers(-128, 128, (4, N)), np.full((1, N), -128), np.full((1, N), 127),
                            rng.choice([-2047, 2047], (1, N)),
                            np.eye(N, dtype=np.int64)[[0, 1, 63, 511, 512, 1023, 1024, 1535, 1536, N - 1]]
                            ]).astype(np.int32)
    got = T().fft_forward_digits(polys)
    ref_fp32 = T().fft_forward(polys, prec=1)
    for i in range(polys.shape[0]):
        ref = closed_form_transform(polys[i])
        scale = max(1.0, np.abs(ref).max())
        err = np.abs(got[i] - ref).max()
        assert err <= 4e-7 * scale * 8, (i, err)
        assert err <= 4 * max(np.abs(ref_fp32[i] - ref).max(), 1e-7 * scale), (i, err)


def test_fft_forward_digits_n2048_rejects_out_of_range_inputs():
    polys = np.zeros((1, N), np.int32)
    polys[0, N - 1] = -2048
    with pytest.raises(T().TfheError, match="inputs must lie"):
        T().fft_forward_digits(polys)


@pytest.mark.parametrize("prec,rtol", [(0, 1e-12), (1, 4e-7)])
def test_fft_inverse_n2048_matches_closed_form(prec, rtol):
    """The inverse transform the bootstrap kernels run at N = 2048 (for the split path: pass 1's
    twiddles in the matrix-core forward's layout, untwist in FP64) applied to exact spectra
    (direct evaluation of known polynomials) returns the polynomials."""
    rng = np.random.default_rng(2051)
    polys = np.concatenate([rand_torus(rng, (2, N)), rng.integers(-128, 128, (2, N)).astype(np.int32),
                            np.eye(N, dtype=np.int32)[[0, 1, 1023, 1024, N - 1]] * 1000])
    F = np.stack([closed_form_transform(p) for p in polys])
    back = T().fft_inverse(F, prec=prec)
    for i in range(polys.shape[0]):
        err = np.abs(back[i] - polys[i]).max()
        assert err <= rtol * max(1.0, np.abs(polys[i]).max()) * 8, (i, err)


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


@pytest.mark.parametrize("cnt", [2, 300])
def test_blind_rotate_n2048_fp64_forms(env2048, cnt):
    """Launches of at most one bootstrap per CU run the latency form (one workgroup per bootstrap,
    a wavefront pair per level), larger ones the batched two-wavefront kernel.  Accumulators 0, 1,
    the last and five spread between them are bit-exact with the oracle; on the small launch the
    latency form and the batched kernel (forced through the FP64 form switch) agree bit for bit."""
    _, dk, keys = env2048
    rng = np.random.default_rng(2055 + cnt)
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P.n)).astype(np.int32)
    bara[0, ::5] = 0
    got = dk.blind_rotate(acc, bara, prec=0)
    for g in sorted({0, 1, cnt - 1} | set(np.linspace(0, cnt - 1, 7).astype(int).tolist())):
        assert np.array_equal(got[g], O.blind_rotate(keys, acc[g], bara[g])), g
    if cnt <= 256:
        T().set_fp64_shared_rows(True)
        try:
            batched = dk.blind_rotate(acc, bara, prec=0)
        finally:
            T().set_fp64_shared_rows("auto")
        assert np.array_equal(batched, got)


@pytest.mark.parametrize("cnt", [2, 300])
def test_blind_rotate_n2048_split32_phase_error(env2048, cnt):
    """Split FP32 at N=2048, latency form (2 accumulators) and batched kernel (300): the TLWE phase
    of 24 accumulators spread over the launch (both ends included) against the noiseless rotation,
    beside the FP64 path's.  At
    this set the key noise is small (sigma 2^-30: FP64 rms about 2^20 units) and the split path's
    own error dominates, about 2^24.5 units rms after the 750 steps."""
    _, dk, keys = env2048
    rng = np.random.default_rng(2057 + cnt)
    acc = np.stack([O.test_vector(P, int(b), T().MU) for b in rng.integers(0, 2 * N, cnt)])
    bara = rng.integers(0, 2 * N, (cnt, P.n)).astype(np.int32)
    res = {p: dk.blind_rotate(acc, bara, prec=p) for p in (0, 1)}
    rms, worst, mean = {}, {}, {}
    checked = sorted({0, 1, cnt - 1} | set(np.linspace(0, cnt - 1, 24).astype(int).tolist()))
    for p in (0, 1):
        errs = []
        for g in checked:
            rot = int((bara[g].astype(np.int64) * keys.lwe_key).sum() % (2 * N))
            ideal = O.poly_mul_xai(rot, acc[g, 1])
            ph = O.tlwe_phase(P, res[p][g], keys.tlwe_key)
            errs.append((ph.astype(np.int64) - ideal.astype(np.int64) + 2**31) % 2**32 - 2**31)
        d = np.concatenate(errs).astype(np.float64)
        e = np.abs(d)
        rms[p], worst[p], mean[p] = np.sqrt((e ** 2).mean()), e.max(), d.mean()
    print(f"N=2048 blind rotation phase noise (2^-32 units), count {cnt}: FP64 rms {rms[0]:.3e} max {worst[0]:.3e}; "
          f"split rms {rms[1]:.3e} max {worst[1]:.3e} mean {mean[1]:.3e}")
    assert rms[0] < 2**22
    # stated bounds: split rms within 2^25 and no coefficient beyond 2^27, half of the 1/16 gate
    # decision margin (2^28 units)
    assert rms[1] < 2**25
    assert worst[1] < 2**27


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cnt", [300, 259])
def test_gate_n2048_batched_nand(env2048, prec, cnt):
    """NAND batches above one bootstrap per CU run the batched kernel, four bootstraps per
    workgroup; 259 leaves one idle slot in the last workgroup (it still meets every barrier).  All
    outputs decrypt correctly; on FP64 the last gate is bit-exact with the oracle."""
    ks, dk, keys = env2048
    rng = np.random.default_rng(2059 + cnt)
    bits = rng.integers(0, 2, (2, cnt)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    out = dk.gate("NAND", ca, cb, prec=prec)
    assert np.array_equal(ks.decrypt(out), 1 - (bits[0] & bits[1]))
    if prec == 0:
        assert np.array_equal(out[cnt - 1], O.gate(keys, "NAND", ca[cnt - 1], cb[cnt - 1], None))


@pytest.mark.parametrize("prec", [0, 1])
def test_upstream_api_n2048(env2048, prec):
    """bootsSymEncrypt / bootsCONSTANT / bootsMUX / bootsNAND / bootsXOR on an N=2048 key set, one
    gate per call (the latency form and the sliced keyswitch), outputs written over inputs."""
    import ctypes
    ks, _, _ = env2048
    L = T().lib()
    L.tfhe_gpu_set_default_precision(prec)
    arr = L.new_gate_bootstrapping_ciphertext_array(4, ks.params_ptr)
    try:
        L.bootsSymEncrypt(ctypes.byref(arr[0]), 1, ks.ptr)
        L.bootsSymEncrypt(ctypes.byref(arr[1]), 0, ks.ptr)
        L.bootsCONSTANT(ctypes.byref(arr[2]), 1, ks.cloud_ptr)
        print(f"[n2048] upstream gates prec={prec}", flush=True)
        # MUX(sel=1, b=0, c=1) = 0
        L.bootsMUX(ctypes.byref(arr[3]), ctypes.byref(arr[0]), ctypes.byref(arr[1]), ctypes.byref(arr[2]),
                   ks.cloud_ptr)
        assert L.bootsSymDecrypt(ctypes.byref(arr[3]), ks.ptr) == 0
        for name in ("bootsNAND", "bootsXOR"):
            getattr(L, name).argtypes = [ctypes.c_void_p] * 4
        L.bootsXOR(ctypes.addressof(arr[1]), ctypes.addressof(arr[1]), ctypes.addressof(arr[2]), ks.cloud_ptr)
        assert L.bootsSymDecrypt(ctypes.byref(arr[1]), ks.ptr) == 1          # XOR(0, 1)
        L.bootsNAND(ctypes.addressof(arr[0]), ctypes.addressof(arr[0]), ctypes.addressof(arr[2]), ks.cloud_ptr)
        assert L.bootsSymDecrypt(ctypes.byref(arr[0]), ks.ptr) == 0          # NAND(1, 1)
    finally:
        L.delete_gate_bootstrapping_ciphertext_array(4, arr)
        L.tfhe_gpu_set_default_precision(0)


def _adder_netlist(bits):
    """Ripple-carry adder as a product circuit plus the recorded netlist: p_i = XOR(a_i, b_i),
    s_i = XOR(p_i, c_i), c_{i+1} = MUX(p_i, c_i, a_i), c_0 = 0."""
    c = T().Circuit(2 * bits)
    ops = [("IN", i) for i in range(2 * bits)]

    def add(op, a, b=-1, cc=-1):
        w = c.add(op, a, b, cc)
        assert w == len(ops)
        ops.append((op, a, b, cc))
        return w

    p = [add("XOR", i, bits + i) for i in range(bits)]
    carry = c.constant(0)
    assert carry == len(ops)
    ops.append(("CONST", 0))
    outs = []
    for i in range(bits):
        s = add("XOR", p[i], carry)
        c.output(s)
        outs.append(s)
        carry = add("MUX", p[i], carry, i)
    c.output(carry)
    outs.append(carry)
    return c, ops, outs


@pytest.mark.parametrize("prec", [0, 1])
def test_circuit_n2048_adder_graph_replay(env2048, prec):
    """A 4-bit adder at N=2048 over 3 instances, levels small enough for the latency form, captured
    in a HIP graph (with the FP64 form autotuner) and replayed on fresh inputs in the bound buffer:
    every replay equals an eager run and decrypts to the plain sums; FP64 outputs of one instance
    are bit-exact with the oracle evaluating the netlist gate by gate."""
    ks, dk, keys = env2048
    rng = np.random.default_rng(2060 + prec)
    c, ops, outs = _adder_netlist(4)
    B, n1 = 3, P.n + 1
    d_in = T().DeviceBuffer(8 * B * n1 * 4)
    d_out = T().DeviceBuffer(5 * B * n1 * 4)
    try:
        for rep in range(2):
            x, y = rng.integers(0, 16, B), rng.integers(0, 16, B)
            bits = np.array([[(int(v) >> i) & 1 for v in vals] for vals in (x, y) for i in range(4)], np.int32)
            ct = np.stack([ks.encrypt(bits[i]) for i in range(8)])
            eager = c.run(dk, ct, prec=prec)
            T().lib().tfhe_gpu_memcpy_h2d(d_in.ptr, ct.ctypes.data_as(T()._vp), ct.nbytes)
            print(f"[n2048] adder graph run {rep} prec={prec}", flush=True)
            c.run_dev(dk, d_in, d_out, B, prec=prec, graph=True, autotune=True)
            T().sync()
            got = d_out.to_array((5, B, n1))
            assert np.array_equal(got, eager), rep
            dec = np.stack([ks.decrypt(got[o]) for o in range(5)])
            sums = [int(sum(int(dec[o][j]) << o for o in range(5))) for j in range(B)]
            assert sums == [int(a + b) for a, b in zip(x, y)], rep
        if prec == 0:
            vals = []
            for op in ops:
                if op[0] == "IN":
                    vals.append(ct[op[1], 0])
                elif op[0] == "CONST":
                    t = np.zeros(n1, np.int32)
                    t[P.n] = T().MU if op[1] else -T().MU
                    vals.append(t)
                else:
                    name, a, b, cc = op
                    vals.append(O.gate(keys, name, vals[a], vals[b], vals[cc] if name == "MUX" else None))
            assert np.array_equal(got[:, 0], np.stack([vals[w] for w in outs]))
    finally:
        d_in.free()
        d_out.free()
        c.free()


def test_split_point_override_n2048(env2048):
    """tfhe_gpu_set_split_bits applies to keys created afterwards: a key made at a 26-bit split
    gives external products within the stated bound, about twice the default 25-bit split's error,
    and correct gates; 0 restores the default for later keys."""
    ks, dk, keys = env2048
    rng = np.random.default_rng(2070)
    acc = rand_torus(rng, (3, 2, N))
    rows = np.array([2, 333, 747], np.int32)
    T().set_split_bits(26)
    try:
        dk26 = T().DeviceKey.from_keyset(ks)
    finally:
        T().set_split_bits(0)
    try:
        rms = {}
        for name, key in (("26", dk26), ("default", dk)):
            got = key.extern_product(acc, rows, prec=1)
            e = np.concatenate([np.abs((got[g].astype(np.int64) - O.extern_product_exact(P, acc[g], keys.bk[rows[g]])
                                        .astype(np.int64) + 2**31) % 2**32 - 2**31).ravel() for g in range(3)])
            rms[name] = float(np.sqrt((e.astype(float) ** 2).mean()))
            print(f"N=2048 split {name}: external product error max {e.max()} rms {rms[name]:.1f}", flush=True)
            assert e.max() < 2**18
        assert 1.5 < rms["26"] / rms["default"] < 2.7
        bits = rng.integers(0, 2, (2, 24)).astype(np.int32)
        out = dk26.gate("NAND", ks.encrypt(bits[0]), ks.encrypt(bits[1]), prec=1)
        assert np.array_equal(ks.decrypt(out), 1 - (bits[0] & bits[1]))
    finally:
        dk26.free()


def test_split_hi_rounding_margin_n2048(env2048):
    """As at N=1024: the hi-half product before rounding stays within 0.4 of its exact integer over
    131,072 coefficients (32 external products), 8 of them against the exact int64 product with the
    hi-half key.  The default split point at N=2048 is 25 bits (hi half 7 bits)."""
    ks, dk, keys = env2048
    assert dk.split_bits == 25
    rng = np.random.default_rng(2425)
    cnt = 32
    acc = rand_torus(rng, (cnt, 2, N))
    rows = rng.integers(0, P.n, cnt).astype(np.int32)
    got, hi = dk.extern_product_split_hi(acc, rows)
    assert np.array_equal(got, dk.extern_product(acc, rows, prec=1))
    dist = np.abs(hi - np.rint(hi))
    worst_exact = 0.0
    for g in range(8):
        hk = ((keys.bk[rows[g]].astype(np.int64) + (1 << 24)) >> 25).astype(np.int32)
        _, exact = O.extern_product_exact(P, acc[g], hk, want_raw=True)
        worst_exact = max(worst_exact, float(np.abs(hi[g] - exact.astype(np.float64)).max()))
    print(f"N=2048 split 25: hi product distance to an integer max {dist.max():.4f} rms {np.sqrt((dist**2).mean()):.4f}"
          f" over {dist.size} coefficients; max |hi - exact| = {worst_exact:.4f}; |hi| max {np.abs(hi).max():.3e}",
          flush=True)
    assert dist.max() < 0.4
    assert worst_exact < 0.4


@pytest.mark.parametrize("prec", [0, 1])
def test_mux_xor_8192_full_workload(env2048, prec):
    """The N=2048 workload itself: 8192 MUX (16,384 bootstraps in one launch) and 8192 XOR on random
    bits through the device-resident gate entry point.  Every output decrypts to its plaintext gate;
    on FP64 the four gates at each end of both batches (MUX: both segment ends of its bootstrap
    launch) are bit-exact against the oracle."""
    ks, dk, keys = env2048
    G = 8192
    n1 = P.n + 1
    rng = np.random.default_rng(5 + prec)
    bits = rng.integers(0, 2, (3, G)).astype(np.int32)
    cts = [ks.encrypt(bits[i]) for i in range(3)]
    bufs = [T().DeviceBuffer.from_array(c) for c in cts]
    outs = [T().DeviceBuffer(G * n1 * 4) for _ in range(2)]
    try:
        print(f"[n2048] 8192 MUX + 8192 XOR prec={prec}", flush=True)
        dk.gate_dev("MUX", outs[0].ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, G, prec=prec)
        dk.gate_dev("XOR", outs[1].ptr, bufs[0].ptr, bufs[1].ptr, None, G, prec=prec)
        T().sync()
        mux = outs[0].to_array((G, n1))
        xor = outs[1].to_array((G, n1))
        assert np.array_equal(ks.decrypt(mux), np.where(bits[0] == 1, bits[1], bits[2]))
        assert np.array_equal(ks.decrypt(xor), bits[0] ^ bits[1])
        if prec == 0:
            idx = [0, 1, 2, 3, G - 4, G - 3, G - 2, G - 1]
            sel = [c[idx] for c in cts]
            want_mux, _ = O.gate_batch(keys, "MUX", sel[0], sel[1], sel[2], nthreads=8)
            want_xor, _ = O.gate_batch(keys, "XOR", sel[0], sel[1], nthreads=8)
            assert np.array_equal(mux[idx], want_mux)
            assert np.array_equal(xor[idx], want_xor)
    finally:
        for b in bufs + outs:
            b.free()


@pytest.mark.parametrize("prec", [1, 0])
def test_mux_xor_chained_million_gates_no_decryption_failure(env2048, prec):
    """The N=2048 failure rate: zero decryption failures in 1,015,808 gates on each path, as 62
    chained levels of the workload's shape (8192 MUX and 8192 XOR).  Each level draws its inputs
    from the previous level's 16384 outputs under random pairings, so every gate sees bootstrapped
    and keyswitched noise.  Decryption on the device; the expected bits follow the decrypted ones,
    so a failure is counted once."""
    ks, dk, _ = env2048
    G, levels = 8192, 62
    n1 = P.n + 1
    rng = np.random.default_rng(2048 + prec)
    bits = rng.integers(0, 2, 2 * G).astype(np.int32)
    cur = T().DeviceBuffer.from_array(ks.encrypt(bits))
    ins = [T().DeviceBuffer(2 * G * n1 * 4) for _ in range(3)]
    out = T().DeviceBuffer(2 * G * n1 * 4)
    half = G * n1 * 4
    fails = 0
    print(f"[n2048] chained MUX/XOR levels prec={prec}", flush=True)
    try:
        for level in range(levels):
            host = cur.to_array((2 * G, n1))
            perms = [rng.permutation(2 * G) for _ in range(3)]
            for buf, pm in zip(ins, perms):
                arr = np.ascontiguousarray(host[pm])
                assert T().lib().tfhe_gpu_memcpy_h2d(buf.ptr, arr.ctypes.data_as(T()._vp), arr.nbytes) == 0
            a, b, c = (buf.ptr.value for buf in ins)
            dk.gate_dev("MUX", out.ptr.value, a, b, c, G, prec=prec)
            dk.gate_dev("XOR", out.ptr.value + half, a + half, b + half, None, G, prec=prec)
            got = T().decrypt_dev(out, ks.lwe_key, 2 * G)
            x, y, z = (bits[pm] for pm in perms)
            want = np.concatenate([np.where(x[:G] == 1, y[:G], z[:G]), x[G:] ^ y[G:]])
            fails += int((got != want).sum())
            bits = got
            cur, out = out, cur
    finally:
        for buf in ins + [cur, out]:
            buf.free()
    assert 2 * G * levels >= 10**6
    assert fails == 0, (prec, fails)


def test_split_digits_fp16_n2048_bit_exact():
    """The N = 2048 split kernels' conversion of its 8-bit gadget digits to FP16 pairs (one byte
    permute, an or and one packed FP16 add) against the oracle's signed decomposition."""
    rng = np.random.default_rng(2052)
    x = np.concatenate([rand_torus(rng, 4096),
                        np.array([0, -1, 2**31 - 1, -2**31, 0x7F7F7F00, -0x7F7F7F00, 0x80808000 - 2**32, 0x7FFFFF],
                                 np.int64).astype(np.int32)])
    x = np.concatenate([x, x[::-1]]).reshape(-1, 2)
    got = T().split_digits(x, P.Bgbit)
    for h in range(2):
        want = O.decompose(x[:, h], P.l, P.Bgbit)
        assert np.array_equal(got[:, :, h].T, want.astype(np.float32)), h


def test_extern_product_n2048_fp64_exact_sweep(env2048):
    """The N = 2048 FP64 external product bit-exact with the exact int64 product over 32 products
    (rows spread over the key; random, extreme-digit and int32-end accumulators)."""
    _, dk, keys = env2048
    rng = np.random.default_rng(2053)
    unit = sum(1 << (32 - (p + 1) * P.Bgbit) for p in range(P.l))
    lo, hi = -(1 << (P.Bgbit - 1)) * unit, ((1 << (P.Bgbit - 1)) - 1) * unit
    alt = np.where(np.arange(N) % 2 == 0, lo, hi)
    special = np.stack([np.full((2, N), lo), np.full((2, N), hi), np.stack([alt, alt[::-1]]),
                        np.full((2, N), 2**31 - 1), np.full((2, N), -2**31)]).astype(np.int64)
    acc = np.concatenate([rand_torus(rng, (27, 2, N)), ((special + 2**31) % 2**32 - 2**31).astype(np.int32)])
    rows = np.linspace(0, P.n - 1, acc.shape[0]).astype(np.int32)
    got, raw = dk.extern_product(acc, rows, prec=0, want_raw=True)
    worst = 0.0
    for g in range(acc.shape[0]):
        want, exact = O.extern_product_exact(P, acc[g], keys.bk[rows[g]], want_raw=True)
        assert np.array_equal(got[g], want), (g, int(rows[g]))
        worst = max(worst, float(np.abs(raw[g] - exact.astype(np.float64)).max()))
    print(f"N=2048 FP64 external product sweep ({acc.shape[0]} products): max |pre-rounding - exact| = {worst:.3e}")
    assert worst < 0.125
