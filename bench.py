"""Benchmark: bootstrapped NAND gates/s on one MI355X, split-FP32 vs FP64 external product.

Workload (one "step"): 16384 independent NAND gates, each a full gate bootstrap (pre-combination,
modulus switch, n = 630 MuxRotate steps, extraction) plus keyswitch, on ciphertexts already
resident in HBM.  Parameters N=1024, k=1, n=630, l=3, Bgbit=7, keyswitch t=8, basebit=2; keys from
seed 0; uniformly random input bits (numpy PCG64, seed 1).

Prints one JSON line.  `value` is the split-FP32 rate; the FP64 rate, their ratio, the chained
10^6-gate decryption audit, the roofline of the bootstrap kernel (timed with HIP events on the
library stream) and the CPU baseline (the oracle's FP64 FFT path on the host cores) ride along,
with two secondary workloads: 4096 independent 32-bit ripple-carry adders evaluated as one
levelised circuit (graph-replayed), and 8192 MUX + 8192 XOR gates at the N=2048 parameter set.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fpga-tfhe_amd"))

METRIC = "bootstrapped NAND gates/sec on 1 MI355X, split-FP32 vs FP64 at 0 fails/1e6"
N, n, k, l, Bgbit = 1024, 630, 1, 3, 7
M = N // 2
# algorithmic work per gate: n * [8 transforms * (5 M log2 M + 6 M) + (k+1) l (k+1) M complex MACs * 8]
FLOP_PER_GATE = n * (8 * (5 * M * 9 + 6 * M) + (k + 1) * l * (k + 1) * M * 8)   # 162,570,240
PEAK_TFLOPS = {"fp64": 78.6, "split32": 157.3}  # MI355X published vector/matrix peaks
FORM_NAMES = {0: "size rule", 1: "key-sharing workgroups (8 bootstraps)", 2: "one bootstrap per workgroup",
              3: "latency form", 4: "2 bootstraps per workgroup", 5: "4 bootstraps per workgroup"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--audit-gates", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--host-reps", type=int, default=3,
                    help="timed calls of the host-array entry point (PCIe-inclusive rate; 0: skip)")
    ap.add_argument("--prec", choices=["both", "split32", "fp64"], default="both")
    ap.add_argument("--adders", type=int, default=4096, help="32-bit adders per circuit run (0: skip)")
    ap.add_argument("--n2048-gates", type=int, default=8192,
                    help="MUX and XOR gates each at the N=2048 set (0: skip)")
    ap.add_argument("--split-variants", action="store_true",
                    help="also measure the split-FP32 forms (24/16-bit split x plain/compensated accumulation)")
    ap.add_argument("--fused-ks-ab", action="store_true",
                    help="also time the keyswitch fused into the split bootstrap kernel against the separate kernel")
    ap.add_argument("--form-sweep", action="store_true",
                    help="also time every bootstrap launch form on the headline and N=2048 workloads")
    ap.add_argument("--launch-form", type=int, default=0,
                    help="force a bootstrap launch form for every launch (0: the size rules; see tfhe_gpu_set_launch_form)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib):
    """Reference kernel of known traffic for the FETCH_SIZE/WRITE_SIZE counters: a device-to-device
    copy moves exactly `mib` MiB in and out per dispatch."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"calibrate_copy_bytes": mib << 20, "dispatches": 3}), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(T, dk, prec, bufs, batch, steps, warmup):
    d_out, d_a, d_b = bufs
    name = "bootstrap_fp64" if prec == T.PREC_FP64 else "bootstrap_split32"
    for _ in range(warmup):
        dk.gate_dev("NAND", d_out.ptr, d_a.ptr, d_b.ptr, None, batch, prec=prec)
    T.sync()
    T.profile_reset()
    T.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        dk.gate_dev("NAND", d_out.ptr, d_a.ptr, d_b.ptr, None, batch, prec=prec)
    T.sync()
    t1 = time.perf_counter()
    T.profile_enable(False)
    boot_ms, boot_n = T.profile_read(name)
    ks_ms, ks_n = T.profile_read("keyswitch")
    return {"ms_per_step": (t1 - t0) * 1e3 / steps, "gates_per_s": batch * steps / (t1 - t0),
            "boot_ms_avg": boot_ms / max(boot_n, 1), "ks_ms_avg": ks_ms / max(ks_n, 1)}


def audit(T, dk, ks, prec, batch, total, rng, lwe_n=None):
    """Chained NAND levels: each level's outputs are the next level's inputs (random pairing), so
    every gate sees bootstrapped-and-keyswitched noise.  Decryption on the GPU, plaintext on host.
    lwe_n: LWE dimension of the key set (the N=1024 set's by default)."""
    nn = n if lwe_n is None else lwe_n
    bits = rng.integers(0, 2, batch).astype(np.int32)
    cur = T.DeviceBuffer.from_array(ks.encrypt(bits))
    nxt_a = T.DeviceBuffer(cur.nbytes)
    nxt_b = T.DeviceBuffer(cur.nbytes)
    out = T.DeviceBuffer(cur.nbytes)
    done = fails = 0
    try:
        while done < total:
            host = cur.to_array((batch, nn + 1))
            perm = rng.permutation(batch)
            T.lib().tfhe_gpu_memcpy_h2d(nxt_b.ptr, np.ascontiguousarray(host[perm]).ctypes.data_as(T._vp),
                                        host.nbytes)
            dk.gate_dev("NAND", out.ptr, cur.ptr, nxt_b.ptr, None, batch, prec=prec)
            got = T.decrypt_dev(out, ks.lwe_key, batch)
            bits = 1 - (bits & bits[perm])
            bad = got != bits
            fails += int(bad.sum())
            bits = got  # follow the decrypted values so one failure is counted once
            cur, out = out, cur
            done += batch
    finally:
        for b in (cur, nxt_a, nxt_b, out):
            b.free()
    return done, fails


def precision_report(T, dk, rng, count=64):
    """Per-coefficient error of the split-FP32 external product.  The reference value is the FP64
    path's result, which is the exact int64 negacyclic product whenever its own pre-rounding
    values lie within 1/2 of an integer (checked here and reported as fp64_max_dist_to_integer;
    the parity tests hold it bit-exact against the int64 oracle)."""
    acc = rng.integers(-2**31, 2**31, (count, 2, N), dtype=np.int64).astype(np.int32)
    rows = rng.integers(0, n, count).astype(np.int32)
    r64, raw64 = dk.extern_product(acc, rows, prec=T.PREC_FP64, want_raw=True)
    r32 = dk.extern_product(acc, rows, prec=T.PREC_SPLIT32)
    dist = float(np.abs(raw64 - np.rint(raw64)).max())
    err = np.abs(((r32.astype(np.int64) - r64.astype(np.int64) + 2**31) % 2**32) - 2**31).ravel()
    hist = np.bincount(np.ceil(np.log2(err.astype(np.float64) + 1)).astype(int), minlength=1)
    return {"samples": int(err.size), "max": int(err.max()), "rms": round(float(np.sqrt((err.astype(float) ** 2).mean())), 1),
            "unit": "2^-32 torus", "log2_histogram": {f"<=2^{i}": int(c) for i, c in enumerate(hist) if c},
            "fp64_max_dist_to_integer": dist, "split_bits": dk.split_bits,
            "split_hi_rounding": hi_margin(dk, acc, rows)}


def hi_margin(dk, acc, rows):
    """The split path is exact while the digits x hi-key-half product, before rounding, stays within
    1/2 of its integer value (a slip would cost 2^split units at once): its largest and rms distance
    to the nearest integer over the sample."""
    _, hi = dk.extern_product_split_hi(acc, rows)
    d = np.abs(hi - np.rint(hi))
    return {"max_dist_to_integer": round(float(d.max()), 5), "rms_dist": round(float(np.sqrt((d ** 2).mean())), 5),
            "samples": int(d.size), "slip_at": 0.5}


def negacyclic_mul_binary(a, s):
    """a * s mod (X^N + 1) for an int64 polynomial a and a binary key s (exact in int64)."""
    Nn = a.shape[-1]
    c = np.convolve(a.astype(np.int64), s.astype(np.int64))
    out = c[:Nn].copy()
    out[:Nn - 1] -= c[Nn:]
    return out


def split_variants_report(T, ks, batch, steps, audit_gates, rng):
    """The split-FP32 external product in four forms: key halves split at 24 bits (hi 8 bits, the
    default) or at 16 bits (two 16-bit halves), each with plain FP32 FMA accumulation or with
    compensated accumulation (TwoProduct + TwoSum).  For each: gates/s on the NAND batch, the
    external-product error against the FP64 path (the exact int64 product; 64 products), the TLWE
    phase error after a full blind rotation (8 accumulators of a launch of 300, against the
    noiseless rotation), and the decryption failures of the chained NAND audit."""
    nn = ks.fp.n
    bits = rng.integers(0, 2, (2, batch)).astype(np.int32)
    d_a = T.DeviceBuffer.from_array(ks.encrypt(bits[0]))
    d_b = T.DeviceBuffer.from_array(ks.encrypt(bits[1]))
    d_out = T.DeviceBuffer(batch * (nn + 1) * 4)
    want = 1 - (bits[0] & bits[1])
    acc = rng.integers(-2**31, 2**31, (64, 2, N), dtype=np.int64).astype(np.int32)
    rows = rng.integers(0, nn, 64).astype(np.int32)
    cnt = 300
    acc0 = np.zeros((cnt, 2, N), np.int32)
    acc0[:, 1, :] = 1 << 29
    bara = rng.integers(0, 2 * N, (cnt, nn)).astype(np.int32)
    skey = ks.tlwe_key[:N].astype(np.int64)
    out = {}
    try:
        ref64 = None
        for bits_split, mac in ((24, 0), (24, 1), (16, 0), (16, 1)):
            name = f"split{bits_split}_{'compensated' if mac else 'plain'}"
            T.set_split_bits(bits_split)
            try:
                dk = T.DeviceKey.from_keyset(ks)
            finally:
                T.set_split_bits(0)
            T.set_split_mac(mac)
            try:
                if ref64 is None:
                    ref64 = dk.extern_product(acc, rows, prec=T.PREC_FP64)
                r = time_path(T, dk, T.PREC_SPLIT32, (d_out, d_a, d_b), batch, steps, 1)
                r["correct"] = bool(np.array_equal(T.decrypt_dev(d_out, ks.lwe_key, batch), want))
                r32 = dk.extern_product(acc, rows, prec=T.PREC_SPLIT32)
                err = np.abs(((r32.astype(np.int64) - ref64.astype(np.int64) + 2**31) % 2**32) - 2**31).ravel()
                hist = np.bincount(np.ceil(np.log2(err.astype(np.float64) + 1)).astype(int), minlength=1)
                rot = dk.blind_rotate(acc0, bara, prec=T.PREC_SPLIT32)
                perr = []
                for g in range(8):
                    ph = (rot[g, 1].astype(np.int64) - negacyclic_mul_binary(rot[g, 0], skey))
                    m = int((bara[g].astype(np.int64) * ks.lwe_key).sum() % (2 * N))
                    ideal = np.where((np.arange(N) - m) % (2 * N) < N, 1, -1).astype(np.int64) << 29
                    perr.append((ph - ideal + 2**31) % 2**32 - 2**31)
                pe = np.abs(np.concatenate(perr)).astype(np.float64)
                done, fails = audit(T, dk, ks, T.PREC_SPLIT32, batch, audit_gates, np.random.default_rng(2))
                out[name] = {"split_bits": bits_split, "accumulation": "compensated" if mac else "plain FMA",
                             "gates_per_s": round(r["gates_per_s"], 1), "bootstrap_kernel_ms": round(r["boot_ms_avg"], 3),
                             "correct_batch": r["correct"],
                             "extern_product_error": {"max": int(err.max()),
                                                      "rms": round(float(np.sqrt((err.astype(float) ** 2).mean())), 1),
                                                      "samples": int(err.size),
                                                      "log2_histogram": {f"<=2^{i}": int(c) for i, c in enumerate(hist) if c}},
                             "phase_error_after_blind_rotation": {"rms": round(float(np.sqrt((pe ** 2).mean())), 1),
                                                                  "max": int(pe.max()), "steps": nn,
                                                                  "samples": int(pe.size), "decision_margin": 2**28},
                             "audit": {"gates": done, "failures": fails, "chained": True},
                             "unit": "2^-32 torus"}
                log(f"[bench] split variant {name}: {out[name]}")
            finally:
                T.set_split_mac(0)
                dk.free()
    finally:
        for b_ in (d_a, d_b, d_out):
            b_.free()
    return out


def form_sweep_report(T, dk, bufs, batch, n2048_gates):
    """Every launch form that applies, forced in turn: the headline NAND step at N=1024 (FP64: key
    sharing, one bootstrap per workgroup; split: one per workgroup) and one MUX + XOR step at N=2048
    (1, 2 or 4 two-wavefront bootstraps per workgroup), median of 3 timed steps after a warm-up."""
    out = {"n1024": {}, "n2048": {}}
    d_out, d_a, d_b = bufs
    try:
        for name, prec, forms in (("fp64", T.PREC_FP64, (1, 2)), ("split32", T.PREC_SPLIT32, (2,))):
            for f in forms:
                T.set_launch_form(f)
                ts = []
                for rep in range(4):
                    T.sync()
                    t0 = time.perf_counter()
                    dk.gate_dev("NAND", d_out.ptr, d_a.ptr, d_b.ptr, None, batch, prec=prec)
                    T.sync()
                    ts.append((time.perf_counter() - t0) * 1e3)
                out["n1024"][f"{name} {FORM_NAMES[f]}"] = round(sorted(ts[1:])[1], 2)
        if n2048_gates > 0:
            ks2 = T.KeySet(2048, seed=(0,))
            dk2 = T.DeviceKey.from_keyset(ks2)
            G = n2048_gates
            rng = np.random.default_rng(9)
            bits = rng.integers(0, 2, (3, G)).astype(np.int32)
            bb = [T.DeviceBuffer.from_array(ks2.encrypt(bits[i])) for i in range(3)]
            oo = [T.DeviceBuffer(G * (ks2.fp.n + 1) * 4) for _ in range(2)]
            try:
                for name, prec in (("fp64", T.PREC_FP64), ("split32", T.PREC_SPLIT32)):
                    for f in (2, 4, 5):
                        T.set_launch_form(f)
                        ts = []
                        for rep in range(4):
                            T.sync()
                            t0 = time.perf_counter()
                            dk2.gate_dev("MUX", oo[0].ptr, bb[0].ptr, bb[1].ptr, bb[2].ptr, G, prec=prec)
                            dk2.gate_dev("XOR", oo[1].ptr, bb[0].ptr, bb[1].ptr, None, G, prec=prec)
                            T.sync()
                            ts.append((time.perf_counter() - t0) * 1e3)
                        ok = bool(np.array_equal(T.decrypt_dev(oo[0], ks2.lwe_key, G), np.where(bits[0] == 1, bits[1], bits[2])))
                        out["n2048"][f"{name} {FORM_NAMES[f]}"] = {"ms_per_step": round(sorted(ts[1:])[1], 2), "correct": ok}
            finally:
                for b_ in bb + oo:
                    b_.free()
                dk2.free()
                ks2.free()
    finally:
        T.set_launch_form(0)
    log(f"[bench] launch-form sweep: {out}")
    return out


def fused_ks_report(T, dk, bufs, batch, steps, want, ks):
    """Same-box A/B of the headline NAND step on the split path: bootstrap kernel + separate
    keyswitch kernel against the keyswitch fused into the bootstrap kernel, alternating A, B, A, B;
    the two outputs must be bit-identical."""
    d_out, d_a, d_b = bufs
    nn = ks.fp.n
    res = {"separate": [], "fused": []}
    outs = {}
    try:
        for rep in range(2):
            for name, on in (("separate", False), ("fused", True)):
                T.set_fused_keyswitch(on)
                r = time_path(T, dk, T.PREC_SPLIT32, bufs, batch, steps, 1)
                res[name].append(round(r["ms_per_step"], 3))
                outs[name] = d_out.to_array((batch, nn + 1))
    finally:
        T.set_fused_keyswitch(False)
    return {"ms_per_step": res, "bit_identical": bool(np.array_equal(outs["separate"], outs["fused"])),
            "fused_correct": bool(np.array_equal(ks.decrypt(outs["fused"]), want)),
            "workload": f"{batch} NAND, split-FP32, N=1024"}


def adders_report(T, dk, ks, count, reps=2):
    """`count` independent 32-bit ripple-carry adders as one circuit over a batch of `count`
    instances: 33 bootstrapped levels (one of 32 XOR, then MUX + XOR along the carry chain), each
    one bootstrap launch and at most two keyswitch launches, replayed from a HIP graph with the FP64
    launch form autotuned per level width.  Operands from numpy PCG64 seed 3; sums checked against
    plain addition after GPU decryption."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 2**32, count, dtype=np.uint64)
    y = rng.integers(0, 2**32, count, dtype=np.uint64)
    bits = np.array([[(int(v) >> i) & 1 for v in vals] for vals in (x, y) for i in range(32)], np.int32)
    t = time.perf_counter()
    ct = np.stack([ks.encrypt(bits[i]) for i in range(64)])
    log(f"[bench] adders: encrypted {bits.size} input bits in {time.perf_counter() - t:.1f} s")
    c = T.ripple_carry_adder(32)
    info = c.info()
    d_in = T.DeviceBuffer.from_array(ct)
    del ct
    d_out = T.DeviceBuffer(33 * count * (n + 1) * 4)
    res = {"workload": f"{count} independent 32-bit ripple-carry adders (p=XOR(a,b), s=XOR(p,c), "
                       f"c'=MUX(p,c,a)), one circuit run over a batch of {count}",
           "levels": info["levels"], "bootstraps_per_adder": info["bootstraps"]}
    try:
        for name, prec in (("split32", T.PREC_SPLIT32), ("fp64", T.PREC_FP64)):
            c.run_dev(dk, d_in, d_out, count, prec=prec, graph=True, autotune=True)   # tune + capture
            T.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                c.run_dev(dk, d_in, d_out, count, prec=prec, graph=True, autotune=True)
            T.sync()
            sec = (time.perf_counter() - t0) / reps
            dec = np.stack([T.decrypt_dev(d_out.ptr.value + i * count * (n + 1) * 4, ks.lwe_key, count)
                            for i in range(33)])
            s_ = [int(sum(int(dec[i][j]) << i for i in range(32))) for j in range(count)]
            ok = s_ == [int((a + b) % 2**32) for a, b in zip(x, y)] and \
                list(dec[32]) == [int((a + b) >> 32) for a, b in zip(x, y)]
            forms = {}
            for nb, form in c.level_plan():
                forms[str(nb * count)] = FORM_NAMES.get(form, str(form))
            res[name] = {"ms_per_run": round(sec * 1e3, 2), "adders_per_s": round(count / sec, 1),
                         "bootstraps_per_s": round(count * info["bootstraps"] / sec, 1), "correct": bool(ok),
                         "autotuned_form_by_level_bootstraps": forms}
            log(f"[bench] adders {name}: {res[name]}")
    finally:
        d_in.free()
        d_out.free()
        c.free()
    return res


def n2048_report(T, gates, steps=2, audit_gates=1_000_000):
    """N=2048 set (n=750, l=3, Bgbit=8, ks t=10): `gates` MUX and `gates` XOR on random bits per
    step (one launch each), both paths; decryption failures counted over every step and over a
    chained audit of at least `audit_gates` NANDs per path (levels of 16384, as at N=1024), and the
    split path's external-product error against the FP64 path."""
    t = time.perf_counter()
    ks = T.KeySet(2048, seed=(0,))
    dk = T.DeviceKey.from_keyset(ks)
    log(f"[bench] N=2048 keys ready in {time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(5)
    nn = ks.fp.n
    bits = rng.integers(0, 2, (3, gates)).astype(np.int32)
    bufs = [T.DeviceBuffer.from_array(ks.encrypt(bits[i])) for i in range(3)]
    outs = [T.DeviceBuffer(gates * (nn + 1) * 4) for _ in range(2)]
    want_mux = np.where(bits[0] == 1, bits[1], bits[2])
    want_xor = bits[0] ^ bits[1]
    res = {"workload": f"{gates} MUX + {gates} XOR per step, N=2048 n=750 l=3 Bgbit=8 ks t=10",
           "bootstraps_per_step": 3 * gates}
    try:
        for name, prec in (("split32", T.PREC_SPLIT32), ("fp64", T.PREC_FP64)):
            fails = 0
            dk.gate_dev("MUX", outs[0].ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, gates, prec=prec)
            T.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                dk.gate_dev("MUX", outs[0].ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, gates, prec=prec)
                dk.gate_dev("XOR", outs[1].ptr, bufs[0].ptr, bufs[1].ptr, None, gates, prec=prec)
            T.sync()
            sec = (time.perf_counter() - t0) / steps
            fails += int((T.decrypt_dev(outs[0], ks.lwe_key, gates) != want_mux).sum())
            fails += int((T.decrypt_dev(outs[1], ks.lwe_key, gates) != want_xor).sum())
            res[name] = {"ms_per_step": round(sec * 1e3, 2), "gates_per_s": round(2 * gates / sec, 1),
                         "bootstraps_per_s": round(3 * gates / sec, 1), "failures": fails,
                         "gates_checked": 2 * gates}
            # one NAND per call from host arrays (the upstream single-gate API's shape): the
            # small-launch forms and the copies, end to end
            a1, b1 = ks.encrypt(bits[0][:1]), ks.encrypt(bits[1][:1])
            dk.gate("NAND", a1, b1, prec=prec)
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                o1 = dk.gate("NAND", a1, b1, prec=prec)
            res[name]["single_gate_ms"] = round((time.perf_counter() - t0) / reps * 1e3, 3)
            res[name]["single_gate_correct"] = bool(ks.decrypt(o1)[0] == 1 - (bits[0][0] & bits[1][0]))
            log(f"[bench] N=2048 {name}: {res[name]}")
        # chained audit: levels of 16384 NANDs whose inputs are the previous level's outputs
        # (bootstrapped noise), every output decrypted on the GPU
        if audit_gates > 0:
            for name, prec in (("split32", T.PREC_SPLIT32), ("fp64", T.PREC_FP64)):
                t = time.perf_counter()
                done, fails = audit(T, dk, ks, prec, 16384, audit_gates, np.random.default_rng(11), lwe_n=nn)
                res[name]["audit"] = {"gates": done, "failures": fails, "chained": True}
                log(f"[bench] N=2048 audit {name}: {res[name]['audit']} in {time.perf_counter() - t:.1f} s")
        acc = rng.integers(-2**31, 2**31, (32, 2, 2048), dtype=np.int64).astype(np.int32)
        rows = rng.integers(0, nn, 32).astype(np.int32)
        r64, raw64 = dk.extern_product(acc, rows, prec=T.PREC_FP64, want_raw=True)
        r32 = dk.extern_product(acc, rows, prec=T.PREC_SPLIT32)
        err = np.abs(((r32.astype(np.int64) - r64.astype(np.int64) + 2**31) % 2**32) - 2**31)
        res["split32_extern_product_error"] = {"samples": int(err.size), "max": int(err.max()),
                                               "rms": round(float(np.sqrt((err.astype(float) ** 2).mean())), 1),
                                               "unit": "2^-32 torus",
                                               "fp64_max_dist_to_integer": float(np.abs(raw64 - np.rint(raw64)).max()),
                                               "split_bits": dk.split_bits,
                                               "split_hi_rounding": hi_margin(dk, acc, rows)}
    finally:
        for b in bufs + outs:
            b.free()
        dk.free()
        ks.free()
    return res


def host_cpu():
    """Host CPU as the box presents it: logical CPUs, the CPUs this process may run on, the OpenMP
    thread limit set in the environment, and the model name (lscpu, else /proc/cpuinfo)."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    omp = os.environ.get("OMP_NUM_THREADS", "")
    info["omp_num_threads"] = int(omp) if omp.isdigit() and int(omp) > 0 else None
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    if not model:
        try:
            model = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")][0]
        except Exception:
            model = "unknown"
    info["model"] = model
    return info


def cpu_baseline(ks, batch_bits, seconds):
    """The oracle's FP64 FFT gate path (bootstrap + keyswitch per gate, one gate per OpenMP thread)
    on every host CPU this process may use -- the affinity set, capped by OMP_NUM_THREADS where the
    environment sets it -- and on one thread, over bounded samples of the same NAND workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    keys = O.Keys(O.PARAMS_1024, ks.lwe_key, ks.tlwe_key, ks.ksk, ks.bk)
    host = host_cpu()
    threads = host["affinity"]
    if host["omp_num_threads"]:
        threads = min(threads, host["omp_num_threads"])
    threads = max(1, threads)
    try:
        cal = 2 * threads
        ca, cb = ks.encrypt(batch_bits[0][:cal]), ks.encrypt(batch_bits[1][:cal])
        _, sec = O.gate_batch(keys, "NAND", ca, cb, nthreads=threads)
        cnt = int(max(cal, min(20000, seconds * cal / max(sec, 1e-6))))
        cnt = (cnt // threads) * threads or threads
        ca, cb = ks.encrypt(batch_bits[0][:cnt]), ks.encrypt(batch_bits[1][:cnt])
        out, sec = O.gate_batch(keys, "NAND", ca, cb, nthreads=threads)
        ok = bool(np.array_equal(ks.decrypt(out), 1 - (batch_bits[0][:cnt] & batch_bits[1][:cnt])))
        # the same path on one core, over a sample of about a fifth of the time
        one = max(1, int(seconds / 5 * cnt / sec / threads))
        out1, sec1 = O.gate_batch(keys, "NAND", ca[:one], cb[:one], nthreads=1)
        ok = ok and bool(np.array_equal(out1, out[:one]))
        return {"value": cnt / sec, "unit": "gates/s", "cores": threads, "kind": "port",
                "sample": f"{cnt} NAND gates (bootstrap + keyswitch, FP64 FFT path), {sec:.1f} s on {threads} threads, "
                          f"{host['model']}",
                "host": host,
                "single_thread": {"value": one / sec1, "unit": "gates/s", "sample": f"{one} of the same gates, {sec1:.1f} s"},
                "correct": ok}
    finally:
        keys.free()


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import tfhe_amd as T
    T.init(0)
    if args.launch_form:
        T.set_launch_form(args.launch_form)   # before the keys are made: form 6 needs its key layout
    dev = T.device_info()
    log(f"[bench] device {dev}")
    t = time.perf_counter()
    ks = T.KeySet(1024, seed=(0,))
    dk = T.DeviceKey.from_keyset(ks)
    log(f"[bench] keys ready in {time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(1)
    B = args.batch
    bits = rng.integers(0, 2, (2, B)).astype(np.int32)
    ca, cb = ks.encrypt(bits[0]), ks.encrypt(bits[1])
    d_a, d_b = T.DeviceBuffer.from_array(ca), T.DeviceBuffer.from_array(cb)
    d_out = T.DeviceBuffer(ca.nbytes)
    want = 1 - (bits[0] & bits[1])

    precs = {"split32": T.PREC_SPLIT32, "fp64": T.PREC_FP64}
    if args.prec != "both":
        precs = {args.prec: precs[args.prec]}
    res = {}
    for name, prec in precs.items():
        r = time_path(T, dk, prec, (d_out, d_a, d_b), B, args.steps, args.warmup)
        r["correct"] = bool(np.array_equal(T.decrypt_dev(d_out, ks.lwe_key, B), want))
        res[name] = r
        log(f"[bench] {name}: {r}")

    head = "split32" if "split32" in res else "fp64"
    # the same batch through the host-array entry point (tfhe_gpu_gate_batch): the ciphertexts are
    # copied in from pageable host memory and the results copied back inside each timed call
    host = None
    if args.host_reps > 0:
        out_h = dk.gate("NAND", ca, cb, prec=precs[head])
        t0 = time.perf_counter()
        for _ in range(args.host_reps):
            out_h = dk.gate("NAND", ca, cb, prec=precs[head])
        sec = (time.perf_counter() - t0) / args.host_reps
        host = {"gates_per_s": round(B / sec, 1), "ms_per_call": round(sec * 1e3, 3),
                "bytes_in": int(ca.nbytes + cb.nbytes), "bytes_out": int(out_h.nbytes),
                "correct": bool(np.array_equal(ks.decrypt(out_h), want)),
                "entry": "tfhe_gpu_gate_batch (pageable host arrays, copies inside the timed call)"}
        log(f"[bench] host-buffer {head}: {host}")
    # one gate per call through the same entry point (the upstream single-gate API's shape): the
    # small-launch forms (FP64 latency form, sliced keyswitch) and the copies, end to end
    single = {}
    if args.host_reps > 0:
        for name, prec in precs.items():
            dk.gate("NAND", ca[:1], cb[:1], prec=prec)
            t0 = time.perf_counter()
            for _ in range(args.host_reps):
                o1 = dk.gate("NAND", ca[:1], cb[:1], prec=prec)
            sec = (time.perf_counter() - t0) / args.host_reps
            single[name] = {"ms_per_gate": round(sec * 1e3, 3),
                            "correct": bool(ks.decrypt(o1)[0] == want[0])}
        log(f"[bench] single gate: {single}")
    aud = {}
    if args.audit_gates > 0:
        for name, prec in precs.items():
            t = time.perf_counter()
            done, fails = audit(T, dk, ks, prec, B, args.audit_gates, np.random.default_rng(2))
            aud[name] = {"gates": done, "failures": fails, "chained": True}
            log(f"[bench] audit {name} {aud[name]} in {time.perf_counter() - t:.1f} s")
    prec_rep = precision_report(T, dk, np.random.default_rng(3)) if len(precs) == 2 else None
    log(f"[bench] split32 external-product error {prec_rep}")

    fused = (fused_ks_report(T, dk, (d_out, d_a, d_b), B, max(2, args.steps // 2), want, ks)
             if args.fused_ks_ab else None)
    if fused:
        log(f"[bench] fused keyswitch A/B: {fused}")
    sweep = form_sweep_report(T, dk, (d_out, d_a, d_b), B, args.n2048_gates) if args.form_sweep else None
    variants = (split_variants_report(T, ks, B, max(2, args.steps // 2), args.audit_gates, np.random.default_rng(4))
                if args.split_variants else None)
    adders = adders_report(T, dk, ks, args.adders) if args.adders > 0 and args.prec == "both" else None
    n2048 = (n2048_report(T, args.n2048_gates, audit_gates=args.audit_gates)
             if args.n2048_gates > 0 and args.prec == "both" else None)

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(ks, bits, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu} in {time.perf_counter() - t:.1f} s")

    vp64, vp32 = T.measure_valu_peak()
    log(f"[bench] measured dense-FMA VALU peak: fp64 {vp64:.1f} TFLOP/s, fp32 {vp32:.1f} TFLOP/s")
    hr = res[head]
    achieved = FLOP_PER_GATE * B / (hr["boot_ms_avg"] * 1e-3) / 1e12
    traffic = None
    # every bootstrap streams the whole FFT-domain key (61.9 MB at N=1024) once: the demand on the
    # cache hierarchy, measured live; the profile adds what L2 and HBM actually served
    key_bytes = n * (k + 1) * l * (k + 1) * M * 16
    key_stream = {"bytes_per_bootstrap": key_bytes, "bytes_per_launch": key_bytes * B,
                  "demand_gbps": round(key_bytes * B / (hr["boot_ms_avg"] * 1e-3) / 1e9, 1)}
    tpath = os.path.join(ROOT, "profiles", f"traffic_bootstrap_{head}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        key_stream["profiled"] = {kk: prof.get(kk) for kk in ("l2_to_l1_gbps", "l2_hit_rate", "hbm_gbps", "avg_ms", "source")}
    line = {
        "metric": METRIC,
        "value": round(hr["gates_per_s"], 1),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "split32" else "f64",
        "data": "synthetic: uniform random bits (PCG64 seed 1) encrypted under keys from seed 0",
        "config": {"workload": f"{B} independent NAND gate bootstraps + keyswitch per step, "
                               f"{'split-FP32' if head == 'split32' else 'FP64'} external product",
                   "N": N, "k": k, "n": n, "l": l, "Bgbit": Bgbit, "ks_t": 8, "ks_basebit": 2, "batch": B},
        "paths": {kk: {"gates_per_s": round(v["gates_per_s"], 1), "ms_per_step": round(v["ms_per_step"], 3),
                       "bootstrap_kernel_ms": round(v["boot_ms_avg"], 3), "keyswitch_kernel_ms": round(v["ks_ms_avg"], 3),
                       "correct": v["correct"]} for kk, v in res.items()},
        "split32_over_fp64": (round(res["split32"]["gates_per_s"] / res["fp64"]["gates_per_s"], 3)
                              if len(res) == 2 else None),
        "host_buffers": host,
        "single_gate_latency": single or None,
        "audit": aud or None,
        "split32_extern_product_error": prec_rep,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS[head], 4),
                     "traffic": traffic, "kernel": f"bootstrap_{head}", "flop_per_gate": FLOP_PER_GATE,
                     "bound_note": "vector-ALU FLOP roofline: the transforms and the pointwise MAC run on the VALU "
                                   "(the split path's digit transforms take their first 8-point pass on the matrix "
                                   "cores); peak is the published FP32 (split path) or FP64 vector rate",
                     "measured_valu_peak": {"fp64": round(vp64, 1), "fp32_unpacked": round(vp32, 1)},
                     "key_stream": key_stream},
        "cpu_baseline": cpu,
        "split32_variants": variants,
        "fused_keyswitch_ab": fused,
        "launch_form_sweep": sweep,
        "circuits_32bit_adders": adders,
        "n2048": n2048,
        "device": dev,
    }
    print(json.dumps(line), flush=True)
    for b in (d_a, d_b, d_out):
        b.free()
    dk.free()
    ks.free()


if __name__ == "__main__":
    main()
