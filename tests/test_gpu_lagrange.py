"""GPU: device keys from FFT-domain bootstrapping keys in upstream order (LagrangeHalfC values in
torus units), through the flat entry point and through an upstream-shaped cloud key set whose
bkFFT polynomials carry their values (bk == NULL).  Both recover the exact time-domain key, so
every gate output equals the one made with the time-domain key, bit for bit, on both paths."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


class Lagrange(ctypes.Structure):            # LagrangeHalfCPolynomial {data, precomp}
    _fields_ = [("data", vp), ("precomp", vp)]


class TLweSampleFFT(ctypes.Structure):
    _fields_ = [("a", ctypes.POINTER(Lagrange)), ("b", ctypes.POINTER(Lagrange)),
                ("current_variance", ctypes.c_double), ("k", ctypes.c_int32)]


class TGswSampleFFT(ctypes.Structure):
    _fields_ = [("all_samples", ctypes.POINTER(TLweSampleFFT)), ("sample", vp),
                ("k", ctypes.c_int32), ("l", ctypes.c_int32)]


class LweBootstrappingKeyFFT(ctypes.Structure):
    _fields_ = [("in_out_params", vp), ("bk_params", vp), ("accum_params", vp), ("extract_params", vp),
                ("bkFFT", ctypes.POINTER(TGswSampleFFT)), ("ks", vp)]


class CloudKeySet(ctypes.Structure):
    _fields_ = [("params", vp), ("bk", vp), ("bkFFT", ctypes.POINTER(LweBootstrappingKeyFFT))]


def T():
    import tfhe_amd
    return tfhe_amd


def gates_equal(ks, dk_ref, dk_new, prec):
    rng = np.random.default_rng(70 + prec)
    bits = rng.integers(0, 2, (3, 40)).astype(np.int32)
    ca, cb, cc = (ks.encrypt(bits[i]) for i in range(3))
    for gate in ("NAND", "XOR", "MUX"):
        c = cc if gate == "MUX" else None
        want = dk_ref.gate(gate, ca, cb, c, prec=prec)
        got = dk_new.gate(gate, ca, cb, c, prec=prec)
        assert np.array_equal(got, want), (gate, prec)


def test_key_from_flat_lagrange_values(gpu_env):
    ks, dk, _ = gpu_env
    lag = T().bk_to_lagrange(ks.fp, ks.bk)
    print("[lagrange] key from flat FFT-domain values", flush=True)
    dk2 = T().DeviceKey.from_lagrange(ks.fp, lag, ks.ksk)
    try:
        for prec in (0, 1):
            gates_equal(ks, dk, dk2, prec)
    finally:
        dk2.free()


def test_key_upload_of_upstream_shaped_cloud_key(gpu_env):
    """The library's own cloud key set supplies the upstream-shaped bkFFT structure; its
    polynomials are pointed at FFT-domain values, and a cloud key set with bk = NULL is uploaded."""
    ks, dk, _ = gpu_env
    L = T().lib()
    own = ctypes.cast(ks.cloud_ptr, ctypes.POINTER(CloudKeySet))[0]
    bkfft = own.bkFFT[0]
    n, N, k, l = ks.fp.n, ks.fp.N, ks.fp.k, ks.fp.l
    kpl = (k + 1) * l
    lag = np.ascontiguousarray(T().bk_to_lagrange(ks.fp, ks.bk))   # [n][kpl][k+1][N/2]
    polys = [bkfft.bkFFT[i].all_samples[r].a[q] for i in range(n) for r in range(kpl) for q in range(k + 1)]
    base = lag.ctypes.data
    try:
        for idx, poly in enumerate(polys):
            poly.data = base + idx * (N // 2) * 16
        upstream = CloudKeySet(own.params, None, own.bkFFT)
        key = ctypes.c_void_p()
        print("[lagrange] upload of a cloud key set with bk = NULL", flush=True)
        rc = L.tfhe_gpu_key_upload(ctypes.byref(upstream), 3, ctypes.byref(key))
        assert rc == 0, L.tfhe_gpu_last_error()
        dk2 = T().DeviceKey.__new__(T().DeviceKey)
        dk2.fp, dk2.ptr = ks.fp, key
        try:
            for prec in (0, 1):
                gates_equal(ks, dk, dk2, prec)
        finally:
            dk2.free()          # also drops the library's record of the cloud key set
    finally:
        for poly in polys:
            poly.data = None
