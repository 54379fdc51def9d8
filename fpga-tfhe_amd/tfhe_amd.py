"""Python binding of libtfhe_mi355x (ctypes over the C ABI in include/tfhe_mi355x.h).

The library is the product: HIP kernels for gfx950 behind the upstream tfhe C API and a flat
batched ABI.  This module only marshals numpy arrays and device pointers; it never falls back
to a CPU implementation -- a missing library or device raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TFHE_MI355X_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("TFHE_MI355X_LIB") or os.path.join(_HERE, "lib", "libtfhe_mi355x.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "tfhe_mi355x.h")

PREC_FP64, PREC_SPLIT32 = 0, 1
PREC_MASK_FP64, PREC_MASK_SPLIT32 = 1, 2
GATES = {"NAND": 0, "AND": 1, "OR": 2, "XOR": 3, "XNOR": 4, "NOR": 5,
         "ANDNY": 6, "ANDYN": 7, "ORNY": 8, "ORYN": 9, "MUX": 10}
MU = 1 << 29  # modSwitchToTorus32(1, 8)
OP_NOT, OP_COPY = 32, 33
CIRCUIT_GRAPH, CIRCUIT_AUTOTUNE = 1, 2


class TfheError(RuntimeError):
    pass


class FlatParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("N", ctypes.c_int32), ("k", ctypes.c_int32),
                ("l", ctypes.c_int32), ("Bgbit", ctypes.c_int32), ("ks_t", ctypes.c_int32),
                ("ks_basebit", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LweSample(ctypes.Structure):
    _fields_ = [("a", ctypes.POINTER(ctypes.c_int32)), ("b", ctypes.c_int32),
                ("current_variance", ctypes.c_double)]


_lib = None
P = ctypes.POINTER
_i32p, _f64p, _vp = P(ctypes.c_int32), P(ctypes.c_double), ctypes.c_void_p


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TfheError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i32 = ctypes.c_int32
    sig = {
        # upstream API
        "tfhe_random_generator_setSeed": (None, [P(ctypes.c_uint32), i32]),
        "new_default_gate_bootstrapping_parameters": (_vp, [i32]),
        "new_gate_bootstrapping_parameters_N2048": (_vp, []),
        "delete_gate_bootstrapping_parameters": (None, [_vp]),
        "new_random_gate_bootstrapping_secret_keyset": (_vp, [_vp]),
        "delete_gate_bootstrapping_secret_keyset": (None, [_vp]),
        "new_gate_bootstrapping_ciphertext_array": (P(LweSample), [i32, _vp]),
        "delete_gate_bootstrapping_ciphertext_array": (None, [i32, P(LweSample)]),
        "bootsSymEncrypt": (None, [P(LweSample), i32, _vp]),
        "bootsSymDecrypt": (i32, [P(LweSample), _vp]),
        "bootsCONSTANT": (None, [P(LweSample), i32, _vp]),
        "bootsNOT": (None, [P(LweSample), P(LweSample), _vp]),
        "bootsMUX": (None, [P(LweSample), P(LweSample), P(LweSample), P(LweSample), _vp]),
        "modSwitchToTorus32": (i32, [i32, i32]),
        "modSwitchFromTorus32": (i32, [i32, i32]),
        "approxPhase": (i32, [i32, i32]),
        "dtot32": (i32, [ctypes.c_double]),
        "tfhe_keyset_export_flat": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p]),
        "tfhe_flat_params_of": (ctypes.c_int, [_vp, P(FlatParams)]),
        "tfhe_cloud_export_flat": (ctypes.c_int, [_vp, _i32p, _i32p]),
        "export_tfheGateBootstrappingParameterSet_toFile": (None, [_vp, _vp]),
        "new_tfheGateBootstrappingParameterSet_fromFile": (_vp, [_vp]),
        "export_tfheGateBootstrappingSecretKeySet_toFile": (None, [_vp, _vp]),
        "new_tfheGateBootstrappingSecretKeySet_fromFile": (_vp, [_vp]),
        "export_tfheGateBootstrappingCloudKeySet_toFile": (None, [_vp, _vp]),
        "new_tfheGateBootstrappingCloudKeySet_fromFile": (_vp, [_vp]),
        "export_gate_bootstrapping_ciphertext_toFile": (None, [_vp, P(LweSample), _vp]),
        "import_gate_bootstrapping_ciphertext_fromFile": (None, [_vp, P(LweSample), _vp]),
        "delete_gate_bootstrapping_cloud_keyset": (None, [_vp]),
        "tfhe_encrypt_bits_flat": (ctypes.c_int, [_vp, _i32p, i32, _i32p]),
        "tfhe_decrypt_bits_flat": (ctypes.c_int, [_vp, _i32p, i32, _i32p]),
        # flat ABI
        "tfhe_gpu_last_error": (ctypes.c_char_p, []),
        "tfhe_gpu_init": (ctypes.c_int, [i32]),
        "tfhe_gpu_device_info": (ctypes.c_int, [ctypes.c_char_p, i32, P(i32), P(ctypes.c_int64)]),
        "tfhe_gpu_sync": (ctypes.c_int, []),
        "tfhe_gpu_set_split_bits": (ctypes.c_int, [i32]),
        "tfhe_gpu_set_split_mac": (ctypes.c_int, [i32]),
        "tfhe_gpu_set_launch_form": (ctypes.c_int, [i32]),
        "tfhe_gpu_set_fused_keyswitch": (ctypes.c_int, [i32]),
        "tfhe_circuit_level_plan": (ctypes.c_int, [_vp, i32, P(i32), P(i32)]),
        "tfhe_gpu_set_default_precision": (ctypes.c_int, [i32]),
        "tfhe_gpu_set_fp64_shared_rows": (ctypes.c_int, [i32]),
        "tfhe_gpu_key_create": (ctypes.c_int, [P(FlatParams), _i32p, _i32p, i32, P(_vp)]),
        "tfhe_gpu_key_upload": (ctypes.c_int, [_vp, i32, P(_vp)]),
        "tfhe_bk_from_lagrange": (ctypes.c_int, [P(FlatParams), P(ctypes.c_double), _i32p]),
        "tfhe_bk_to_lagrange": (ctypes.c_int, [P(FlatParams), _i32p, P(ctypes.c_double)]),
        "tfhe_gpu_key_create_lagrange": (ctypes.c_int, [P(FlatParams), P(ctypes.c_double), _i32p, i32, P(_vp)]),
        "tfhe_gpu_key_free": (ctypes.c_int, [_vp]),
        "tfhe_gpu_key_params": (ctypes.c_int, [_vp, P(FlatParams)]),
        "tfhe_gpu_gate_batch": (ctypes.c_int, [_vp, i32, i32, _i32p, _i32p, _i32p, _i32p, i32]),
        "tfhe_gpu_gate_batch_dev": (ctypes.c_int, [_vp, i32, i32, _vp, _vp, _vp, _vp, i32]),
        "tfhe_gpu_gate_batch_dev_noalloc": (ctypes.c_int, [_vp, i32, i32, _vp, _vp, _vp, _vp, i32]),
        "tfhe_gpu_boots_batch": (ctypes.c_int, [i32, P(LweSample), P(LweSample), P(LweSample),
                                                P(LweSample), i32, _vp]),
        "tfhe_gpu_bootstrap_woks_batch": (ctypes.c_int, [_vp, i32, _i32p, _i32p, i32, i32]),
        "tfhe_gpu_bootstrap_woks_acc_batch": (ctypes.c_int, [_vp, i32, _i32p, _i32p, _i32p, i32, i32]),
        "tfhe_gpu_bootstrap_batch": (ctypes.c_int, [_vp, i32, _i32p, _i32p, i32, i32]),
        "tfhe_gpu_blind_rotate_batch": (ctypes.c_int, [_vp, i32, _i32p, _i32p, i32]),
        "tfhe_gpu_keyswitch_batch": (ctypes.c_int, [_vp, _i32p, _i32p, i32]),
        "tfhe_gpu_extract_batch": (ctypes.c_int, [P(FlatParams), _i32p, _i32p, i32]),
        "tfhe_gpu_extern_product_batch": (ctypes.c_int, [_vp, i32, _i32p, _i32p, _f64p, i32]),
        "tfhe_gpu_extern_product_split_hi_batch": (ctypes.c_int, [_vp, _i32p, _i32p, _f64p, i32]),
        "tfhe_gpu_key_split_bits": (ctypes.c_int, [_vp, P(i32)]),
        "tfhe_gpu_rotate_decompose_batch": (ctypes.c_int, [P(FlatParams), _i32p, _i32p, _i32p, _i32p, i32]),
        "tfhe_gpu_fft_forward_batch": (ctypes.c_int, [i32, i32, _f64p, _i32p, i32]),
        "tfhe_gpu_fft_forward_digits_batch": (ctypes.c_int, [i32, _f64p, _i32p, i32]),
        "tfhe_gpu_split_digits_batch": (ctypes.c_int, [i32, i32, P(ctypes.c_float), _i32p, i32]),
        "tfhe_gpu_fft_inverse_batch": (ctypes.c_int, [i32, i32, _f64p, _f64p, i32]),
        "tfhe_gpu_fft_q4_batch": (ctypes.c_int, [i32, _f64p, _vp, i32]),
        "tfhe_gpu_malloc": (ctypes.c_int, [P(_vp), ctypes.c_size_t]),
        "tfhe_gpu_free": (ctypes.c_int, [_vp]),
        "tfhe_gpu_memcpy_h2d": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
        "tfhe_gpu_memcpy_d2h": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
        "tfhe_gpu_decrypt_batch_dev": (ctypes.c_int, [_i32p, _vp, _i32p, i32, i32]),
        "tfhe_gpu_measure_valu_peak": (ctypes.c_int, [P(ctypes.c_double), P(ctypes.c_double)]),
        "tfhe_gpu_profile_enable": (ctypes.c_int, [i32]),
        "tfhe_gpu_profile_reset": (ctypes.c_int, []),
        "tfhe_gpu_profile_read": (ctypes.c_int, [ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_int64)]),
        "tfhe_circuit_create": (ctypes.c_int, [i32, P(_vp)]),
        "tfhe_circuit_free": (ctypes.c_int, [_vp]),
        "tfhe_circuit_add": (ctypes.c_int, [_vp, i32, i32, i32, i32, P(i32)]),
        "tfhe_circuit_constant": (ctypes.c_int, [_vp, i32, P(i32)]),
        "tfhe_circuit_output": (ctypes.c_int, [_vp, i32]),
        "tfhe_circuit_info": (ctypes.c_int, [_vp, P(i32), P(i32), P(ctypes.c_int64), P(i32)]),
        "tfhe_circuit_level_of": (ctypes.c_int, [_vp, i32, P(i32)]),
        "tfhe_circuit_run": (ctypes.c_int, [_vp, _vp, i32, _vp, _vp, i32, i32]),
        "tfhe_gpu_circuit_level": (ctypes.c_int, [_vp, i32, i32, _i32p, P(_vp), P(_vp), P(_vp), P(_vp), i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ck(rc, what=""):
    if rc != 0:
        msg = lib().tfhe_gpu_last_error()
        raise TfheError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def i32p(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_i32p)


def f64p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_f64p)


# ---------------------------------------------------------------------------------------------
# host-side key sets (upstream API)
# ---------------------------------------------------------------------------------------------
class KeySet:
    """A secret key set made by new_random_gate_bootstrapping_secret_keyset, with flat views."""

    def __init__(self, ring=1024, seed=(0,)):
        L = lib()
        self.params_ptr = (L.new_default_gate_bootstrapping_parameters(110) if ring == 1024
                           else L.new_gate_bootstrapping_parameters_N2048())
        s = (ctypes.c_uint32 * len(seed))(*seed)
        L.tfhe_random_generator_setSeed(s, len(seed))
        self.ptr = L.new_random_gate_bootstrapping_secret_keyset(self.params_ptr)
        fp = FlatParams()
        _ck(L.tfhe_flat_params_of(self.params_ptr, ctypes.byref(fp)), "tfhe_flat_params_of")
        self.fp = fp
        self.p = fp.as_dict()
        n, N, k, l, t, bb = fp.n, fp.N, fp.k, fp.l, fp.ks_t, fp.ks_basebit
        self.lwe_key = np.zeros(n, np.int32)
        self.tlwe_key = np.zeros(k * N, np.int32)
        self.bk = np.zeros((n, (k + 1) * l, k + 1, N), np.int32)
        self.ksk = np.zeros((k * N, t, 1 << bb, n + 1), np.int32)
        _ck(L.tfhe_keyset_export_flat(self.ptr, i32p(self.lwe_key), i32p(self.tlwe_key), i32p(self.bk),
                                      i32p(self.ksk)), "tfhe_keyset_export_flat")

    @property
    def cloud_ptr(self):
        # TFheGateBootstrappingSecretKeySet: params, lwe_key, tgsw_key, then the embedded cloud set
        return self.ptr + 3 * ctypes.sizeof(ctypes.c_void_p)

    def encrypt(self, bits):
        bits = np.ascontiguousarray(np.asarray(bits).reshape(-1), dtype=np.int32)
        out = np.zeros((len(bits), self.fp.n + 1), np.int32)
        _ck(lib().tfhe_encrypt_bits_flat(self.ptr, i32p(bits), len(bits), i32p(out)), "encrypt")
        return out

    def decrypt(self, samples):
        samples = np.ascontiguousarray(samples, dtype=np.int32)
        bits = np.zeros(samples.shape[0], np.int32)
        _ck(lib().tfhe_decrypt_bits_flat(self.ptr, i32p(samples), samples.shape[0], i32p(bits)), "decrypt")
        return bits

    def free(self):
        if self.ptr:
            lib().delete_gate_bootstrapping_secret_keyset(self.ptr)
            lib().delete_gate_bootstrapping_parameters(self.params_ptr)
            self.ptr = None


# ---------------------------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------------------------
def init(device=0):
    _ck(lib().tfhe_gpu_init(device), "tfhe_gpu_init")


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int32()
    mem = ctypes.c_int64()
    _ck(lib().tfhe_gpu_device_info(name, 256, ctypes.byref(cu), ctypes.byref(mem)), "device_info")
    return {"name": name.value.decode(), "cu": cu.value, "hbm_bytes": mem.value}


def sync():
    _ck(lib().tfhe_gpu_sync(), "sync")


def set_split_bits(b):
    _ck(lib().tfhe_gpu_set_split_bits(int(b)), "set_split_bits")


def set_split_mac(mode):
    """Split-FP32 accumulation at N=1024: 0 plain FMAs (default), 1 compensated (TwoProduct + TwoSum)."""
    _ck(lib().tfhe_gpu_set_split_mac(int(mode)), "set_split_mac")


def set_launch_form(form):
    """Bootstrap launch form for launches that name none (0: size rules; 1 FP64 key sharing at
    N=1024; 2 one bootstrap per workgroup; 3 latency form; 4, 5 N=2048 with 2 / 4 per workgroup;
    6 split FP32 at N=1024 with 4 bootstraps per wavefront and the products on the matrix cores)."""
    _ck(lib().tfhe_gpu_set_launch_form(int(form)), "set_launch_form")


def set_fused_keyswitch(on):
    """Split-FP32 gate batches at N=1024 (not MUX): keyswitch inside the bootstrap kernel (True) or in
    the separate keyswitch kernel (False, the default)."""
    _ck(lib().tfhe_gpu_set_fused_keyswitch(1 if on else 0), "set_fused_keyswitch")


def bk_to_lagrange(flat_params, bk):
    """FFT-domain values (upstream LagrangeHalfC order, torus units) of every bk polynomial."""
    fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
    bk = np.ascontiguousarray(bk, dtype=np.int32)
    out = np.zeros(bk.shape[:-1] + (fp.N // 2,), np.complex128)
    _ck(lib().tfhe_bk_to_lagrange(ctypes.byref(fp), i32p(bk), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
        "bk_to_lagrange")
    return out


def bk_from_lagrange(flat_params, lagrange):
    """Exact time-domain bk from FFT-domain values; raises TfheError when they are not the
    transform of torus32 polynomials."""
    fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
    lag = np.ascontiguousarray(lagrange, dtype=np.complex128)
    out = np.zeros(lag.shape[:-1] + (fp.N,), np.int32)
    _ck(lib().tfhe_bk_from_lagrange(ctypes.byref(fp), lag.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), i32p(out)),
        "bk_from_lagrange")
    return out


def set_fp64_shared_rows(on):
    """FP64 launch form: True (key-sharing workgroups), False (one bootstrap per workgroup) or
    "auto" (the default: by launch size).  At N=2048, True and False both select the batched
    two-wavefront kernel and "auto" adds the latency form for small launches."""
    _ck(lib().tfhe_gpu_set_fp64_shared_rows(2 if on == "auto" else (1 if on else 0)), "set_fp64_shared_rows")


class DeviceKey:
    def __init__(self, flat_params, bk, ksk, prec_mask=PREC_MASK_FP64 | PREC_MASK_SPLIT32):
        fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
        self.fp = fp
        self.ptr = ctypes.c_void_p()
        bk = np.ascontiguousarray(bk, dtype=np.int32)
        ksk = np.ascontiguousarray(ksk, dtype=np.int32)
        _ck(lib().tfhe_gpu_key_create(ctypes.byref(fp), i32p(bk), i32p(ksk), prec_mask, ctypes.byref(self.ptr)),
            "tfhe_gpu_key_create")

    @classmethod
    def from_keyset(cls, ks: KeySet, prec_mask=PREC_MASK_FP64 | PREC_MASK_SPLIT32):
        return cls(ks.fp, ks.bk, ks.ksk, prec_mask)

    @classmethod
    def from_lagrange(cls, flat_params, lagrange, ksk, prec_mask=PREC_MASK_FP64 | PREC_MASK_SPLIT32):
        """Device key from an FFT-domain bk in upstream LagrangeHalfC order (complex128
        [n][(k+1)l][k+1][N/2], torus units), as tfhe_gpu_key_create_lagrange."""
        self = cls.__new__(cls)
        fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
        self.fp = fp
        self.ptr = ctypes.c_void_p()
        lag = np.ascontiguousarray(lagrange, dtype=np.complex128)
        ksk = np.ascontiguousarray(ksk, dtype=np.int32)
        _ck(lib().tfhe_gpu_key_create_lagrange(ctypes.byref(fp), lag.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               i32p(ksk), prec_mask, ctypes.byref(self.ptr)),
            "tfhe_gpu_key_create_lagrange")
        return self

    def free(self):
        if self.ptr:
            lib().tfhe_gpu_key_free(self.ptr)
            self.ptr = None

    # ---- batches on host arrays ----
    def gate(self, gate, a, b, c=None, prec=PREC_FP64):
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        c = None if c is None else np.ascontiguousarray(c, np.int32)
        out = np.zeros_like(a)
        _ck(lib().tfhe_gpu_gate_batch(self.ptr, prec, GATES[gate] if isinstance(gate, str) else gate,
                                      i32p(out), i32p(a), i32p(b), i32p(c), a.shape[0]), "gate_batch")
        return out

    def bootstrap_woks(self, x, mu=MU, prec=PREC_FP64, want_acc=False):
        x = np.ascontiguousarray(x, np.int32)
        cnt = x.shape[0]
        u = np.zeros((cnt, self.fp.k * self.fp.N + 1), np.int32)
        if want_acc:
            acc = np.zeros((cnt, self.fp.k + 1, self.fp.N), np.int32)
            _ck(lib().tfhe_gpu_bootstrap_woks_acc_batch(self.ptr, prec, i32p(u), i32p(acc), i32p(x), mu, cnt),
                "bootstrap_woks_acc")
            return u, acc
        _ck(lib().tfhe_gpu_bootstrap_woks_batch(self.ptr, prec, i32p(u), i32p(x), mu, cnt), "bootstrap_woks")
        return u

    def bootstrap(self, x, mu=MU, prec=PREC_FP64):
        x = np.ascontiguousarray(x, np.int32)
        out = np.zeros_like(x)
        _ck(lib().tfhe_gpu_bootstrap_batch(self.ptr, prec, i32p(out), i32p(x), mu, x.shape[0]), "bootstrap")
        return out

    def blind_rotate(self, acc, bara, prec=PREC_FP64):
        acc = np.ascontiguousarray(acc, np.int32).copy()
        bara = np.ascontiguousarray(bara, np.int32)
        _ck(lib().tfhe_gpu_blind_rotate_batch(self.ptr, prec, i32p(acc), i32p(bara), acc.shape[0]), "blind_rotate")
        return acc

    def extern_product(self, acc, row_set, prec=PREC_FP64, want_raw=False):
        acc = np.ascontiguousarray(acc, np.int32).copy()
        rs = np.ascontiguousarray(row_set, np.int32)
        raw = np.zeros(acc.shape, np.float64) if want_raw else None
        _ck(lib().tfhe_gpu_extern_product_batch(self.ptr, prec, i32p(acc), i32p(rs), f64p(raw), acc.shape[0]),
            "extern_product")
        return (acc, raw) if want_raw else acc

    def extern_product_split_hi(self, acc, row_set):
        """Split-FP32 external product and the hi-half product before rounding (units of 2^split)."""
        acc = np.ascontiguousarray(acc, np.int32).copy()
        rs = np.ascontiguousarray(row_set, np.int32)
        raw_hi = np.zeros(acc.shape, np.float64)
        _ck(lib().tfhe_gpu_extern_product_split_hi_batch(self.ptr, i32p(acc), i32p(rs), f64p(raw_hi), acc.shape[0]),
            "extern_product_split_hi")
        return acc, raw_hi

    @property
    def split_bits(self):
        b = ctypes.c_int32()
        _ck(lib().tfhe_gpu_key_split_bits(self.ptr, ctypes.byref(b)), "key_split_bits")
        return b.value

    def keyswitch(self, u):
        u = np.ascontiguousarray(u, np.int32)
        out = np.zeros((u.shape[0], self.fp.n + 1), np.int32)
        _ck(lib().tfhe_gpu_keyswitch_batch(self.ptr, i32p(out), i32p(u), u.shape[0]), "keyswitch")
        return out

    # ---- device-resident batches ----
    def gate_dev(self, gate, d_out, d_a, d_b, d_c, count, prec=PREC_FP64):
        _ck(lib().tfhe_gpu_gate_batch_dev(self.ptr, prec, GATES[gate] if isinstance(gate, str) else gate,
                                          d_out, d_a, d_b, d_c, count), "gate_batch_dev")


def extract(flat_params, acc):
    fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
    acc = np.ascontiguousarray(acc, np.int32)
    u = np.zeros((acc.shape[0], fp.k * fp.N + 1), np.int32)
    _ck(lib().tfhe_gpu_extract_batch(ctypes.byref(fp), i32p(u), i32p(acc), acc.shape[0]), "extract")
    return u


def rotate_decompose(flat_params, acc, bara):
    fp = FlatParams(**flat_params) if isinstance(flat_params, dict) else flat_params
    acc = np.ascontiguousarray(acc, np.int32)
    bara = np.ascontiguousarray(bara, np.int32)
    cnt = acc.shape[0]
    rot = np.zeros((cnt, 2, fp.N), np.int32)
    dig = np.zeros((cnt, 2 * fp.l, fp.N), np.int32)
    _ck(lib().tfhe_gpu_rotate_decompose_batch(ctypes.byref(fp), i32p(rot), i32p(dig), i32p(acc), i32p(bara), cnt),
        "rotate_decompose")
    return rot, dig


def fft_forward(polys, prec=PREC_FP64):
    polys = np.ascontiguousarray(polys, np.int32)
    cnt, N = polys.shape
    out = np.zeros((cnt, N), np.float64)
    _ck(lib().tfhe_gpu_fft_forward_batch(N, prec, f64p(out), i32p(polys), cnt), "fft_forward")
    return out[:, 0::2] + 1j * out[:, 1::2]


def split_digits(x, Bgbit, variant=0):
    """The split kernels' gadget digits of coefficient pairs as FP16 (returned as float32):
    x [count][2] int32 -> [count][3][2], digit lvl of each coefficient of the pair (l = 3)."""
    x = np.ascontiguousarray(x, np.int32)
    cnt = x.shape[0]
    out = np.zeros((cnt, 3, 2), np.float32)
    _ck(lib().tfhe_gpu_split_digits_batch(Bgbit, variant, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           i32p(x), cnt), "split_digits")
    return out


def fft_forward_digits(polys):
    """The split path's forward transform of gadget digits (N = 1024 or 2048, |p_j| < 2048), its
    first pass on the matrix cores, as the bootstrap kernels run it."""
    polys = np.ascontiguousarray(polys, np.int32)
    cnt, N = polys.shape
    out = np.zeros((cnt, N), np.float64)
    _ck(lib().tfhe_gpu_fft_forward_digits_batch(N, f64p(out), i32p(polys), cnt), "fft_forward_digits")
    return out[:, 0::2] + 1j * out[:, 1::2]


def fft_inverse(F, prec=PREC_FP64):
    F = np.asarray(F, np.complex128)
    cnt, M = F.shape
    inp = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1).reshape(cnt, 2 * M))
    out = np.zeros((cnt, 2 * M), np.float64)
    _ck(lib().tfhe_gpu_fft_inverse_batch(2 * M, prec, f64p(out), f64p(inp), cnt), "fft_inverse")
    return out


def fft_q4_forward(polys):
    """Forward transform (N=1024, FP32) in launch form 6's layout: int32 [count][1024] -> complex F."""
    polys = np.ascontiguousarray(polys, np.int32)
    cnt = polys.shape[0]
    out = np.zeros((cnt, 1024), np.float64)
    _ck(lib().tfhe_gpu_fft_q4_batch(0, f64p(out), polys.ctypes.data_as(_vp), cnt), "fft_q4_forward")
    return out[:, 0::2] + 1j * out[:, 1::2]


def fft_q4_inverse(F):
    """Inverse of fft_q4_forward to integer-unit doubles [count][1024]."""
    F = np.asarray(F, np.complex128)
    cnt = F.shape[0]
    inp = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1).reshape(cnt, 1024))
    out = np.zeros((cnt, 1024), np.float64)
    _ck(lib().tfhe_gpu_fft_q4_batch(1, f64p(out), inp.ctypes.data_as(_vp), cnt), "fft_q4_inverse")
    return out


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        _ck(lib().tfhe_gpu_malloc(ctypes.byref(self.ptr), self.nbytes), "malloc")

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        _ck(lib().tfhe_gpu_memcpy_h2d(b.ptr, arr.ctypes.data_as(_vp), arr.nbytes), "h2d")
        return b

    def to_array(self, shape, dtype=np.int32):
        out = np.zeros(shape, dtype)
        _ck(lib().tfhe_gpu_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().tfhe_gpu_free(self.ptr)
            self.ptr = None


def decrypt_dev(d_samples, lwe_key, count):
    key = np.ascontiguousarray(lwe_key, np.int32)
    bits = np.zeros(count, np.int32)
    _ck(lib().tfhe_gpu_decrypt_batch_dev(i32p(bits), d_samples.ptr if isinstance(d_samples, DeviceBuffer)
                                         else d_samples, i32p(key), key.shape[0], count), "decrypt_dev")
    return bits


def profile_enable(on=True):
    _ck(lib().tfhe_gpu_profile_enable(1 if on else 0), "profile_enable")


def profile_reset():
    _ck(lib().tfhe_gpu_profile_reset(), "profile_reset")


def profile_read(kernel):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    _ck(lib().tfhe_gpu_profile_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)), "profile_read")
    return ms.value, n.value


def measure_valu_peak():
    """Attainable dense-FMA VALU TFLOP/s (fp64, fp32) on this device."""
    a, b = ctypes.c_double(), ctypes.c_double()
    _ck(lib().tfhe_gpu_measure_valu_peak(ctypes.byref(a), ctypes.byref(b)), "measure_valu_peak")
    return a.value, b.value


class Circuit:
    """A netlist over wires (inputs 0..n_inputs-1, then one wire per added gate or constant),
    evaluated level by level on the device over a batch of independent instances."""

    def __init__(self, n_inputs):
        self.ptr = ctypes.c_void_p()
        _ck(lib().tfhe_circuit_create(int(n_inputs), ctypes.byref(self.ptr)), "circuit_create")
        self.n_inputs = int(n_inputs)
        self.n_outputs = 0

    def add(self, op, a, b=-1, c=-1):
        w = ctypes.c_int32()
        code = {"NOT": OP_NOT, "COPY": OP_COPY}.get(op, None) if isinstance(op, str) else op
        if code is None:
            code = GATES[op]
        _ck(lib().tfhe_circuit_add(self.ptr, code, int(a), int(b), int(c), ctypes.byref(w)), "circuit_add")
        return w.value

    def constant(self, value):
        w = ctypes.c_int32()
        _ck(lib().tfhe_circuit_constant(self.ptr, int(value), ctypes.byref(w)), "circuit_constant")
        return w.value

    def output(self, wire):
        _ck(lib().tfhe_circuit_output(self.ptr, int(wire)), "circuit_output")
        self.n_outputs += 1

    def info(self):
        lv, mx, no = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        nb = ctypes.c_int64()
        _ck(lib().tfhe_circuit_info(self.ptr, ctypes.byref(lv), ctypes.byref(mx), ctypes.byref(nb),
                                    ctypes.byref(no)), "circuit_info")
        return {"levels": lv.value, "max_level_bootstraps": mx.value, "bootstraps": nb.value, "outputs": no.value}

    def level_plan(self):
        """[(bootstraps per instance, launch form of the last run)] for levels 1..n_levels."""
        out = []
        for lv in range(1, self.info()["levels"] + 1):
            nb, form = ctypes.c_int32(), ctypes.c_int32()
            _ck(lib().tfhe_circuit_level_plan(self.ptr, lv, ctypes.byref(nb), ctypes.byref(form)), "circuit_level_plan")
            out.append((nb.value, form.value))
        return out

    def level_of(self, wire):
        lv = ctypes.c_int32()
        _ck(lib().tfhe_circuit_level_of(self.ptr, int(wire), ctypes.byref(lv)), "circuit_level_of")
        return lv.value

    def run_dev(self, key, d_in, d_out, batch, prec=PREC_FP64, graph=False, autotune=False):
        flags = (CIRCUIT_GRAPH if graph else 0) | (CIRCUIT_AUTOTUNE if autotune else 0)
        din = d_in.ptr if isinstance(d_in, DeviceBuffer) else d_in
        dout = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
        _ck(lib().tfhe_circuit_run(self.ptr, key.ptr, prec, din, dout, int(batch), flags), "circuit_run")

    def run(self, key, inputs, prec=PREC_FP64, graph=False, autotune=False):
        """inputs: int32 [n_inputs][batch][n+1] -> outputs [n_outputs][batch][n+1] (synchronous)."""
        inputs = np.ascontiguousarray(inputs, np.int32)
        batch = inputs.shape[1]
        d_in = DeviceBuffer.from_array(inputs)
        d_out = DeviceBuffer(max(4, self.n_outputs * batch * inputs.shape[2] * 4))
        try:
            self.run_dev(key, d_in, d_out, batch, prec, graph, autotune)
            sync()
            return d_out.to_array((self.n_outputs, batch, inputs.shape[2]))
        finally:
            d_in.free()
            d_out.free()

    def free(self):
        if self.ptr:
            lib().tfhe_circuit_free(self.ptr)
            self.ptr = None


def ripple_carry_adder(bits):
    """Circuit of a `bits`-wide ripple-carry adder: inputs a_0..a_{bits-1}, b_0..b_{bits-1} (LSB
    first); outputs s_0..s_{bits-1} and the carry out.  p_i = XOR(a_i, b_i), s_i = XOR(p_i, c_i),
    c_{i+1} = MUX(p_i, c_i, a_i) with c_0 = 0."""
    c = Circuit(2 * bits)
    p = [c.add("XOR", i, bits + i) for i in range(bits)]
    carry = c.constant(0)
    sums = []
    for i in range(bits):
        sums.append(c.add("XOR", p[i], carry))
        carry = c.add("MUX", p[i], carry, i)
    for w in sums:
        c.output(w)
    c.output(carry)
    return c


def circuit_level(key, ops, a, b, c, out, batch, prec=PREC_FP64):
    """One level of independent gates over device arrays (lists of DeviceBuffer or pointers)."""
    n = len(ops)
    ptrs = lambda xs: (_vp * n)(*[(x.ptr if isinstance(x, DeviceBuffer) else x) if x is not None else None
                                  for x in xs])
    codes = np.array([GATES[o] if isinstance(o, str) else o for o in ops], np.int32)
    cc = ptrs(c) if c is not None else None
    _ck(lib().tfhe_gpu_circuit_level(key.ptr, prec, n, i32p(codes), ptrs(a), ptrs(b), cc, ptrs(out), int(batch)),
        "circuit_level")
