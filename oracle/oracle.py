"""ctypes binding of the CPU oracle (oracle/_build/libtfhe_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtfhe_oracle.so")

GATES = {"NAND": 0, "AND": 1, "OR": 2, "XOR": 3, "XNOR": 4, "NOR": 5,
         "ANDNY": 6, "ANDYN": 7, "ORNY": 8, "ORYN": 9, "MUX": 10}

# Truth tables of the gate set, as plaintext functions of (a, b, c).
TRUTH = {
    "NAND": lambda a, b, c: 1 - (a & b),
    "AND": lambda a, b, c: a & b,
    "OR": lambda a, b, c: a | b,
    "XOR": lambda a, b, c: a ^ b,
    "XNOR": lambda a, b, c: 1 - (a ^ b),
    "NOR": lambda a, b, c: 1 - (a | b),
    "ANDNY": lambda a, b, c: (1 - a) & b,
    "ANDYN": lambda a, b, c: a & (1 - b),
    "ORNY": lambda a, b, c: (1 - a) | b,
    "ORYN": lambda a, b, c: a | (1 - b),
    "MUX": lambda a, b, c: b if a else c,
}


class OrcParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("N", ctypes.c_int32), ("k", ctypes.c_int32),
                ("l", ctypes.c_int32), ("Bgbit", ctypes.c_int32), ("ks_t", ctypes.c_int32),
                ("ks_basebit", ctypes.c_int32), ("lwe_alpha", ctypes.c_double),
                ("bk_alpha", ctypes.c_double)]


@dataclass(frozen=True)
class Params:
    n: int
    N: int
    k: int
    l: int
    Bgbit: int
    ks_t: int
    ks_basebit: int
    lwe_alpha: float
    bk_alpha: float

    def c(self):
        return OrcParams(self.n, self.N, self.k, self.l, self.Bgbit, self.ks_t,
                         self.ks_basebit, self.lwe_alpha, self.bk_alpha)

    @property
    def kpl(self):
        return (self.k + 1) * self.l


# kernel/src/include/constants.h:26-36
PARAMS_1024 = Params(n=630, N=1024, k=1, l=3, Bgbit=7, ks_t=8, ks_basebit=2,
                     lwe_alpha=2.0 ** -15, bk_alpha=2.0 ** -25)
# the larger ring of the N=2048 workload
PARAMS_2048 = Params(n=750, N=2048, k=1, l=3, Bgbit=8, ks_t=10, ks_basebit=2,
                     lwe_alpha=2.0 ** -17, bk_alpha=2.0 ** -30)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        i32p, i64p, u32p, f64p = P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_uint32), P(ctypes.c_double)
        pp = P(OrcParams)
        vp = ctypes.c_void_p
        i32, f64 = ctypes.c_int32, ctypes.c_double
        sig = {
            "orc_set_seed": (None, [u32p, i32]),
            "orc_dtot32": (i32, [f64]),
            "orc_t32tod": (f64, [i32]),
            "orc_gaussian32": (i32, [i32, f64]),
            "orc_uniform_torus32": (i32, []),
            "orc_approx_phase": (i32, [i32, i32]),
            "orc_modswitch_from_torus32": (i32, [i32, i32]),
            "orc_modswitch_to_torus32": (i32, [i32, i32]),
            "orc_decomp_offset": (ctypes.c_uint32, [i32, i32]),
            "orc_keygen": (None, [pp, u32p, i32, i32p, i32p, i32p, i32p]),
            "orc_lwe_encrypt": (None, [i32p, i32, f64, i32p, i32]),
            "orc_lwe_phase": (i32, [i32p, i32p, i32]),
            "orc_boots_encrypt": (None, [pp, i32p, i32, i32p]),
            "orc_boots_decrypt": (i32, [pp, i32p, i32p]),
            "orc_tlwe_phase": (None, [pp, i32p, i32p, i32p]),
            "orc_poly_mul_xai": (None, [i32p, i32, i32p, i32]),
            "orc_poly_mul_xai_minus_one": (None, [i32p, i32, i32p, i32]),
            "orc_decompose": (None, [i32p, i32p, i32, i32, i32]),
            "orc_negacyclic_mul_i64": (None, [i64p, i32p, i32p, i32]),
            "orc_negacyclic_mul_naive": (None, [i32p, i32p, i32p, i32]),
            "orc_fft_forward": (None, [f64p, i32p, i32]),
            "orc_fft_inverse_raw": (None, [f64p, f64p, i32]),
            "orc_lagrange_half_reference": (None, [f64p, i32p, i32]),
            "orc_torus_from_lagrange_reference": (None, [i32p, f64p, i32]),
            "orc_extern_product_exact": (None, [pp, i32p, i32p, i64p]),
            "orc_fft_key_create": (vp, [pp, i32p]),
            "orc_fft_key_free": (None, [vp]),
            "orc_extern_product_fft": (None, [pp, i32p, vp, i32, f64p, i32]),
            "orc_blind_rotate": (None, [pp, i32p, i32p, i32p, vp]),
            "orc_test_vector": (None, [pp, i32p, i32, i32]),
            "orc_extract": (None, [pp, i32p, i32p]),
            "orc_keyswitch": (None, [pp, i32p, i32p, i32p]),
            "orc_bootstrap_woks": (None, [pp, i32p, i32p, i32, i32p, vp]),
            "orc_bootstrap": (None, [pp, i32p, i32p, i32, i32p, vp, i32p]),
            "orc_gate_combine": (None, [pp, i32, i32p, i32p, i32p]),
            "orc_gate": (None, [pp, i32, i32p, i32p, i32p, i32p, i32p, vp, i32p]),
            "orc_gate_batch": (f64, [pp, i32, i32p, i32p, i32p, i32p, i32, vp, i32p, i32]),
            "orc_set_fault": (None, [i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


def i32p(a):
    assert a is None or (a.dtype == np.int32 and a.flags["C_CONTIGUOUS"])
    return _p(a, ctypes.c_int32)


def i64p(a):
    assert a is None or (a.dtype == np.int64 and a.flags["C_CONTIGUOUS"])
    return _p(a, ctypes.c_int64)


def f64p(a):
    assert a is None or (a.dtype == np.float64 and a.flags["C_CONTIGUOUS"])
    return _p(a, ctypes.c_double)


def set_seed(values):
    v = np.ascontiguousarray(values, dtype=np.uint32)
    lib().orc_set_seed(_p(v, ctypes.c_uint32), len(v))


def set_fault(f):
    lib().orc_set_fault(int(f))


# ---- scalar numerics ----
def dtot32(d):
    return lib().orc_dtot32(float(d))


def t32tod(x):
    return lib().orc_t32tod(int(x))


def gaussian32(message, sigma):
    return lib().orc_gaussian32(int(message), float(sigma))


def approx_phase(phase, msize):
    return lib().orc_approx_phase(int(phase), int(msize))


def modswitch_from(phase, msize):
    return lib().orc_modswitch_from_torus32(int(phase), int(msize))


def modswitch_to(mu, msize):
    return lib().orc_modswitch_to_torus32(int(mu), int(msize))


def decomp_offset(l, Bgbit):
    return lib().orc_decomp_offset(l, Bgbit)


# ---- keys ----
@dataclass
class Keys:
    params: Params
    lwe_key: np.ndarray
    tlwe_key: np.ndarray
    ksk: np.ndarray
    bk: np.ndarray
    _fkey: object = None

    @property
    def fkey(self):
        if self._fkey is None:
            pc = self.params.c()
            self._fkey = lib().orc_fft_key_create(ctypes.byref(pc), i32p(self.bk))
        return self._fkey

    def free(self):
        if self._fkey is not None:
            lib().orc_fft_key_free(self._fkey)
            self._fkey = None


def key_shapes(p: Params):
    base = 1 << p.ks_basebit
    return dict(lwe_key=(p.n,), tlwe_key=(p.k * p.N,),
                ksk=(p.k * p.N, p.ks_t, base, p.n + 1),
                bk=(p.n, p.kpl, p.k + 1, p.N))


def keygen(p: Params, seed=(0,)):
    sh = key_shapes(p)
    arrs = {k: np.zeros(v, dtype=np.int32) for k, v in sh.items()}
    s = np.ascontiguousarray(seed, dtype=np.uint32)
    pc = p.c()
    lib().orc_keygen(ctypes.byref(pc), _p(s, ctypes.c_uint32), len(s), i32p(arrs["lwe_key"]),
                     i32p(arrs["tlwe_key"]), i32p(arrs["ksk"]), i32p(arrs["bk"]))
    return Keys(p, **arrs)


def boots_encrypt(p: Params, bits, lwe_key):
    bits = np.asarray(bits).reshape(-1)
    out = np.zeros((len(bits), p.n + 1), dtype=np.int32)
    pc = p.c()
    for i, b in enumerate(bits):
        lib().orc_boots_encrypt(ctypes.byref(pc), i32p(out[i]), int(b), i32p(lwe_key))
    return out


def lwe_phase(samples, key):
    """Phase b - <a, s> (mod 2^32) of each row, computed in numpy int64 (independent of the C code)."""
    s = np.asarray(samples, dtype=np.int64)
    k = np.asarray(key, dtype=np.int64)
    n = k.shape[0]
    ph = (s[..., n] - (s[..., :n] * k).sum(axis=-1)) & 0xFFFFFFFF
    return ph.astype(np.uint32).view(np.int32)


def boots_decrypt(samples, key):
    return (lwe_phase(samples, key) > 0).astype(np.int32)


# ---- polynomial primitives ----
def poly_mul_xai(a, src):
    src = np.ascontiguousarray(src, dtype=np.int32)
    res = np.zeros_like(src)
    lib().orc_poly_mul_xai(i32p(res), int(a), i32p(src), len(src))
    return res


def poly_mul_xai_minus_one(a, src):
    src = np.ascontiguousarray(src, dtype=np.int32)
    res = np.zeros_like(src)
    lib().orc_poly_mul_xai_minus_one(i32p(res), int(a), i32p(src), len(src))
    return res


def decompose(src, l, Bgbit):
    src = np.ascontiguousarray(src, dtype=np.int32)
    res = np.zeros((l, len(src)), dtype=np.int32)
    lib().orc_decompose(i32p(res), i32p(src), len(src), l, Bgbit)
    return res


def negacyclic_mul_i64(ip, tp):
    ip = np.ascontiguousarray(ip, dtype=np.int32)
    tp = np.ascontiguousarray(tp, dtype=np.int32)
    res = np.zeros(len(ip), dtype=np.int64)
    lib().orc_negacyclic_mul_i64(i64p(res), i32p(ip), i32p(tp), len(ip))
    return res


def negacyclic_mul_naive(ip, tp):
    ip = np.ascontiguousarray(ip, dtype=np.int32)
    tp = np.ascontiguousarray(tp, dtype=np.int32)
    res = np.zeros(len(ip), dtype=np.int32)
    lib().orc_negacyclic_mul_naive(i32p(res), i32p(ip), i32p(tp), len(ip))
    return res


def fft_forward(poly):
    poly = np.ascontiguousarray(poly, dtype=np.int32)
    out = np.zeros(len(poly), dtype=np.float64)
    lib().orc_fft_forward(f64p(out), i32p(poly), len(poly))
    return out[0::2] + 1j * out[1::2]


def fft_inverse_raw(F, N):
    F = np.asarray(F, dtype=np.complex128)
    inp = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1).reshape(-1))
    out = np.zeros(N, dtype=np.float64)
    lib().orc_fft_inverse_raw(f64p(out), f64p(inp), N)
    return out


def lagrange_half_reference(poly):
    poly = np.ascontiguousarray(poly, dtype=np.int32)
    out = np.zeros(len(poly), dtype=np.float64)
    lib().orc_lagrange_half_reference(f64p(out), i32p(poly), len(poly))
    return out[0::2] + 1j * out[1::2]


def torus_from_lagrange_reference(values):
    """TorusPolynomial_fft.cpp:4-43 on N/2 complex values (truncating)."""
    v = np.ascontiguousarray(values, dtype=np.complex128)
    N = 2 * len(v)
    out = np.zeros(N, dtype=np.int32)
    lib().orc_torus_from_lagrange_reference(i32p(out), f64p(v.view(np.float64)), N)
    return out


# ---- external product / bootstrapping ----
def extern_product_exact(p: Params, acc, tgsw, want_raw=False):
    acc = np.ascontiguousarray(acc, dtype=np.int32).copy()
    tgsw = np.ascontiguousarray(tgsw, dtype=np.int32)
    raw = np.zeros(acc.shape, dtype=np.int64) if want_raw else None
    pc = p.c()
    lib().orc_extern_product_exact(ctypes.byref(pc), i32p(acc), i32p(tgsw), i64p(raw))
    return (acc, raw) if want_raw else acc


def extern_product_fft(keys: Keys, acc, row_set, truncate=False, want_raw=False):
    p = keys.params
    acc = np.ascontiguousarray(acc, dtype=np.int32).copy()
    raw = np.zeros(acc.shape, dtype=np.float64) if want_raw else None
    pc = p.c()
    lib().orc_extern_product_fft(ctypes.byref(pc), i32p(acc), keys.fkey, int(row_set), f64p(raw),
                                 1 if truncate else 0)
    return (acc, raw) if want_raw else acc


def blind_rotate(keys: Keys, acc, bara, exact=False):
    p = keys.params
    acc = np.ascontiguousarray(acc, dtype=np.int32).copy()
    bara = np.ascontiguousarray(bara, dtype=np.int32)
    pc = p.c()
    lib().orc_blind_rotate(ctypes.byref(pc), i32p(acc), i32p(bara), i32p(keys.bk),
                           None if exact else keys.fkey)
    return acc


def test_vector(p: Params, barb, mu):
    acc = np.zeros((p.k + 1, p.N), dtype=np.int32)
    pc = p.c()
    lib().orc_test_vector(ctypes.byref(pc), i32p(acc), int(barb), int(mu))
    return acc


def tlwe_phase(p: Params, sample, tlwe_key):
    sample = np.ascontiguousarray(sample, dtype=np.int32)
    ph = np.zeros(p.N, dtype=np.int32)
    pc = p.c()
    lib().orc_tlwe_phase(ctypes.byref(pc), i32p(ph), i32p(sample), i32p(tlwe_key))
    return ph


def extract(p: Params, acc):
    acc = np.ascontiguousarray(acc, dtype=np.int32)
    u = np.zeros(p.k * p.N + 1, dtype=np.int32)
    pc = p.c()
    lib().orc_extract(ctypes.byref(pc), i32p(u), i32p(acc))
    return u


def keyswitch(keys: Keys, u):
    p = keys.params
    u = np.ascontiguousarray(u, dtype=np.int32)
    res = np.zeros(p.n + 1, dtype=np.int32)
    pc = p.c()
    lib().orc_keyswitch(ctypes.byref(pc), i32p(res), i32p(keys.ksk), i32p(u))
    return res


def bootstrap_woks(keys: Keys, x, mu, exact=False):
    p = keys.params
    x = np.ascontiguousarray(x, dtype=np.int32)
    u = np.zeros(p.k * p.N + 1, dtype=np.int32)
    pc = p.c()
    lib().orc_bootstrap_woks(ctypes.byref(pc), i32p(u), i32p(x), int(mu), i32p(keys.bk),
                             None if exact else keys.fkey)
    return u


def gate_combine(p: Params, gate, a, b):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    tmp = np.zeros(p.n + 1, dtype=np.int32)
    pc = p.c()
    lib().orc_gate_combine(ctypes.byref(pc), GATES[gate], i32p(tmp), i32p(a), i32p(b))
    return tmp


def gate(keys: Keys, name, a, b, c=None, exact=False):
    p = keys.params
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    c = None if c is None else np.ascontiguousarray(c, dtype=np.int32)
    out = np.zeros(p.n + 1, dtype=np.int32)
    pc = p.c()
    lib().orc_gate(ctypes.byref(pc), GATES[name], i32p(out), i32p(a), i32p(b), i32p(c),
                   i32p(keys.bk), None if exact else keys.fkey, i32p(keys.ksk))
    return out


def gate_batch(keys: Keys, name, a, b, c=None, nthreads=0):
    """Batch of independent gates on the FP64 FFT path (OpenMP).  Returns (outputs, seconds)."""
    p = keys.params
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    c = None if c is None else np.ascontiguousarray(c, dtype=np.int32)
    out = np.zeros_like(a)
    pc = p.c()
    sec = lib().orc_gate_batch(ctypes.byref(pc), GATES[name], i32p(out), i32p(a), i32p(b), i32p(c),
                               a.shape[0], keys.fkey, i32p(keys.ksk), int(nthreads))
    return out, sec
