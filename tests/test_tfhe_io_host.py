"""CPU: key-set, parameter-set and ciphertext files (upstream tfhe_io names, this library's own
byte format -- parity with upstream's bytes is unpinned, as upstream is absent).  Round trips must
reproduce every key array and ciphertext exactly; malformed files must be refused; the client side
of the demo flow (tests/c/demo_flow.c) must write files the library reads back."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

import tfhe_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
libc = ctypes.CDLL(None)
libc.fopen.restype = ctypes.c_void_p
libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
libc.fclose.argtypes = [ctypes.c_void_p]


def fopen(path, mode):
    f = libc.fopen(str(path).encode(), mode.encode())
    assert f, path
    return f


def flat_cloud(cloud_ptr, ks):
    bk = np.zeros_like(ks.bk)
    ksk = np.zeros_like(ks.ksk)
    assert T.lib().tfhe_cloud_export_flat(cloud_ptr, T.i32p(bk), T.i32p(ksk)) == 0
    return bk, ksk


@pytest.fixture(scope="module")
def keyset():
    ks = T.KeySet(1024, seed=(7,))
    yield ks
    ks.free()


def test_secret_key_round_trip(keyset, tmp_path):
    L = T.lib()
    f = fopen(tmp_path / "secret.key", "wb")
    L.export_tfheGateBootstrappingSecretKeySet_toFile(f, keyset.ptr)
    libc.fclose(f)
    f = fopen(tmp_path / "secret.key", "rb")
    k2 = L.new_tfheGateBootstrappingSecretKeySet_fromFile(f)
    libc.fclose(f)
    assert k2
    try:
        lk = np.zeros_like(keyset.lwe_key)
        tk = np.zeros_like(keyset.tlwe_key)
        bk = np.zeros_like(keyset.bk)
        ksk = np.zeros_like(keyset.ksk)
        assert L.tfhe_keyset_export_flat(k2, T.i32p(lk), T.i32p(tk), T.i32p(bk), T.i32p(ksk)) == 0
        for got, want in ((lk, keyset.lwe_key), (tk, keyset.tlwe_key), (bk, keyset.bk), (ksk, keyset.ksk)):
            assert np.array_equal(got, want)
        # the imported key decrypts what the original encrypts
        bits = np.array([0, 1, 1, 0, 1], np.int32)
        ct = keyset.encrypt(bits)
        got = np.zeros(5, np.int32)
        assert L.tfhe_decrypt_bits_flat(k2, T.i32p(ct), 5, T.i32p(got)) == 0
        assert np.array_equal(got, bits)
    finally:
        L.delete_gate_bootstrapping_secret_keyset(k2)


def test_cloud_key_round_trip(keyset, tmp_path):
    L = T.lib()
    f = fopen(tmp_path / "cloud.key", "wb")
    L.export_tfheGateBootstrappingCloudKeySet_toFile(f, keyset.cloud_ptr)
    libc.fclose(f)
    size = os.path.getsize(tmp_path / "cloud.key")
    assert size == 8 + (8 + 7 * 4 + 4 * 8) + keyset.bk.nbytes + keyset.ksk.nbytes
    f = fopen(tmp_path / "cloud.key", "rb")
    c2 = L.new_tfheGateBootstrappingCloudKeySet_fromFile(f)
    libc.fclose(f)
    assert c2
    try:
        bk, ksk = flat_cloud(c2, keyset)
        assert np.array_equal(bk, keyset.bk)
        assert np.array_equal(ksk, keyset.ksk)
        fp = T.FlatParams()
        params = ctypes.cast(c2, ctypes.POINTER(ctypes.c_void_p))[0]   # cloud->params
        assert L.tfhe_flat_params_of(params, ctypes.byref(fp)) == 0
        assert fp.as_dict() == keyset.fp.as_dict()
    finally:
        L.delete_gate_bootstrapping_cloud_keyset(c2)


def test_parameter_set_round_trip(keyset, tmp_path):
    L = T.lib()
    for ring in (1024, 2048):
        p = L.new_default_gate_bootstrapping_parameters(110) if ring == 1024 else L.new_gate_bootstrapping_parameters_N2048()
        f = fopen(tmp_path / "p.bin", "wb")
        L.export_tfheGateBootstrappingParameterSet_toFile(f, p)
        libc.fclose(f)
        f = fopen(tmp_path / "p.bin", "rb")
        p2 = L.new_tfheGateBootstrappingParameterSet_fromFile(f)
        libc.fclose(f)
        a, b = T.FlatParams(), T.FlatParams()
        L.tfhe_flat_params_of(p, ctypes.byref(a))
        L.tfhe_flat_params_of(p2, ctypes.byref(b))
        assert a.as_dict() == b.as_dict()
        # alpha values travel too: LweParams {int32 n; double alpha_min, alpha_max}
        lwe = lambda q: ctypes.cast(ctypes.cast(q + 8, ctypes.POINTER(ctypes.c_void_p))[0],
                                    ctypes.POINTER(ctypes.c_double * 3))[0]
        assert list(lwe(p))[1:] == list(lwe(p2))[1:]
        L.delete_gate_bootstrapping_parameters(p)
        L.delete_gate_bootstrapping_parameters(p2)


def test_ciphertext_round_trip(keyset, tmp_path):
    L = T.lib()
    n = keyset.fp.n
    cnt = 20
    src = L.new_gate_bootstrapping_ciphertext_array(cnt, keyset.params_ptr)
    dst = L.new_gate_bootstrapping_ciphertext_array(cnt, keyset.params_ptr)
    rng = np.random.default_rng(9)
    for i in range(cnt):
        L.bootsSymEncrypt(ctypes.byref(src[i]), int(rng.integers(0, 2)), keyset.ptr)
    f = fopen(tmp_path / "ct.bin", "wb")
    for i in range(cnt):
        L.export_gate_bootstrapping_ciphertext_toFile(f, ctypes.byref(src[i]), keyset.params_ptr)
    libc.fclose(f)
    f = fopen(tmp_path / "ct.bin", "rb")
    for i in range(cnt):
        L.import_gate_bootstrapping_ciphertext_fromFile(f, ctypes.byref(dst[i]), keyset.params_ptr)
    libc.fclose(f)
    for i in range(cnt):
        assert [src[i].a[j] for j in range(n)] == [dst[i].a[j] for j in range(n)]
        assert src[i].b == dst[i].b and src[i].current_variance == dst[i].current_variance
        assert L.bootsSymDecrypt(ctypes.byref(src[i]), keyset.ptr) == L.bootsSymDecrypt(ctypes.byref(dst[i]), keyset.ptr)
    L.delete_gate_bootstrapping_ciphertext_array(cnt, src)
    L.delete_gate_bootstrapping_ciphertext_array(cnt, dst)


def test_malformed_files_are_refused(keyset, tmp_path):
    L = T.lib()
    (tmp_path / "junk").write_bytes(b"NOPE" + b"\0" * 64)
    for reader in (L.new_tfheGateBootstrappingSecretKeySet_fromFile, L.new_tfheGateBootstrappingCloudKeySet_fromFile,
                   L.new_tfheGateBootstrappingParameterSet_fromFile):
        f = fopen(tmp_path / "junk", "rb")
        assert not reader(f)
        libc.fclose(f)
        assert L.tfhe_gpu_last_error()
    # a cloud key cut short
    f = fopen(tmp_path / "cloud.key", "wb")
    L.export_tfheGateBootstrappingCloudKeySet_toFile(f, keyset.cloud_ptr)
    libc.fclose(f)
    data = (tmp_path / "cloud.key").read_bytes()
    (tmp_path / "short.key").write_bytes(data[: len(data) // 2])
    f = fopen(tmp_path / "short.key", "rb")
    assert not L.new_tfheGateBootstrappingCloudKeySet_fromFile(f)
    libc.fclose(f)
    # a secret key file is not a cloud key file
    f = fopen(tmp_path / "s.key", "wb")
    L.export_tfheGateBootstrappingSecretKeySet_toFile(f, keyset.ptr)
    libc.fclose(f)
    f = fopen(tmp_path / "s.key", "rb")
    assert not L.new_tfheGateBootstrappingCloudKeySet_fromFile(f)
    libc.fclose(f)



def params_header(n, N, k, l, Bgbit, ks_t, ks_basebit):
    """A TMPS object with the given sizes (the layout tfhe_io.cpp documents)."""
    return (b"TMPS" + struct.pack("<I", 1) + struct.pack("<7i", n, N, k, l, Bgbit, ks_t, ks_basebit)
            + struct.pack("<4d", 2.0**-15, 2.0**-15, 2.0**-25, 2.0**-25))


@pytest.mark.parametrize("sizes", [
    (630, 1024, 1, 3, 7, 1, 31),          # keyswitch base 2^31: must be refused before it is formed
    (630, 1024, 1, 3, 7, 1, 32),
    (4096, 65536, 4, 16, 2, 8, 2),        # within the per-field limits, but a ~430 GB bootstrapping key
    (4096, 65536, 1, 3, 7, 16, 2),        # a ~69 GB keyswitch key
])
def test_crafted_headers_are_refused(tmp_path, sizes):
    """Parameter headers whose keys could not be allocated are refused with NULL and an error
    message, for each reader, before any key storage is allocated."""
    L = T.lib()
    hdr = params_header(*sizes)
    files = [(L.new_tfheGateBootstrappingParameterSet_fromFile, hdr),
             (L.new_tfheGateBootstrappingCloudKeySet_fromFile, b"TMCK" + struct.pack("<I", 1) + hdr),
             (L.new_tfheGateBootstrappingSecretKeySet_fromFile, b"TMSK" + struct.pack("<I", 1) + hdr)]
    for reader, data in files:
        (tmp_path / "crafted").write_bytes(data + b"\0" * 256)
        f = fopen(tmp_path / "crafted", "rb")
        assert not reader(f), sizes
        libc.fclose(f)
        assert L.tfhe_gpu_last_error()
    # the supported sets still read back
    for ok in ((630, 1024, 1, 3, 7, 8, 2), (750, 2048, 1, 3, 8, 10, 2)):
        (tmp_path / "ok").write_bytes(params_header(*ok))
        f = fopen(tmp_path / "ok", "rb")
        p = L.new_tfheGateBootstrappingParameterSet_fromFile(f)
        libc.fclose(f)
        assert p
        L.delete_gate_bootstrapping_parameters(p)

def build_demo(tmp_path):
    exe = tmp_path / "demo_flow"
    libdir = os.path.join(ROOT, "fpga-tfhe_amd", "lib")
    subprocess.run(["gcc", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "c", "demo_flow.c"),
                    "-I", os.path.join(ROOT, "include"), "-L", libdir, "-ltfhe_mi355x", f"-Wl,-rpath,{libdir}"],
                   check=True)
    return exe


def test_demo_client_side(tmp_path):
    """The client writes keys and ciphertexts; the library reads them back and they decrypt to the
    operands (499 and 388, the reference demo's values)."""
    exe = build_demo(tmp_path)
    d = tmp_path / "io"
    d.mkdir()
    subprocess.run([str(exe), "client", str(d), "499", "388"], check=True, timeout=120)
    L = T.lib()
    f = fopen(d / "secret.key", "rb")
    key = L.new_tfheGateBootstrappingSecretKeySet_fromFile(f)
    libc.fclose(f)
    assert key
    params = ctypes.cast(key, ctypes.POINTER(ctypes.c_void_p))[0]
    arr = L.new_gate_bootstrapping_ciphertext_array(32, params)
    f = fopen(d / "cloud.data", "rb")
    for i in range(32):
        L.import_gate_bootstrapping_ciphertext_fromFile(f, ctypes.byref(arr[i]), params)
    libc.fclose(f)
    vals = [sum(L.bootsSymDecrypt(ctypes.byref(arr[16 * w + i]), key) << i for i in range(16)) for w in (0, 1)]
    assert vals == [499, 388]
    L.delete_gate_bootstrapping_ciphertext_array(32, arr)
    L.delete_gate_bootstrapping_secret_keyset(key)
