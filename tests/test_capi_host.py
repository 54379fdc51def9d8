"""CPU: the C-ABI library loads, exports every function include/tfhe_mi355x.h declares, and its
host-side logic (parameters, key generation, encryption) matches the oracle bit for bit.  No
GPU entry point is called here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "tfhe_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"typedef struct [^;]*?\{.*?\}\s*\w+\s*;", " ", text, flags=re.S)
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name not in ("if", "while", "sizeof"):
            names.add(name)
    return sorted(names)


def test_header_declares_the_upstream_gate_api():
    fns = header_functions()
    for must in ("bootsNAND", "bootsAND", "bootsOR", "bootsXOR", "bootsXNOR", "bootsMUX", "tfhe_bootstrap_FFT",
                 "tfhe_blindRotate_FFT", "lweKeySwitch", "tfhe_gpu_key_upload", "tfhe_gpu_boots_batch"):
        assert must in fns


def test_library_exports_every_declared_function():
    import tfhe_amd as T
    lib = ctypes.CDLL(T.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", T.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(header_functions()) <= exported


def test_integration_guide_indexes_every_declared_function():
    """The entry-point index of INTEGRATION.md names every function the header declares, and names
    nothing the header lacks."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    index = text[text.index("## 5. Entry-point index"):]
    declared = set(header_functions())
    missing = [f for f in sorted(declared) if not re.search(r"`%s`" % re.escape(f), index)]
    assert not missing, missing
    named = set(re.findall(r"`([A-Za-z_][A-Za-z0-9_]*)`", index))
    modules = {"tfhe_io"}   # upstream module names, not functions
    unknown = sorted(n for n in named - modules
                     if n not in declared and re.match(r"(boots|tfhe_|lwe|new_|delete_|export_|import_)", n))
    assert not unknown, unknown


def test_default_parameters_match_reference_constants():
    """kernel/src/include/constants.h:26-36 through the upstream constructor."""
    import tfhe_amd as T
    L = T.lib()
    p = L.new_default_gate_bootstrapping_parameters(110)
    fp = T.FlatParams()
    assert L.tfhe_flat_params_of(p, ctypes.byref(fp)) == 0
    assert fp.as_dict() == dict(n=630, N=1024, k=1, l=3, Bgbit=7, ks_t=8, ks_basebit=2)
    L.delete_gate_bootstrapping_parameters(p)
    p2 = L.new_gate_bootstrapping_parameters_N2048()
    assert L.tfhe_flat_params_of(p2, ctypes.byref(fp)) == 0
    assert fp.as_dict() == dict(n=750, N=2048, k=1, l=3, Bgbit=8, ks_t=10, ks_basebit=2)
    L.delete_gate_bootstrapping_parameters(p2)


def test_product_numerics_match_oracle():
    import tfhe_amd as T
    L = T.lib()
    rng = np.random.default_rng(9)
    for x in list(rng.integers(-2**31, 2**31, 200)) + [0, -1, 2**31 - 1, -2**31]:
        for m in (8, 2048, 4096):
            assert L.modSwitchFromTorus32(int(x), m) == O.modswitch_from(int(x), m)
            assert L.approxPhase(int(x), m) == O.approx_phase(int(x), m)
    for mu in range(-8, 9):
        assert L.modSwitchToTorus32(mu, 8) == O.modswitch_to(mu, 8)
    for d in (0.1, -0.3, 2.0 ** -15, 0.75, -1.5):
        assert L.dtot32(d) == O.dtot32(d)


def test_product_keygen_matches_oracle_bit_exact():
    """Host key generation of the product (upstream draw order) equals the oracle's, seed 0."""
    import tfhe_amd as T
    ks = T.KeySet(1024, seed=(0,))
    K = O.keygen(O.PARAMS_1024, (0,))
    try:
        for name in ("lwe_key", "tlwe_key", "ksk", "bk"):
            assert np.array_equal(getattr(ks, name), getattr(K, name)), name
        bits = np.array([1, 0, 0, 1, 1], np.int32)
        assert np.array_equal(ks.encrypt(bits), O.boots_encrypt(O.PARAMS_1024, bits, K.lwe_key))
        ct = O.boots_encrypt(O.PARAMS_1024, bits, K.lwe_key)
        assert np.array_equal(ks.decrypt(ct), bits)
    finally:
        K.free()
        ks.free()


def test_flat_batch_rejects_bad_arguments_without_gpu_work():
    """Argument checks run before any device call: a null key is an error, count 0 is a no-op."""
    import tfhe_amd as T
    L = T.lib()
    z = np.zeros((1, 631), np.int32)
    rc = L.tfhe_gpu_gate_batch(None, 0, 0, T.i32p(z), T.i32p(z), T.i32p(z), None, 1)
    assert rc != 0


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
