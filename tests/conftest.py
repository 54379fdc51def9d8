"""Shared fixtures.  Tests marked `gpu` need a GPU and the built library; the rest run on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fpga-tfhe_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def orc_keys_1024():
    """Oracle key set, seed 0 (CPU-only tests)."""
    k = O.keygen(O.PARAMS_1024, (0,))
    yield k
    k.free()


@pytest.fixture(scope="session")
def gpu_env():
    """Product key set (seed 0), its device key, and the same arrays wrapped for the oracle."""
    import tfhe_amd as T
    T.init(0)
    print(f"[gpu_env] device: {T.device_info()}", flush=True)
    ks = T.KeySet(1024, seed=(0,))
    dk = T.DeviceKey.from_keyset(ks)
    keys = O.Keys(O.PARAMS_1024, ks.lwe_key, ks.tlwe_key, ks.ksk, ks.bk)
    yield ks, dk, keys
    dk.free()
    keys.free()
    ks.free()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
