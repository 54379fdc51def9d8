"""GPU: the client -> server -> client demo flow through key and ciphertext files
(tests/c/demo_flow.c, in the shape of the reference's alice.c / cloud.c / verify.c): the server
reads the cloud key from its file and evaluates 16 upstream bootsOR gates on the GPU; the client
decrypts OR(499, 388) = 503, the reference demo's answer."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

from test_tfhe_io_host import build_demo  # noqa: E402


def test_demo_flow_through_files(tmp_path):
    exe = build_demo(tmp_path)
    d = tmp_path / "io"
    d.mkdir()
    print("[tfhe_io] client", flush=True)
    subprocess.run([str(exe), "client", str(d), "499", "388"], check=True, timeout=120)
    print("[tfhe_io] server (GPU)", flush=True)
    subprocess.run([str(exe), "server", str(d)], check=True, timeout=120)
    print("[tfhe_io] verify", flush=True)
    out = subprocess.run([str(exe), "verify", str(d)], check=True, timeout=120, capture_output=True, text=True)
    assert out.stdout.strip() == "503"
    assert os.path.getsize(d / "answer.data") == 16 * (8 + 4 + 631 * 4 + 8)
