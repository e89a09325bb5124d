"""The C-ABI's host side under AddressSanitizer (CPU only): the library is
rebuilt with -fsanitize=address on its host code, and the host-side ABI tests
(argument validation, workspace sizing and offsets, view preparation, tile
maps, descriptor checks) run against it in a child process with the
sanitizer's runtime preloaded; any invalid host memory access aborts the
child and fails this test."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        return None
    r = subprocess.run([hipcc, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_host_abi_under_address_sanitizer():
    runtime = _asan_runtime()
    if runtime is None:
        pytest.skip("hipcc or the AddressSanitizer runtime is not available")
    from dmgs_amd import build as B
    lib = B.build(asan=True)
    env = dict(os.environ, DMGS_LIBRARY=lib, LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        "tests/test_native_abi.py", "tests/test_batched_host.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr, tail
    assert " passed" in r.stdout, tail
