"""The C++ server API (include/gpu_protocol.h) driven the way the
reference's own callers drive it: tests/cpp/test_gpu_api.cpp, built by
inspire-gpu_amd/Makefile, run as a child process (setup -> gpu_preprocess
-> gpu_setup_server -> gpu_answer / gpu_answer_batch -> extract, caps,
partial batches, reruns, rejected calls, and per-group precomp in the
reference's layout adopted and answering like the arena server)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "inspire-gpu_amd", "build", "test_gpu_api")


def test_cpp_api_program():
    assert os.path.exists(BIN), f"{BIN} is not built: run __graft_entry__.build() (make -C inspire-gpu_amd)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "all C++ API checks passed" in r.stdout
