"""CPU model of the work-sharing protocol of the persistent engines (task reservation, ready-flag
publication, busy/idle termination, bounded waits) on many host threads, and an exhaustive check
of the split arithmetic the GPU engines compile from the same header (ws_split.h). The model is
tests/ws_protocol_model.cpp; it stops at its first failure."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "ws_protocol_model.cpp")
INC = os.path.join(ROOT, "accmultiwayjoins_amd", "csrc")


def _build(tmp_path, extra):
    exe = str(tmp_path / "ws_model")
    subprocess.check_call(["g++", "-std=c++17", "-pthread", "-I", INC, *extra, SRC, "-o", exe])
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_ws_protocol_model(tmp_path):
    exe = _build(tmp_path, ["-O2", "-Wall"])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    assert out.stdout.strip().endswith("OK")
    assert "children once each" in out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_ws_protocol_model_tsan(tmp_path):
    try:
        exe = _build(tmp_path, ["-O1", "-g", "-fsanitize=thread"])
    except subprocess.CalledProcessError:
        pytest.skip("ThreadSanitizer is not available to g++ here")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "ThreadSanitizer" not in out.stderr, out.stderr[-4000:]
    assert out.stdout.strip().endswith("OK")
