"""bench.py's command line, checked on the CPU: it measures one GPU only and
refuses more with a message, and it exposes the far-field choice of the
precision forms (the FP32 split kernel by default, FP64 on request)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120)


def test_refuses_more_than_one_gpu():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_far_field_choice():
    r = run("--help")
    assert r.returncode == 0
    assert "--far-field {fp32,fp64}" in r.stdout
    r = run("--far-field", "fp16")
    assert r.returncode != 0 and "invalid choice" in r.stderr
