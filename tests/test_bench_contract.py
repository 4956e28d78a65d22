"""bench.py's command-line contract, checked without a GPU: one GPU only
(more exits with a message before any device work) and its options parse."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=60, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    for n in ("2", "8"):
        r = _run("--gpus", n)
        assert r.returncode == 2
        assert r.stdout == ""
        assert f"--gpus {n}" in r.stderr and "one GPU" in r.stderr


def test_help_names_the_contract_options():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu-baseline", "--copy-probe"):
        assert opt in r.stdout
