"""bench.py's command line (no GPU needed): more than one GPU is refused with
a message and exit status 2 before anything touches a device, and the
options the driver passes are accepted by the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "3", "--warmup", "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr
    assert out.stdout == ""


def test_help_lists_contract_options():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--mode"):
        assert opt in out.stdout
