"""bench.py's command line, on the CPU.

The bench measures one GPU: asked for more it must exit with status 2 and a
message, before it loads the library or touches a device.
"""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in (2, 8):
        run = _bench("--gpus", str(n), "--steps", "1", "--warmup", "0")
        assert run.returncode == 2, run.stderr
        assert f"--gpus {n} is not supported" in run.stderr
        assert run.stdout == ""


def test_help_lists_the_contract_flags():
    run = _bench("--help")
    assert run.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in run.stdout
