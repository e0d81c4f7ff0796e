"""bench.py's command line on the CPU: it measures one GPU and says so for
any other count, before touching a device; its options parse."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "8")
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_legs():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--size", "--only", "--only-c2-series",
                "--only-c3-series", "--no-cpu-baseline"):
        assert opt in r.stdout
