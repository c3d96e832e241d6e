"""bench.py's command-line contract, checked without a GPU: it measures one
GPU only and says so for --gpus above 1; --help lists the flags the
profiles' commands use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_help_lists_the_profile_flags():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--profile-steps", "--no-cpu", "--batch-pairs",
                 "--warm-pairs", "--no-other", "--copy-too", "--batch-reps", "--no-batch-workers"):
        assert flag in r.stdout, flag
