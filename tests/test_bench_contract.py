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


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_hash_ignores_comments_not_code():
    """The measured HBM traffic is reported only for the kernel code it was
    measured on: comments and layout do not count, code and strings do."""
    b = _bench_module()
    base = 'int f(int a) { return a * 2; }  // twice\n/* note */ const char* s = "a // b";\n'
    same = '/* header */\nint f(int a)\n{\n    return a * 2;   // doubled\n}\nconst char* s = "a // b";'
    assert b._code_only(base) == b._code_only(same)
    assert b._code_only(base) != b._code_only(base.replace("a * 2", "a * 3"))
    assert b._code_only(base) != b._code_only(base.replace('"a // b"', '"a b"'))
    assert all(os.path.exists(os.path.join(ROOT, f)) for f in b.TRAFFIC_SOURCES)
    assert len(b._sources_hash()) == 16
