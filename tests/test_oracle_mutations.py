"""The closed-form checks of tests/pins.py must catch plausible restatement
mistakes.  Each case below plants one mistake in a private copy of
oracle/oracle.c (CPU only), builds it, runs every check against it in a
child process and requires the named checks to fail.  The unmodified
oracle passes all of them (tests/test_oracle_pins.py)."""
import json
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")

# (name, [(text in oracle.c, replacement)], checks that must fail)
MUTATIONS = [
    ("g1_from_other_component",     # the T17 norm built from u1x and u2y
     [("hypotf(div1[i], v1[i])", "hypotf(div1[i], v2[i])")],
     {"t17_one_step", "t17_dual_feasible"}),
    ("p12_divided_by_g2",
     [("p12[i] = q12 / g1;", "p12[i] = q12 / g2;")],
     {"t17_one_step", "t17_dual_feasible"}),
    ("taut_is_tau",                  # tau instead of tau/theta
     [("const float taut = tau / theta;", "const float taut = tau;")],
     {"t17_one_step"}),
    ("error_not_divided_by_n",
     [("error = s / (double)size;", "error = s;"), ("error_f /= (float)size;", "")],
     {"t16_stop_threshold"}),
    ("u_update_without_theta",
     [("u1[p] = v1[p] + theta * div1[p];", "u1[p] = v1[p] + div1[p];")],
     {"t16_u_update"}),
    ("hs_upscale_without_half_pixel",
     [("const float x = ((float)ix - 0.5f) * 0.5f;", "const float x = ((float)ix) * 0.5f;")],
     {"h6_upscale_ramp"}),
    ("hs_upscale_without_scale",
     [("orc_bilinear_at(src, w, h, s, x, y) * scale;", "orc_bilinear_at(src, w, h, s, x, y);")],
     {"h6_upscale_constant", "h6_upscale_ramp"}),
    ("hs_downscale_without_quarter",
     [("out[(size_t)ix + (size_t)iy * ns] = 0.25f * (((a + b) + c) + d);",
       "out[(size_t)ix + (size_t)iy * ns] = 0.5f * (((a + b) + c) + d);")],
     {"h6_downscale_ramp"}),
    ("th_step_sign",                 # d = +lambda*theta*grad where rho > lambda*theta*|grad|^2
     [("d1 = -l_t * wx;", "d1 = l_t * wx;")],
     {"t16_error_term"}),
]


# The same for the FP64 oracle (oracle64.c): each mistake must fail the
# named tests of tests/test_oracle64.py.
MUTATIONS64 = [
    ("fp64_g1_from_other_component",
     [("const double g1 = 1.0 + taut * hypot(d1[i], v1[i]);",
       "const double g1 = 1.0 + taut * hypot(d1[i], v2[i]);")],
     {"test_t17_one_step_closed_form"}),
    ("fp64_p12_divided_by_g2",
     [("p12[i] = (p12[i] + taut * v1[i]) / g1;", "p12[i] = (p12[i] + taut * v1[i]) / g2;")],
     {"test_t17_one_step_closed_form"}),
    ("fp64_taut_is_tau",
     [("const double lt = lambda * theta, taut = tau / theta;",
       "const double lt = lambda * theta, taut = tau;")],
     {"test_t17_one_step_closed_form"}),
    ("fp64_error_not_divided_by_n",
     [("                                       prm->tau, prm->lambda, prm->theta, scratch) /\n"
       "                    (double)n;",
       "                                       prm->tau, prm->lambda, prm->theta, scratch);")],
     {"test_t16_stop_threshold"}),
    ("fp64_u_update_without_theta",
     [("const double a = v1[i] + theta * d1[i], b = v2[i] + theta * d2[i];",
       "const double a = v1[i] + d1[i], b = v2[i] + theta * d2[i];")],
     {"test_t16_u_update_and_error_term"}),
    ("fp64_hs_upscale_without_half_pixel",
     [("bilinear64(src, w, h, (ix - 0.5) * 0.5, (iy - 0.5) * 0.5) * scale;",
       "bilinear64(src, w, h, ix * 0.5, (iy - 0.5) * 0.5) * scale;")],
     {"test_h6_pyramid_steps"}),
    ("fp64_jacobi_alpha_squared",
     [("/ (Ix[p] * Ix[p] + Iy[p] * Iy[p] + alpha);", "/ (Ix[p] * Ix[p] + Iy[p] * Iy[p] + alpha * alpha);")],
     {"test_hs_jacobi_uniform_closed_form"}),
    ("fp64_gaussian_high_side_whole_sample",
     [("if (xr >= nx) xr = 2 * nx - 1 - xr;", "if (xr >= nx) xr = 2 * nx - 2 - xr;")],
     {"test_tvl1_agrees_with_float_oracle"}),
    ("fp64_keys_cubic_term",
     [("x * (3.0 * (v[1] - v[2]) + v[3] - v[0])", "x * (3.0 * (v[1] - v[2]) + v[3] + v[0])")],
     {"test_bicubic_reproduces_quadratics_and_samples"}),
]


def _build(tmp, name, edits, target="oracle.c"):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    for fname in ("oracle.c", "oracle64.c"):
        with open(os.path.join(ORACLE, fname)) as f:
            src = f.read()
        if fname == target:
            for old, new in edits:
                assert src.count(old) == 1, f"{name}: mutation site {old!r} not found exactly once"
                src = src.replace(old, new)
        with open(os.path.join(d, fname), "w") as f:
            f.write(src)
    shutil.copy(os.path.join(ORACLE, "oracle.h"), d)
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-O2", "-fPIC", "-ffp-contract=off",
                    "-fopenmp", "-shared", "-o", lib, os.path.join(d, "oracle.c"),
                    os.path.join(d, "oracle64.c"), "-lm"],
                   check=True, capture_output=True)
    return lib


def _run_pins(lib):
    env = dict(os.environ, AMDFLOW_ORACLE_LIB=lib, OMP_NUM_THREADS="2",
               PYTHONPATH=os.pathsep.join([HERE, ORACLE, os.path.join(ROOT, "sycl-optic-flow_amd")]))
    res = subprocess.run([sys.executable, os.path.join(HERE, "pins.py")], env=env,
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr[-2000:]
    return json.loads(res.stdout.strip().splitlines()[-1])


def _run_fp64_tests(lib):
    """Names of the tests of tests/test_oracle64.py that fail on `lib`."""
    env = dict(os.environ, AMDFLOW_ORACLE_LIB=lib, OMP_NUM_THREADS="2",
               PYTHONPATH=os.pathsep.join([HERE, ORACLE, os.path.join(ROOT, "sycl-optic-flow_amd")]))
    res = subprocess.run([sys.executable, "-m", "pytest", os.path.join(HERE, "test_oracle64.py"),
                          "-q", "--tb=no", "-rf", "-p", "no:cacheprovider", "-p", "no:xdist"],
                         env=env, capture_output=True, text=True, timeout=300, cwd=ROOT)
    failed = set()
    for line in res.stdout.splitlines():
        if line.startswith("FAILED "):
            failed.add(line.split("::")[1].split(" ")[0].split("[")[0])
    return failed


@pytest.fixture(scope="module")
def mutated64(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants64"))
    with ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda m: _build(tmp, m[0], m[1], "oracle64.c"), MUTATIONS64))
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(_run_fp64_tests, libs))
    return {m[0]: r for m, r in zip(MUTATIONS64, results)}


@pytest.mark.parametrize("name,expect", [(m[0], m[2]) for m in MUTATIONS64])
def test_fp64_mutation_is_caught(mutated64, name, expect):
    failed = mutated64[name]
    assert expect <= failed, f"{name}: expected {sorted(expect)} to fail, failed {sorted(failed)}"


def test_fp64_sites_match_the_source():
    with open(os.path.join(ORACLE, "oracle64.c")) as f:
        src = f.read()
    for name, edits, _ in MUTATIONS64:
        for old, _new in edits:
            assert src.count(old) == 1, f"{name}: {old!r}"


@pytest.fixture(scope="module")
def mutated(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    with ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda m: _build(tmp, m[0], m[1]), MUTATIONS))
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(_run_pins, libs))
    return {m[0]: r for m, r in zip(MUTATIONS, results)}


def test_sites_match_the_oracle_source():
    with open(os.path.join(ORACLE, "oracle.c")) as f:
        src = f.read()
    for name, edits, _ in MUTATIONS:
        for old, _new in edits:
            assert src.count(old) == 1, f"{name}: {old!r}"


@pytest.mark.parametrize("name,expect", [(m[0], m[2]) for m in MUTATIONS])
def test_mutation_is_caught(mutated, name, expect):
    res = mutated[name]
    failed = {k for k, v in res.items() if v != "ok"}
    assert expect <= failed, f"{name}: expected {sorted(expect)} to fail, failed {sorted(failed)}"
