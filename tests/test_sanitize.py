"""AddressSanitizer + UndefinedBehaviorSanitizer runs (CPU) of the library's host layer and of the oracle.

tests/sanitize/Makefile builds, with -fsanitize=address,undefined and no recovery, the product's host
sources threepoint_amd/csrc/tp_cubature.cpp and tp_io.cpp with a driver (host_sanitize.cpp: the cubature on
analytic integrands in 1-4 dimensions, every error norm, maxEval cuts; the readers on edge-case files) and
the oracle oracle/tp_oracle.cpp with a driver (oracle_sanitize.cpp: tables, the n(z) lookup at its last
sample, B_kappa at squeezed and zero-side triangles, a <Map^3> cubature, the Gamma integrands, power spectra,
<Map^2>, the covariance pieces, the tabulated P(k)).  A sanitizer finding aborts the driver.  The class
facade (tp_classes.cpp) calls the HIP library on every method, so it runs only in the GPU tests.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(SAN, "_build")


@pytest.mark.parametrize("driver", ["host_sanitize", "oracle_sanitize"])
def test_sanitized_run_is_clean(built, driver, tmp_path):
    env = dict(os.environ, TMPDIR=str(tmp_path), OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(built, driver)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
