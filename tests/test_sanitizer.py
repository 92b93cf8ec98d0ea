"""The C ABI under AddressSanitizer and UndefinedBehaviorSanitizer (host code
of libjc_limber_asan.so and the check program jax_cosmo_amd/csrc/check/
jc_abi_check.cpp, built by `make sanitize`; device code unchanged): argument
checks everywhere, and with a GPU every compute entry point and the request
loop with two submitting threads, an injected batch failure and teardown.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "jax_cosmo_amd", "jc_abi_check_asan")


def run_check():
    if not os.path.exists(EXE):
        pytest.skip("the sanitizer build is not there (make -C jax_cosmo_amd/csrc sanitize)")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "abort_on_error=0:detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.returncode == 0
    return r.stdout


def test_abi_host_checks_under_sanitizers():
    import torch

    if torch.cuda.is_available():
        pytest.skip("covered by the GPU run")
    assert "host-only checks, 0 failed" in run_check()


@pytest.mark.gpu
def test_abi_and_server_under_sanitizers():
    out = run_check()
    assert "jc_abi_check: 0 failed" in out
