"""The surrogate's training sequence on the CPU under AddressSanitizer.

tests/mlp_host_harness.cpp runs the step sequence the device runs
(forcesmith_amd/csrc/mlp_seq.hpp) over host buffers with the device buffers'
growth rule, on the retraining schedule of the screened DE run that faulted
the GPU (minibatches of 1008, 1848, then 2048 rows, screening 672 candidates
in between). The former order of the backward pass (the output pointer read
before the forward pass sized the workspace) writes into the freed buffer when
the 1848-row retrain reallocates it, and AddressSanitizer reports it; the
product order runs clean. The harness also checks the backward pass against
central differences of the loss.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("mlph") / "mlp_host_harness"
    src = os.path.join(ROOT, "tests", "mlp_host_harness.cpp")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", "-fno-omit-frame-pointer",
                    src, "-o", str(out)], check=True)
    return str(out)


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    return subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=120)


def test_product_order_is_clean(harness):
    r = _run(harness, "schedule", "fixed")
    assert r.returncode == 0, r.stderr
    assert "AddressSanitizer" not in r.stderr
    assert "1848-row" in r.stdout and "2048-row" in r.stdout


def test_former_order_writes_freed_workspace(harness):
    r = _run(harness, "schedule", "prefix")
    assert r.returncode != 0
    assert "heap-use-after-free" in r.stderr and "WRITE" in r.stderr
    assert "in head" in r.stderr
    # the first retrain (1008 rows) is clean; the fault comes with the growth
    assert "1008-row" in r.stdout and "1848-row" not in r.stdout


def test_backward_matches_central_differences(harness):
    r = _run(harness, "grad")
    assert r.returncode == 0, r.stdout + r.stderr
