"""Runs the C++ source-compatibility tests of the stdgpu facade (tests/cpp/test_facade.hip,
test_atomic.hip, test_deque.hip, test_bitset.hip, test_vector.hip and test_reference_fixture.hip, built
by `python -m stdgpu_amd.build`) on the GPU: user kernels call the native device API, host code the
agnostic API, as in the reference's own test suite (test_reference_fixture.hip on the reference's own
fixture types)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


@pytest.mark.parametrize("name", ["test_facade", "test_atomic", "test_deque", "test_bitset", "test_vector",
                                  "test_reference_fixture"])
def test_cpp_facade_cases(name):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), "build the C++ tests first: python -m stdgpu_amd.build"
    try:
        r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=180)
    except subprocess.TimeoutExpired as e:
        out = e.stdout.decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
        pytest.fail(f"{name} did not finish in 180 s; cases reported so far:\n{out}")
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "OK: 0 failure(s)" in r.stdout
