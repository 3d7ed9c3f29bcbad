"""The C ABI from C++: tests/cpp/abi_hbn, compiled against include/edus_gpu.h
as the reference-side binding would be (INTEGRATION.md's sketch), runs the
reference's hBN ci-test through the low-level context (edus_create,
edus_set_state, edus_observables before each edus_rk4_steps step).  Its
Time, Population and Velocity rows must match the reference's committed
outputs (ci-test/outputs/hBN) under compare.py's 1e-10 rule."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "build", "abi_hbn")


def test_cpp_caller_reproduces_hbn_fixture(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/build/abi_hbn is not built (__graft_entry__.build())"
    print("case: hBN 50x50, 102 RK4 steps through the C ABI from C++", flush=True)
    r = subprocess.run([BIN, os.path.join(GOLDEN, "models", "hBN_gap7.25eV_a2.5A"), str(tmp_path)],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    for name in ("Population.txt", "Velocity.txt"):
        got = np.loadtxt(tmp_path / name)
        gold = np.loadtxt(os.path.join(GOLDEN, "hBN", name))
        assert got.shape == gold.shape, name
        bad = (np.abs(got - gold) > 1e-10) & ((np.abs(got) > 1e-10) | (np.abs(gold) > 1e-10))
        print(name, "max abs diff", np.abs(got - gold).max(), flush=True)
        assert not bad.any(), name
    t = np.loadtxt(tmp_path / "Time.txt")
    assert np.allclose(t, np.loadtxt(os.path.join(GOLDEN, "hBN", "Time.txt")), rtol=5e-6, atol=1e-12)


SEAMS = os.path.join(ROOT, "tests", "cpp", "build", "seams_hbn")


def test_cpp_seam_classes(tmp_path):
    """include/edus_operator.hpp's Operator<T> and FourierTransform from C++
    on the hBN set-up (tests/cpp/seams_hbn.cpp): go_to_R against a direct DFT,
    the round trips, go_to_bloch against edus_get_state_bloch, copies."""
    assert os.path.exists(SEAMS), "tests/cpp/build/seams_hbn is not built (__graft_entry__.build())"
    print("case: hBN 50x50, 30 RK4 steps, then the seam classes from C++", flush=True)
    r = subprocess.run([SEAMS, os.path.join(GOLDEN, "models", "hBN_gap7.25eV_a2.5A")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "all checks passed" in r.stdout
