"""The C++ drop-in API (include/H2_3D_Tree.hpp, kernel_Types.hpp, compute.hpp)
driven by a PBBFMM3D-style caller (tests/cpp/dropin_example.cpp)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "dropin_example")


def _run(*args):
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build())"
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)


def test_precompute_runs_without_gpu():
    import pbbfmm3d_amd as pbb
    if pbb.capi.lib().pbb_device_count() > 0:
        pytest.skip("a device is visible")
    r = _run("-N", "2000", "-level", "3", "-n", "4")
    assert r.returncode == 3, r.stdout + r.stderr
    assert "ranks" in r.stdout and "no HIP device" in r.stdout


@pytest.mark.gpu
def test_dropin_example_on_gpu(gpu):
    r = _run("-N", "20000", "-level", "4", "-n", "5", "-m", "2", "-eps", "1e-7")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
