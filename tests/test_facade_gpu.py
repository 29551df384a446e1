"""C++ RobustSolver facade tests (tests/cpp/*.cpp), run on the GPU.

Each binary mirrors cases of the reference's C++ tests (testPcm,
testPcmSimple, testPcmDoOptimize, testGnc, testLoadGraph, testTemporary,
testMultiRobot, testIgnorePrefix, testRemoveLastLoopClosure, testLandmark,
testMultirobotFrameAlign) through
the facade library libkimera_rpgo.so, which drives the HIP kernels through
the C ABI only.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")
CASES = ["test_pcm", "test_pcm_simple", "test_robust_solver", "test_multirobot"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_facade_binary(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} missing: run __graft_entry__.build()"
    env = dict(os.environ, RPGO_DATASET_PATH=os.path.join(ROOT, "tests", "golden", "data"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, f"{name} failed:\n{r.stdout}\n{r.stderr}"
