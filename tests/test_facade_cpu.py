"""CPU checks of the C++ facade: libkimera_rpgo.so loads and exports the
RobustSolver / Pcm / dataset API that include/kimera_rpgo_amd/*.h declares,
and the CPU-only C++ test binary (g2o I/O, pose algebra) passes."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kimera-rpgo_amd", "lib", "libkimera_rpgo.so")

API = [
    "kimera_rpgo_amd::RobustSolver::RobustSolver(kimera_rpgo_amd::RobustSolverParams const&)",
    "kimera_rpgo_amd::RobustSolver::update(",
    "kimera_rpgo_amd::RobustSolver::forceUpdate(",
    "kimera_rpgo_amd::RobustSolver::loadGraph(",
    "kimera_rpgo_amd::RobustSolver::addGraph(",
    "kimera_rpgo_amd::RobustSolver::getNumLC() const",
    "kimera_rpgo_amd::RobustSolver::getNumLCInliers() const",
    "kimera_rpgo_amd::RobustSolver::removeLastLoopClosure(char, char)",
    "kimera_rpgo_amd::RobustSolver::removeLastLoopClosure()",
    "kimera_rpgo_amd::RobustSolver::ignorePrefix(char)",
    "kimera_rpgo_amd::RobustSolver::revivePrefix(char)",
    "kimera_rpgo_amd::RobustSolver::getIgnoredPrefixes() const",
    "kimera_rpgo_amd::RobustSolver::removePriorFactorsWithPrefix(",
    "kimera_rpgo_amd::RobustSolver::saveData(",
    "kimera_rpgo_amd::RobustSolver::updateTempFactorsValues(",
    "kimera_rpgo_amd::RobustSolver::replaceTempFactorsValues(",
    "kimera_rpgo_amd::RobustSolver::clearTempFactorsValues()",
    "kimera_rpgo_amd::Pcm::Pcm(int,",
    "kimera_rpgo_amd::Pcm::removeOutliers(",
    "kimera_rpgo_amd::load3D(",
    "kimera_rpgo_amd::load2D(",
    "kimera_rpgo_amd::writeG2o(",
]


def test_facade_library_exports_api():
    assert os.path.exists(LIB), "run __graft_entry__.build()"
    ctypes.CDLL(LIB)  # resolves librpgo_hip.so through its $ORIGIN rpath
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    missing = [a for a in API if a not in out]
    assert not missing, missing


def test_dataset_binary_cpu():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_dataset")
    env = dict(os.environ, RPGO_DATASET_PATH=os.path.join(ROOT, "tests", "golden", "data"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
