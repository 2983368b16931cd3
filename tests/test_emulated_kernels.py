"""The HIP kernels compiled for the host (tests/emu, fake HIP runtime + AddressSanitizer), driven
through the real sdpvibe_amd host code and compared with the oracle. Catches out-of-bounds global
accesses and logic errors on the CPU, before any GPU run."""
import subprocess
import sys

import pytest

from tests.emu import harness


@pytest.fixture(scope="module")
def emu_built():
    if harness.asan_runtime() is None:
        pytest.skip("no ASan runtime in this toolchain")
    harness.build()


@pytest.mark.parametrize("case", ["gemm_small", "gemm_trans", "gemm_batched", "bgemm", "tile", "chol", "crt", "gram", "linalg", "eig", "product", "residues", "carry", "small_round", "bucketed", "gemm_regwords", "gemm_ksplit"])
def test_emulated(emu_built, case):
    r = subprocess.run([sys.executable, "tests/emu/run_cases.py", case], env=harness.asan_env(),
                       capture_output=True, text=True, timeout=900, cwd=harness.HERE + "/../..")
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
