"""The HIP kernels compiled for the host (tests/emu, fake HIP runtime + AddressSanitizer), driven
through the real sdpvibe_amd host code and compared with the oracle. Catches out-of-bounds global
accesses and logic errors on the CPU, before any GPU run.

Every case is its own process; the cases run side by side (longest first) on a few host cores, and
each test waits for its own case only."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

from tests.emu import harness

# longest first, so that the short cases fill in behind the long ones
CASES = ["bucketed", "product_a", "product_b", "product_c", "chol", "tile_wide", "small_round", "linalg", "gram", "bgemm",
         "gemm_regwords", "tile", "carry", "gemm_batched", "gemm_ksplit", "gemm_ksplit_long", "eig", "crt",
         "residues", "gemm_trans", "gemm_small", "bgemm_narrow", "panels", "ksum"]
WORKERS = max(1, min(6, (os.cpu_count() or 2) - 2))


def _run(case):
    return subprocess.run([sys.executable, "tests/emu/run_cases.py", case], env=harness.asan_env(),
                          capture_output=True, text=True, timeout=900, cwd=harness.HERE + "/../..")


@pytest.fixture(scope="module")
def emu_runs():
    if harness.asan_runtime() is None:
        pytest.skip("no ASan runtime in this toolchain")
    harness.build()
    with ThreadPoolExecutor(WORKERS) as ex:
        futures = {c: ex.submit(_run, c) for c in CASES}
        yield futures
        for f in futures.values():
            f.cancel()


@pytest.mark.parametrize("case", CASES)
def test_emulated(emu_runs, case):
    r = emu_runs[case].result()
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
