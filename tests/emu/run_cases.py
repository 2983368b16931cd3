"""Runs sdpvibe_amd operations through the emulated kernels (ASan) and compares them with the
oracle. Usage: python tests/emu/run_cases.py <case> [<case> ...]   (run with asan_env())."""
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from tests.emu import harness  # noqa: E402

harness.install()
from oracle import oracle as O  # noqa: E402
from sdpvibe_amd import mpf as F  # noqa: E402
from sdpvibe_amd import mpmat  # noqa: E402
from tests.intutil import EZ, rand_mpf_ints, to_planes  # noqa: E402


def gemm_case(L, nb, M, K, N, transA, transB, seed=0):
    rng = random.Random(seed)
    shpA = (nb, K, M) if transA else (nb, M, K)
    shpB = (nb, N, K) if transB else (nb, K, N)
    a, ea = rand_mpf_ints(rng, int(np.prod(shpA)), L, 60)
    b, eb = rand_mpf_ints(rng, int(np.prod(shpB)), L, 60)
    Am = to_planes(a, L).reshape((L,) + shpA)
    Bm = to_planes(b, L).reshape((L,) + shpB)
    Ae = np.array(ea, np.int64).reshape(shpA)
    Be = np.array(eb, np.int64).reshape(shpB)
    Ae[np.asarray(a).reshape(shpA) == 0] = EZ
    Be[np.asarray(b).reshape(shpB) == 0] = EZ
    C = F.gemm(F.MPF(torch.from_numpy(Am), torch.from_numpy(Ae)), F.MPF(torch.from_numpy(Bm), torch.from_numpy(Be)),
               transA, transB)
    rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be, transA, transB)
    ok = np.array_equal(C.m.numpy(), rm) and np.array_equal(C.e.numpy(), re_)
    print(f"gemm L={L} nb={nb} {M}x{K}x{N} tA={transA} tB={transB}: {'ok' if ok else 'MISMATCH'}", flush=True)
    return ok


CASES = {
    "gemm_small": lambda: gemm_case(1, 1, 3, 4, 5, False, False),
    "gemm_trans": lambda: all(gemm_case(2, 1, 5, 7, 6, ta, tb, 3) for ta in (False, True) for tb in (False, True)),
    "gemm_batched": lambda: gemm_case(3, 2, 33, 20, 40, False, True, 4),
}

def gpu_test_gemm_cases():
    """the exact inputs of tests/test_gpu_kernels.py::test_mpf_gemm"""
    from tests.test_gpu_kernels import _gemm_case
    ok = True
    for (L, nb, M, K, N) in [(1, 1, 3, 4, 5), (2, 1, 33, 70, 31), (10, 2, 40, 33, 70)]:
        for ta in (False, True):
            for tb in (False, True):
                Am, Ae, Bm, Be = _gemm_case(L, nb, M, K, N, ta, tb, L * 7919 + M + 2 * ta + tb)
                rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be, ta, tb)
                C = F.gemm(F.MPF(torch.from_numpy(Am), torch.from_numpy(Ae)),
                           F.MPF(torch.from_numpy(Bm), torch.from_numpy(Be)), ta, tb)
                good = np.array_equal(C.m.numpy(), rm) and np.array_equal(C.e.numpy(), re_)
                print(f"gpu-test gemm L={L} nb={nb} {M}x{K}x{N} tA={ta} tB={tb}: {'ok' if good else 'MISMATCH'}",
                      flush=True)
                ok &= good
    return ok


CASES["gpu_gemm"] = gpu_test_gemm_cases


if __name__ == "__main__":  # noqa: E305
    ok = True
    for c in sys.argv[1:] or list(CASES):
        ok &= bool(CASES[c]())
    sys.exit(0 if ok else 1)

