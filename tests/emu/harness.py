"""CPU emulation harness (test infrastructure only).

Points the sdpvibe_amd host code at tests/emu/libsdpvibe_emu.so -- the same HIP kernel sources
compiled for the host against a fake HIP runtime, with AddressSanitizer -- and lets it take CPU
tensors. The product package itself has no CPU path; the redirection lives only in this test
process. Run scripts that use it with the ASan runtime preloaded (see asan_env()).
"""
import ctypes
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "libsdpvibe_emu.so")


def asan_runtime():
    c = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    return c[0] if c else None


def asan_env():
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    # the emulated threads are fibers on their own stacks: no fake stacks for use-after-return
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1:detect_stack_use_after_return=0"
    env["OMP_NUM_THREADS"] = "1"
    return env


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def install():
    import torch
    from sdpvibe_amd import _lib, basecase, batched, limbops, mpf, mpmat, rns, sdp, solver, steplen

    h = ctypes.CDLL(EMU)
    for name, (args, res) in _lib.SIGNATURES.items():
        fn = getattr(h, name)
        fn.argtypes, fn.restype = args, res
    _lib._lib = h

    def emu_ptr(t):
        if t is None:
            return None
        assert not t.is_cuda and t.is_contiguous()
        return ctypes.c_void_p(t.data_ptr())

    def emu_ptr_strided(t):
        assert not t.is_cuda and t.dim() == 4 and t.stride(-1) == 1 and min(t.stride()) >= 0
        return ctypes.c_void_p(t.data_ptr())

    torch.cuda.current_device = lambda: 0
    for mod in (limbops, mpf, mpmat, rns, batched, basecase, solver, steplen, sdp):
        if hasattr(mod, "ptr"):
            mod.ptr = emu_ptr
        if hasattr(mod, "ptr_strided"):
            mod.ptr_strided = emu_ptr_strided
        if hasattr(mod, "stream"):
            mod.stream = lambda: None
        if hasattr(mod, "dev"):
            mod.dev = "cpu"
    _lib.ptr = emu_ptr
    _lib.ptr_strided = emu_ptr_strided
    return h
