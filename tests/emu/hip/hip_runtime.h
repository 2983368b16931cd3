// Host emulation of the small HIP surface used by sdpvibe_amd/csrc (CPU testing only).
// Each kernel block runs as blockDim cooperative fibers with real barriers (a deadlock is
// reported); MFMA and cross-lane shuffles are emulated per 64-lane wave. Built with
// AddressSanitizer, it turns an out-of-bounds global access of a kernel into a host report
// instead of a GPU fault.
#pragma once
#include <cmath>
#include <cstdlib>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static  // static LDS arrays: one copy shared by the block being run
#define __align__(x) __attribute__((aligned(x)))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef void* hipStream_t;
typedef int hipError_t;
#define hipSuccess 0
#define hipFuncAttributeMaxDynamicSharedMemorySize 8
inline hipError_t hipGetLastError() { return hipSuccess; }
// one emulated device; a __device__ array is a host array, its "device address" its host address
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return hipSuccess; }
#define hipHostMallocDefault 0
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { *p = malloc(n); return *p ? hipSuccess : 1; }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }

extern thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;
void emu_syncthreads();
inline void __syncthreads() { emu_syncthreads(); }
using std::isfinite;
using std::frexp;
using std::ldexp;
inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline long long atomicMax(long long* p, long long v) {
    long long cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
long long emu_shfl_xor64(long long v, int mask);
long long emu_shfl_idx64(long long v, int src);
inline int __shfl(int v, int src, int width = 64) { (void)width; return (int)emu_shfl_idx64(v, src); }
inline long long __shfl_xor(long long v, int mask, int width = 64) { (void)width; return emu_shfl_xor64(v, mask); }
inline long __shfl_xor(long v, int mask, int width = 64) { (void)width; return (long)emu_shfl_xor64(v, mask); }
inline int __shfl_xor(int v, int mask, int width = 64) { (void)width; return (int)emu_shfl_xor64(v, mask); }
inline float __shfl_xor(float v, int mask, int width = 64) {
    (void)width;
    int b;
    std::memcpy(&b, &v, 4);
    b = (int)emu_shfl_xor64(b, mask);
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}
inline double __shfl_xor(double v, int mask, int width = 64) {
    (void)width;
    long long b;
    std::memcpy(&b, &v, 8);
    b = emu_shfl_xor64(b, mask);
    double r;
    std::memcpy(&r, &b, 8);
    return r;
}
struct emu_u2 {
    unsigned v[2];
    unsigned operator[](int i) const { return v[i]; }
};
inline emu_u2 __builtin_amdgcn_permlane32_swap(unsigned x, unsigned y, bool, bool) {
    const unsigned tx = (unsigned)emu_shfl_xor64((long long)x, 32), ty = (unsigned)emu_shfl_xor64((long long)y, 32);
    return (threadIdx.x & 63) < 32 ? emu_u2{{x, tx}} : emu_u2{{ty, y}};
}
void emu_wave_barrier();
unsigned long long emu_ballot(int pred);
#define __ballot(p) emu_ballot((int)(p))
inline unsigned __builtin_amdgcn_perm(unsigned s0, unsigned s1, unsigned sel) {
    unsigned long long v = ((unsigned long long)s0 << 32) | s1;   // bytes 0-3 from s1, 4-7 from s0
    unsigned r = 0;
    for (int b = 0; b < 4; b++) {
        unsigned q = (sel >> (8 * b)) & 0xffu, byte;
        if (q < 8) byte = (unsigned)(v >> (8 * q)) & 0xffu;
        else if (q == 12) byte = 0;
        else if (q == 13) byte = 0xff;
        else byte = 0;  // sign-replicating selectors are not used by the kernels
        r |= byte << (8 * b);
    }
    return r;
}
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)

typedef int emu_v4i __attribute__((ext_vector_type(4)));
typedef int emu_v16i __attribute__((ext_vector_type(16)));
emu_v16i emu_mfma_i32_32x32x32_i8(emu_v4i a, emu_v4i b, emu_v16i c, int, int, int);
#define __builtin_amdgcn_mfma_i32_32x32x32_i8 emu_mfma_i32_32x32x32_i8

// Direct global -> LDS load of 16 bytes per lane to ldsbase + 16 * lane; vmcnt waits are no-ops
// (the emulated load completes at once).
#define SDPV_LDS_DMA16(gptr, ldsbase) \
    std::memcpy((unsigned char*)(ldsbase) + 16 * (threadIdx.x & 63), (const void*)(gptr), 16)
#define SDPV_WAIT_VMCNT(n) ((void)(n))
#define __builtin_amdgcn_s_setprio(p) ((void)(p))
#define __builtin_amdgcn_readfirstlane(x) (x)

// Dynamic shared memory of the block being run (blocks run one at a time).
extern unsigned char* emu_dynamic_lds;
#define SDPV_DYNAMIC_LDS(type, name) type* name = reinterpret_cast<type*>(emu_dynamic_lds)
void emu_run_grid(dim3 grid, dim3 block, size_t lds_bytes, const std::function<void()>& body);

#define hipLaunchKernelGGL(kernel, grid, block, ldsb, stream, ...) \
    emu_run_grid((grid), (block), (ldsb), [&]() { kernel(__VA_ARGS__); })
