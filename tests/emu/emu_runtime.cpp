// Host runtime for the HIP emulation (see hip/hip_runtime.h). Blocks run one at a time.
#include "hip/hip_runtime.h"
#include <cstdio>
#include <cstdlib>
#include <memory>

thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;
static constexpr size_t LDS_MAX = 160 * 1024;
unsigned char* emu_dynamic_lds = nullptr;

namespace {
struct Wave {
    std::unique_ptr<std::barrier<>> bar;
    emu_v4i a[64], b[64];
    long long x[64];
};
std::unique_ptr<std::barrier<>> g_block;
std::vector<Wave> g_waves;
size_t g_lds_bytes = 0;
}  // namespace

void emu_syncthreads() { g_block->arrive_and_wait(); }

static Wave& my_wave() { return g_waves[threadIdx.x / 64]; }

void emu_wave_barrier() { my_wave().bar->arrive_and_wait(); }

unsigned long long emu_ballot(int pred) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = pred ? 1 : 0;
    w.bar->arrive_and_wait();
    unsigned long long m = 0;
    for (int l = 0; l < 64; l++)
        if (w.x[l]) m |= 1ull << l;
    w.bar->arrive_and_wait();
    return m;
}

long long emu_shfl_xor64(long long v, int mask) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = v;
    w.bar->arrive_and_wait();
    long long r = w.x[(lane ^ mask) & 63];
    w.bar->arrive_and_wait();
    return r;
}

long long emu_shfl_idx64(long long v, int src) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = v;
    w.bar->arrive_and_wait();
    long long r = w.x[src & 63];
    w.bar->arrive_and_wait();
    return r;
}

// D = A B + C with the fragment mapping the kernels assume (verified on the GPU by
// sdpv_mfma_selftest): A row = lane%32 over bytes of lanes (row, row+32); B column = lane%32.
emu_v16i emu_mfma_i32_32x32x32_i8(emu_v4i a, emu_v4i b, emu_v16i c, int, int, int) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.a[lane] = a;
    w.b[lane] = b;
    w.bar->arrive_and_wait();
    emu_v16i d = c;
    for (int j = 0; j < 16; j++) {
        int row = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3), col = lane & 31;
        long long s = 0;
        for (int h = 0; h < 2; h++) {
            const int8_t* pa = (const int8_t*)&w.a[row + 32 * h];
            const int8_t* pb = (const int8_t*)&w.b[col + 32 * h];
            for (int u = 0; u < 16; u++) s += (long long)pa[u] * pb[u];
        }
        d[j] += (int)s;
    }
    w.bar->arrive_and_wait();
    return d;
}

void emu_run_grid(dim3 grid, dim3 block, size_t lds_bytes, const std::function<void()>& body) {
    if (lds_bytes > LDS_MAX) {
        fprintf(stderr, "emu: dynamic LDS %zu exceeds %zu\n", lds_bytes, LDS_MAX);
        abort();
    }
    unsigned nt = block.x * block.y * block.z;
    g_lds_bytes = lds_bytes;
    // exactly the requested size on the heap, so AddressSanitizer reports any access past it
    std::unique_ptr<unsigned char[]> lds(new unsigned char[lds_bytes ? lds_bytes : 16]);
    emu_dynamic_lds = lds.get();
    for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
            for (unsigned bx = 0; bx < grid.x; bx++) {
                memset(lds.get(), 0xA5, lds_bytes);   // LDS starts as garbage
                g_block = std::make_unique<std::barrier<>>(nt);
                unsigned nw = (nt + 63) / 64;
                g_waves = std::vector<Wave>(nw);
                for (unsigned w = 0; w < nw; w++)
                    g_waves[w].bar = std::make_unique<std::barrier<>>(std::min(64u, nt - 64 * w));
                std::vector<std::thread> th;
                th.reserve(nt);
                for (unsigned t = 0; t < nt; t++)
                    th.emplace_back([&, t]() {
                        threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                        blockIdx = dim3(bx, by, bz);
                        blockDim = block;
                        gridDim = grid;
                        body();
                        g_waves[t / 64].bar->arrive_and_drop();
                        g_block->arrive_and_drop();
                    });
                for (auto& x : th) x.join();
            }
}
