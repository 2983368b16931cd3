// Host runtime for the HIP emulation (see hip/hip_runtime.h). Blocks run one at a time; the
// threads of a block are fibers on one OS thread, switched cooperatively at barriers (block,
// wave, and the wave-wide operations built on them): a switch saves and restores the callee-saved
// registers and the stack pointer (x86-64), with no system call, and the scheduler resumes only
// fibers whose barrier has completed.
#include "hip/hip_runtime.h"
#include <sanitizer/asan_interface.h>
#include <sys/mman.h>
#include <cstdio>
#include <cstdlib>
#include <memory>

thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;
static constexpr size_t LDS_MAX = 160 * 1024;
unsigned char* emu_dynamic_lds = nullptr;

// emu_ctx_switch(save, load): pushes the callee-saved registers and the MXCSR / x87 control words,
// stores the stack pointer in *save, switches to the stack `load` and pops the same frame from it.
extern "C" void emu_ctx_switch(void** save, void* load);
asm(R"(
    .text
    .globl emu_ctx_switch
    .type emu_ctx_switch, @function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_ctx_switch, .-emu_ctx_switch
)");

namespace {
struct Barrier {
    int expected = 0, count = 0;
    unsigned gen = 0;
};
struct Wave {
    Barrier bar;
    emu_v4i a[64], b[64];
    long long x[64];
};
struct Fiber {
    void* sp = nullptr;
    Barrier* wait = nullptr;   // blocked until wait->gen != wait_gen
    unsigned wait_gen = 0;
    bool done = false;
};
constexpr size_t STACK = 1 << 20;   // bytes per fiber stack (touched pages only)
Barrier g_block;
std::vector<Wave> g_waves;
std::vector<Fiber> g_fibers;
char* g_stacks = nullptr;
size_t g_nstacks = 0;
void* g_sched_sp = nullptr;
unsigned g_cur = 0;
const std::function<void()>* g_body = nullptr;

void yield() { emu_ctx_switch(&g_fibers[g_cur].sp, g_sched_sp); }

void wait(Barrier& b) {
    const unsigned my = b.gen;
    if (++b.count == b.expected) {
        b.count = 0;
        b.gen++;
        return;
    }
    Fiber& f = g_fibers[g_cur];
    f.wait = &b;
    f.wait_gen = my;
    yield();
    f.wait = nullptr;
}

// a finished thread leaves the barrier: the waiting ones complete it when it was the last one missing
void drop(Barrier& b) {
    b.expected--;
    if (b.count > 0 && b.count == b.expected) {
        b.count = 0;
        b.gen++;
    }
}

void fiber_entry() {
    (*g_body)();
    drop(g_waves[g_cur / 64].bar);
    drop(g_block);
    g_fibers[g_cur].done = true;
    yield();   // never resumed
    abort();
}

void* new_stack(unsigned t) {
    char* top = g_stacks + (size_t)(t + 1) * STACK;   // 16-byte aligned
    __asan_unpoison_memory_region(top - STACK, STACK);
    void** p = (void**)top;
    *--p = nullptr;                                     // return slot of fiber_entry (never used)
    *--p = (void*)&fiber_entry;                         // where the first switch returns to
    for (int r = 0; r < 6; r++) *--p = nullptr;         // rbp rbx r12 r13 r14 r15
    --p;
    unsigned mx;
    unsigned short cw;
    asm volatile("stmxcsr %0" : "=m"(mx));
    asm volatile("fnstcw %0" : "=m"(cw));
    memcpy(p, &mx, 4);
    memcpy((char*)p + 4, &cw, 2);
    return p;
}
}  // namespace

void emu_syncthreads() { wait(g_block); }

static Wave& my_wave() { return g_waves[threadIdx.x / 64]; }

void emu_wave_barrier() { wait(my_wave().bar); }

unsigned long long emu_ballot(int pred) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = pred ? 1 : 0;
    wait(w.bar);
    unsigned long long m = 0;
    for (int l = 0; l < 64; l++)
        if (w.x[l]) m |= 1ull << l;
    wait(w.bar);
    return m;
}

long long emu_shfl_xor64(long long v, int mask) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = v;
    wait(w.bar);
    long long r = w.x[(lane ^ mask) & 63];
    wait(w.bar);
    return r;
}

long long emu_shfl_idx64(long long v, int src) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.x[lane] = v;
    wait(w.bar);
    long long r = w.x[src & 63];
    wait(w.bar);
    return r;
}

// D = A B + C with the fragment mapping the kernels assume (verified on the GPU by
// sdpv_mfma_selftest): A row = lane%32 over bytes of lanes (row, row+32); B column = lane%32.
emu_v16i emu_mfma_i32_32x32x32_i8(emu_v4i a, emu_v4i b, emu_v16i c, int, int, int) {
    Wave& w = my_wave();
    int lane = threadIdx.x & 63;
    w.a[lane] = a;
    w.b[lane] = b;
    wait(w.bar);
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
    wait(w.bar);
    return d;
}

void emu_run_grid(dim3 grid, dim3 block, size_t lds_bytes, const std::function<void()>& body) {
    if (lds_bytes > LDS_MAX) {
        fprintf(stderr, "emu: dynamic LDS %zu exceeds %zu\n", lds_bytes, LDS_MAX);
        abort();
    }
    const unsigned nt = block.x * block.y * block.z;
    if (nt > g_nstacks) {
        if (g_stacks) munmap(g_stacks, g_nstacks * STACK);
        g_stacks = (char*)mmap(nullptr, nt * STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                               -1, 0);
        if (g_stacks == MAP_FAILED) {
            fprintf(stderr, "emu: no memory for %u fiber stacks\n", nt);
            abort();
        }
        g_nstacks = nt;
    }
    // exactly the requested size on the heap, so AddressSanitizer reports any access past it
    std::unique_ptr<unsigned char[]> lds(new unsigned char[lds_bytes ? lds_bytes : 16]);
    emu_dynamic_lds = lds.get();
    g_body = &body;
    blockDim = block;
    gridDim = grid;
    const unsigned nw = (nt + 63) / 64;
    for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
            for (unsigned bx = 0; bx < grid.x; bx++) {
                memset(lds.get(), 0xA5, lds_bytes);   // LDS starts as garbage
                blockIdx = dim3(bx, by, bz);
                g_block = Barrier();
                g_block.expected = (int)nt;
                g_waves = std::vector<Wave>(nw);
                for (unsigned w = 0; w < nw; w++) g_waves[w].bar.expected = (int)std::min(64u, nt - 64 * w);
                g_fibers = std::vector<Fiber>(nt);
                for (unsigned t = 0; t < nt; t++) g_fibers[t].sp = new_stack(t);
                unsigned alive = nt;
                while (alive) {
                    bool ran = false;
                    for (unsigned t = 0; t < nt; t++) {
                        Fiber& f = g_fibers[t];
                        if (f.done || (f.wait && f.wait->gen == f.wait_gen)) continue;
                        g_cur = t;
                        threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                        emu_ctx_switch(&g_sched_sp, f.sp);
                        ran = true;
                        if (f.done) alive--;
                    }
                    if (!ran) {
                        fprintf(stderr, "emu: deadlock in block (%u, %u, %u): %u threads wait at barriers no "
                                        "other thread reaches\n", bx, by, bz, alive);
                        abort();
                    }
                }
            }
}
