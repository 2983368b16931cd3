// FP64 pieces of the step length on gfx950: MPF <-> double conversions and a batched symmetric
// eigensolver (parallel cyclic Jacobi, one workgroup per matrix). The reference copies the blocks
// to the host and calls LAPACK (sdpvibe/solver.py:996-1093); here the float64 copy never leaves
// HBM, only the per-block minimum and the error-bound norm do.
#include <mutex>

#include "common.hpp"

namespace {

constexpr int ET = 256;  // threads per eigen workgroup

// value of entry i as MPF.to_float (sdpvibe/mpf.py:210-215) after multiplying by 2^-shift:
// top two limbs as a double, scaled by 2^(e - shift + 32(L-2)) with the exponent clamped to
// [-1100, 1100]; zero (e = EZ) stays zero.
__global__ void __launch_bounds__(256) k_mpf_to_double(const i32* __restrict__ m, const i64* __restrict__ e, int L,
                                                      i64 n, const i64* __restrict__ shift, i64 per,
                                                      double* __restrict__ out) {
    const i64 i = blockIdx.x * (i64)256 + threadIdx.x;
    if (i >= n) return;
    const i64 ex = e[i];
    if (ex == EZV) {
        out[i] = 0.0;
        return;
    }
    const double hi = (double)m[(i64)(L - 1) * n + i];
    const double top = L > 1 ? hi * 4294967296.0 + (double)(u32)m[(i64)(L - 2) * n + i] : hi;
    i64 sc = ex - (shift ? shift[i / per] : 0) + 32 * (i64)(L - 2);
    if (sc < -1100) sc = -1100;
    if (sc > 1100) sc = 1100;
    out[i] = top * ldexp(1.0, (int)sc);
}

// exact MPF of a double (53-bit mantissa into NB = 32L - 2 bits): m = mant * 2^(NB - 53),
// e = exponent - (NB - 53); zero -> (0, EZ). Non-finite inputs give zero.
__global__ void __launch_bounds__(256) k_mpf_from_double(const double* __restrict__ v, i64 n, int L,
                                                        i32* __restrict__ m, i64* __restrict__ e) {
    const i64 i = blockIdx.x * (i64)256 + threadIdx.x;
    if (i >= n) return;
    const double x = v[i];
    int ex = 0;
    const double fr = frexp(x, &ex);                    // x = fr 2^ex, 0.5 <= |fr| < 1
    i64 mi = (x == 0.0 || !isfinite(x)) ? 0 : (i64)ldexp(fr, 53);
    const int NB = 32 * L - 2;
    if (mi == 0) {
        for (int k = 0; k < L; k++) m[(i64)k * n + i] = 0;
        e[i] = EZV;
        return;
    }
    // word k of the two's complement of mi * 2^sh: bits [32k, 32k + 32) = bits from lo = 32k - sh of mi
    const int sh = NB - 53;
    const u32 ext = mi < 0 ? 0xffffffffu : 0u;
    for (int k = 0; k < L; k++) {
        const int lo = 32 * k - sh;
        u32 w;
        if (lo >= 64) w = ext;
        else if (lo <= -32) w = 0u;
        else if (lo >= 0) w = (u32)(mi >> lo);
        else w = (u32)((u64)mi << (-lo));
        m[(i64)k * n + i] = (i32)w;
    }
    e[i] = (i64)ex - 53 - sh;
}

// Batched symmetric eigen-decomposition, one workgroup per n x n matrix: W = (A + A^T)/2 copied to
// the work area, then Brent-Luk parallel cyclic Jacobi (n/2 disjoint rotations per step, n-1
// steps per sweep) until the off-diagonal mass is below 2^-110 of ||W||_F^2 or 40 sweeps.
// The final off-diagonal Frobenius norm is returned per matrix (off): by Weyl's inequality every
// sorted diagonal entry is within ||offdiag(W)||_2 <= off of the matching exact eigenvalue of W,
// so callers add it to their error bound (a matrix that did not converge gets a wide interval
// instead of a silently wrong one). The rotation pairs live in dynamic LDS (24 bytes per pair).
// Rotation of pair (p, q), Golub & Van Loan sym.schur2: theta = (w_qq - w_pp) / (2 w_pq),
// t = sign(theta) / (|theta| + sqrt(1 + theta^2)), c = 1/sqrt(1 + t^2), s = t c; W <- J^T W J,
// V <- V J. Outputs: eigenvalues ascending (w), eigenvectors as columns of V (same order) when
// V is given, and ||W||_F (fro).
__global__ void __launch_bounds__(ET) k_symeig(const double* __restrict__ A, i64 nb, int n, double* __restrict__ Wk,
                                              double* __restrict__ w, double* __restrict__ V,
                                              double* __restrict__ fro, double* __restrict__ offn) {
    __shared__ double red[ET];
    __shared__ int flag;
    SDPV_DYNAMIC_LDS(double, pairs);
    const int t = threadIdx.x;
    const i64 b = blockIdx.x;
    const int m = n + (n & 1);                          // even player count (dummy index n if odd)
    const int h = m / 2;
    double* rc = pairs;                                 // rotation cosines (h)
    double* rs = pairs + h;                             // rotation sines (h)
    int* pp = reinterpret_cast<int*>(pairs + 2 * h);    // pair rows p < q (h each)
    int* qq = pp + h;
    const double* a = A + b * (i64)n * n;
    double* W = Wk + b * (i64)m * m;
    double* Vb = V ? V + b * (i64)n * n : nullptr;
    double* Vw = Wk + nb * (i64)m * m + b * (i64)m * m;  // V work (m x m)

    // symmetrised copy + Frobenius norm
    double s2 = 0.0;
    for (int idx = t; idx < m * m; idx += ET) {
        const int i = idx / m, j = idx % m;
        double x = 0.0;
        if (i < n && j < n) x = 0.5 * (a[(i64)i * n + j] + a[(i64)j * n + i]);
        W[idx] = x;
        Vw[idx] = (i == j) ? 1.0 : 0.0;
        s2 += x * x;
    }
    red[t] = s2;
    __syncthreads();
    for (int st = ET / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    const double fro2 = red[0];
    __syncthreads();
    if (t == 0 && fro) fro[b] = sqrt(fro2);

    for (int sweep = 0; sweep < 40 && m > 1; sweep++) {
        // off-diagonal mass
        double off = 0.0;
        for (int idx = t; idx < m * m; idx += ET) {
            const int i = idx / m, j = idx % m;
            if (i != j) off += W[idx] * W[idx];
        }
        red[t] = off;
        __syncthreads();
        for (int st = ET / 2; st > 0; st >>= 1) {
            if (t < st) red[t] += red[t + st];
            __syncthreads();
        }
        if (t == 0) flag = !(red[0] > fro2 * 0x1p-110);
        __syncthreads();
        if (flag) break;
        for (int step = 0; step < m - 1; step++) {
            // round-robin pairing: position 0 fixed, the others rotate by `step`
            for (int k = t; k < h; k += ET) {
                auto player = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + step) % (m - 1); };
                int p = player(k), q = player(m - 1 - k);
                if (p > q) { const int x = p; p = q; q = x; }
                pp[k] = p;
                qq[k] = q;
                const double apq = W[(i64)p * m + q];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    const double app = W[(i64)p * m + p], aqq = W[(i64)q * m + q];
                    const double theta = (aqq - app) / (2.0 * apq);
                    const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                }
                rc[k] = c;
                rs[k] = s;
            }
            __syncthreads();
            // rows p, q of every pair: W <- J^T W
            for (int idx = t; idx < h * m; idx += ET) {
                const int k = idx / m, j = idx % m;
                const double c = rc[k], s = rs[k];
                if (s == 0.0) continue;
                double* rp = W + (i64)pp[k] * m;
                double* rq = W + (i64)qq[k] * m;
                const double x = rp[j], y = rq[j];
                rp[j] = c * x - s * y;
                rq[j] = s * x + c * y;
            }
            __syncthreads();
            // columns p, q: W <- W J, V <- V J
            for (int idx = t; idx < h * m; idx += ET) {
                const int k = idx / m, i = idx % m;
                const double c = rc[k], s = rs[k];
                if (s == 0.0) continue;
                const int p = pp[k], q = qq[k];
                double* ri = W + (i64)i * m;
                const double x = ri[p], y = ri[q];
                ri[p] = c * x - s * y;
                ri[q] = s * x + c * y;
                double* vi = Vw + (i64)i * m;
                const double vx = vi[p], vy = vi[q];
                vi[p] = c * vx - s * vy;
                vi[q] = s * vx + c * vy;
            }
            __syncthreads();
            for (int k = t; k < h; k += ET) {
                if (rs[k] == 0.0) continue;
                W[(i64)pp[k] * m + qq[k]] = 0.0;
                W[(i64)qq[k] * m + pp[k]] = 0.0;
            }
            __syncthreads();
        }
    }
    // remaining off-diagonal mass (bounds the distance of the diagonal to the spectrum)
    {
        double off = 0.0;
        for (int idx = t; idx < m * m; idx += ET) {
            const int i = idx / m, j = idx % m;
            if (i != j) off += W[idx] * W[idx];
        }
        red[t] = off;
        __syncthreads();
        for (int st = ET / 2; st > 0; st >>= 1) {
            if (t < st) red[t] += red[t + st];
            __syncthreads();
        }
        if (t == 0 && offn) offn[b] = sqrt(red[0]);
        __syncthreads();
    }
    // ascending order of the diagonal (ties by index), eigenvectors permuted alike
    for (int i = t; i < n; i += ET) {
        const double di = W[(i64)i * m + i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
            const double dj = W[(i64)j * m + j];
            rank += (dj < di) || (dj == di && j < i);
        }
        w[b * (i64)n + rank] = di;
        if (Vb)
            for (int r = 0; r < n; r++) Vb[(i64)r * n + rank] = Vw[(i64)r * m + i];
    }
}

// Device-resident workspace of sdpv_symeig_static: the Jacobi work area (2 nb m^2 doubles) of callers
// that cannot allocate. One copy per GPU, held until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int64_t SYMEIG_SCRATCH_DOUBLES = SDPV_SYMEIG_STATIC_WORK;
__device__ __align__(256) double scratch[SYMEIG_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double* scratch_of_device[MAX_DEVICES];
std::once_flag scratch_once[MAX_DEVICES];
hipError_t scratch_err[MAX_DEVICES];

// address of the current device's scratch array (looked up once per device, with it current)
double* device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::call_once(scratch_once[dev], [dev] {
        void* p = nullptr;
        scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        scratch_of_device[dev] = scratch_err[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
    });
    return scratch_of_device[dev];
}

}  // namespace

extern "C" {

int sdpv_mpf_to_double(const int32_t* m, const int64_t* e, int64_t L, int64_t n, const int64_t* shift, int64_t per,
                       double* out, void* stream) {
    if (L < 1 || n < 0 || (shift && per < 1)) return SDPV_EARG;
    if (n == 0) return SDPV_OK;
    if (!m || !e || !out) return SDPV_EARG;
    hipLaunchKernelGGL(k_mpf_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), m, e,
                       (int)L, n, shift, per, out);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mpf_from_double(const double* v, int64_t n, int64_t L, int32_t* m, int64_t* e, void* stream) {
    if (L < 2 || n < 0) return SDPV_EARG;
    if (n == 0) return SDPV_OK;
    if (!v || !m || !e) return SDPV_EARG;
    hipLaunchKernelGGL(k_mpf_from_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), v, n,
                       (int)L, m, e);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_symeig(const double* A, int64_t nb, int64_t n, double* work, double* w, double* V, double* fro,
                double* off, void* stream) {
    if (nb < 0 || n < 1) return SDPV_EARG;
    if (n > SDPV_SYMEIG_NMAX || nb > 0x7fffffffLL) return SDPV_ESIZE;
    if (nb == 0) return SDPV_OK;
    if (!A || !work || !w) return SDPV_EARG;
    const i64 h = (n + (n & 1)) / 2;
    const size_t lds = (size_t)h * (2 * sizeof(double) + 2 * sizeof(int));
    hipLaunchKernelGGL(k_symeig, dim3((unsigned)nb), dim3(ET), lds, as_stream(stream), A, nb, (int)n, work, w, V, fro,
                       off);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_symeig_static(const double* A, int64_t nb, int64_t n, double* w, double* V, double* fro, double* off,
                       void* stream) {
    if (nb < 0 || n < 1) return SDPV_EARG;
    if (n > SDPV_SYMEIG_NMAX || nb > 0x7fffffffLL) return SDPV_ESIZE;
    const int64_t m = n + (n & 1);
    if (nb > 0 && 2 * m * m > SYMEIG_SCRATCH_DOUBLES / nb) return SDPV_ESIZE;   // 2 nb m^2 doubles
    if (nb == 0) return SDPV_OK;
    if (!A || !w) return SDPV_EARG;
    double* work = device_scratch();
    if (!work) return SDPV_ELAUNCH;
    const size_t lds = (size_t)(m / 2) * (2 * sizeof(double) + 2 * sizeof(int));
    hipLaunchKernelGGL(k_symeig, dim3((unsigned)nb), dim3(ET), lds, as_stream(stream), A, nb, (int)n, work, w, V, fro,
                       off);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

}  // extern "C"
