"""Benchmark on one MI355X. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

step       : one Mehrotra predictor-corrector iteration of the interior-point solver (sdpvibe_amd.ipm)
             on the synthetic polynomial-matrix SDP of the IPM configuration: 300 dimension-1 and 60
             dimension-3 positivity constraints of degree 26 (27 rescaled Laguerre points, orthonormal
             bilinear bases), N = 3000 free variables, 768 bits (L = 24), strictly feasible by
             construction (sdp.SDP.polynomial, seeded host PCG64 data). W iterations run untimed, then
             K iterations are timed, each between two device synchronisations.
value      : seconds per IPM iteration of the batched solver (lower is better)
plain      : the same SDP in the per-block layout (one launch sequence per block), timed over the
             same iterations; its records must be bitwise those of the batched run
mp_gemm    : the exact multiprecision GEMM of the plain-kernel configuration (512 x 512 x 512 at
             L = 24) as % of the dense INT8 MFMA peak, counted with the reference algorithm's
             operation count (2 [4 n M N K + 2 n 4L (M + N) K + 2 E n M N], n = 112, E = 216)
roofline   : the product kernel of that GEMM, timed with HIP events on its launch stream
cpu_baseline: the reference's own CPU path (cpu_baseline/mpf_baseline.c, GMP mpf naive loops,
             compiled from source into oracle/_ref) at 512 / 768 bits on this process's CPU share
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT8_TOPS = 5000.0       # MI355X dense INT8 MFMA, TOP/s
PEAK_HBM_GBS = 8000.0         # MI355X HBM3E, GB/s
METRIC = "s/IPM iteration @768-bit N=3000 SDP (batched vs plain); mp-GEMM % INT8 MFMA peak"


def _host_threads():
    """host threads this process may use: its CPU affinity set (the machine's CPU share for this
    GPU), capped by OMP_NUM_THREADS when that is set"""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() and int(omp) > 0 else n


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed IPM iterations")
    ap.add_argument("--warmup", type=int, default=2, help="untimed IPM iterations before them")
    ap.add_argument("--gemm-steps", type=int, default=50)
    ap.add_argument("--plain-iters", type=int, default=2, help="iterations timed in the per-block layout")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--stream-iters", type=int, default=2,
                    help="iterations timed with every B and P streamed from host memory (0: skip)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=_host_threads())
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-block Cholesky comparison")
    ap.add_argument("--blocks", type=int, default=400)
    ap.add_argument("--no-gram", action="store_true", help="skip the Gram accumulation measurement")
    ap.add_argument("--no-ipm", action="store_true",
                    help="skip the IPM iterations (for counter runs of the GEMM leg; value is then null)")
    ap.add_argument("--gemm-only", action="store_true",
                    help="only the 512^3 GEMM leg and its roofline (for the kernel trace and counter runs of the "
                         "roofline kernel); value is then null")
    ap.add_argument("--chol-only", type=int, default=0, metavar="N",
                    help="only time mpf.Chol of an N x N SPD matrix (G G^T + N I, the Schur complement's size "
                         "in the IPM at N = 3000) twice and print its line (for profiling the tile kernels)")
    return ap.parse_args()


def chol_only(N, L):
    """two mpf.Chol calls on Q = G G^T + N I (G an N x N Gaussian MPF): the first includes the product
    autotuner's runs at new shapes, the second is the steady state"""
    import torch
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd.synthetic import gaussian_mpf
    gm, ge = gaussian_mpf((N, N), L, seed=3)
    G = F.MPF(torch.from_numpy(gm).cuda(), torch.from_numpy(ge).cuda())
    Q = F.add(F.gemm(G, G, transB=True), F.eye(N, L, val=N))
    del G
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        F.Chol(Q)
        torch.cuda.synchronize()
        times.append(round(time.perf_counter() - t, 4))
    return {"workload": f"mpf.Chol of G G^T + {N} I ({N} x {N}, L={L})", "seconds": times}


def gram_acc(L, N=3000, rows=1024, chunks=4):
    """Schur-complement assembly Q = sum_c P_c^T P_c (mpf.GramAcc: residues of ctx(L + 1), syrk
    products on the lower tiles, one CRT at the end) for N columns and `chunks` chunks of `rows`
    Gaussian MPF rows (one chunk drawn, added `chunks` times). Reports mp-MAC/s over the lower
    triangle and the time of the final reconstruction."""
    import numpy as np
    import torch
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd.synthetic import gaussian_mpf

    pm, pe = gaussian_mpf((rows, N), L, seed=7)
    P = F.MPF(torch.from_numpy(pm).cuda(), torch.from_numpy(pe).cuda())
    del pm, pe

    def run():
        G = F.GramAcc(N, L)
        for _ in range(chunks):
            G.add(P)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Q = G.result()
        torch.cuda.synchronize()
        return Q, time.perf_counter() - t1

    run()                                                 # warm (autotuner, attributes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Q, t_crt = run()
    dt = time.perf_counter() - t0
    mac = chunks * rows * N * (N + 1) // 2
    R = F.ctx(L + 1)
    ops = 2 * (4 * R.n * mac + 2 * R.n * 4 * (L + 1) * chunks * rows * N + 2 * R.E * R.n * N * (N + 1) // 2)
    del Q
    return {"workload": f"GramAcc Q = sum P^T P, N={N}, {chunks} chunks of {rows} rows, L={L} (residues of L+1 "
                        f"limbs, {R.n} primes)", "seconds": round(dt, 4), "crt_seconds": round(t_crt, 4),
            "mp_mac_per_s": mac / dt, "pct_int8_peak": round(100.0 * ops / dt / (PEAK_INT8_TOPS * 1e12), 3)}


def batched_chol(L, nblocks, seed=0, rhs=4):
    """Uneven-block Cholesky + solve at L limbs (synthetic.uneven_spd_blocks: n uniform in 8..96,
    A = D (B B^T + n I) D, D = diag(10^u), u uniform in [-20, 20], `rhs` right-hand sides per block).
    Times the plain per-block form (one Chol and one solve per block) against the bucketed form
    (one batched factorisation and solve per size class); results are bitwise equal."""
    import torch
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd import synthetic as SY

    dims, blocks, rhs_l = SY.uneven_spd_blocks(nblocks, L, seed=seed, rhs=rhs)
    torch.cuda.synchronize()

    def plain():
        return [F.Chol(A).solve(b) for A, b in zip(blocks, rhs_l)]

    def bucketed():
        return F.BucketedChol(blocks).solve(rhs_l)

    out = {}
    for name, f in (("plain", plain), ("bucketed", bucketed)):
        f()                                               # warm (autotuner, attributes)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = f()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0, res)
    same = all(torch.equal(a.m, b.m) and torch.equal(a.e, b.e) for a, b in zip(out["plain"][1], out["bucketed"][1]))
    classes = len({max(32, -(-n // 32) * 32) for n in dims})
    return {"workload": f"{nblocks} SPD blocks, n uniform in 8..96, A = D (B B^T + n I) D, D = diag(10^u), "
                        f"u uniform in [-20, 20]; factor + solve with {rhs} RHS each, L={L}",
            "plain_ms": round(out["plain"][0] * 1e3, 2), "bucketed_ms": round(out["bucketed"][0] * 1e3, 2),
            "speedup": round(out["plain"][0] / out["bucketed"][0], 2), "size_classes": classes,
            "bitwise_equal": bool(same)}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "mpgemm_512_L24_hbm_traffic.md")


def gemm_source_hash():
    """64-bit hash of the exact GEMM's kernel code (rns.hip and common.hpp; the C ABI header holds
    declarations only); the library loader guarantees that the loaded library was built from these
    files"""
    import hashlib
    h = hashlib.sha256()
    for f in ("sdpvibe_amd/csrc/common.hpp", "sdpvibe_amd/csrc/rns.hip"):
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return "%#x" % int(h.hexdigest()[:16], 16)


def measured_traffic(kernel, config, n, L):
    """HBM bytes per launch of `kernel` (product tile configuration `config`) at this workload, from
    the rocprofv3 FETCH_SIZE/WRITE_SIZE summary, if it was measured on these kernel sources
    (gemm_source_hash); None otherwise.
    Lines: `traffic <kernel> config=<c> n=<n> L=<L> bytes=<B> source=<hash>`."""
    from sdpvibe_amd import _lib
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    _lib.lib()
    src = gemm_source_hash()
    pat = re.compile(r"^traffic (\S+) config=(-?\d+) n=(\d+) L=(\d+) bytes=(\d+) source=(0x[0-9a-f]+)\s*$")
    for line in open(TRAFFIC_PROFILE):
        m = pat.match(line.strip())
        if m and m.group(1) == kernel and int(m.group(2)) == config and int(m.group(3)) == n \
                and int(m.group(4)) == L and m.group(6) == src:
            return int(m.group(5))
    return None


def cpu_baseline(n, threads, L):
    """The reference's CPU path on the same shape: mpf_baseline <prec> <n> <nrhs> <threads> (GMP mpf at
    32L bits): Crout Cholesky of an n x n SPD matrix, TRSM with n right-hand sides and an n x n x n
    GEMM, each timed by the program itself. `value` is its GEMM rate; the Cholesky and TRSM rates
    of the same run are reported beside it."""
    exe = os.path.join(ROOT, "oracle", "_ref", "mpf_baseline")
    if not os.path.exists(exe):
        return None
    prec = 32 * L
    t0 = time.time()
    r = subprocess.run([exe, str(prec), str(n), str(n), str(threads)], capture_output=True, text=True,
                       timeout=900, env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    wall = time.time() - t0
    rates = {k: re.search(k + r"\s+([\d.]+) s\s+([\d.]+) M mp-MAC/s", r.stdout) for k in ("cholesky", "trsm", "gemm")}
    if r.returncode != 0 or not all(rates.values()):
        return None
    sec = {k: float(m.group(1)) for k, m in rates.items()}
    rate = {k: float(m.group(2)) * 1e6 for k, m in rates.items()}
    # one IPM iteration of the N = 3000 polynomial SDP at these rates, counting only its three dominant
    # contractions (a lower bound): Q += P^T P over sum P_j = 17820 rows (lower triangle), P = L_S^-1 B
    # per block (P_j^2 N / 2; 300 blocks of P = 27, 60 of P = 162) and chol(Q) (N^3 / 6)
    N, rows = 3000, 300 * 27 + 60 * 162
    work = {"gemm": rows * N * (N + 1) / 2, "trsm": (300 * 27 ** 2 + 60 * 162 ** 2) * N / 2, "cholesky": N ** 3 / 6}
    ipm_est = sum(work[k] / rate[k] for k in work)
    return {"value": rate["gemm"], "unit": "mp-MAC/s (768-bit GEMM)", "cores": threads, "kind": "reference",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "cholesky_mp_mac_per_s": rate["cholesky"], "trsm_mp_mac_per_s": rate["trsm"],
            "seconds": sec,
            "ipm_s_per_iteration_lower_bound": round(ipm_est, 1),
            "ipm_lower_bound_mp_mac": {k: float(f"{v:.4g}") for k, v in work.items()},
            "sample": f"cpu_baseline/mpf_baseline.c {prec} {n} {n} {threads} (GMP mpf naive loops, OpenMP "
                      f"{threads} threads = this process's CPU share): Cholesky {n}x{n} {sec['cholesky']:.2f} s, "
                      f"TRSM {n} RHS {sec['trsm']:.2f} s, GEMM {n}x{n}x{n} {sec['gemm']:.2f} s; "
                      f"{wall:.1f} s wall in total"}


def plain_chol(n, L, reps=3):
    """GPU Cholesky of the plain-kernel configuration's SPD matrix A = B B^T + n I and its two solves
    with n right-hand sides (L^-1 R, then L^-T of that), averaged over `reps` runs (ms), with the
    rates in mp-MAC/s counted as the CPU baseline counts them (n^3/6 for the Cholesky, n^3/2 per
    triangular solve of n right-hand sides)."""
    import torch
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd import synthetic as SY
    _, A = SY.plain_config(n, L, seed=0)
    Rm, Re = SY.gaussian_mpf((n, n), L, seed=1)
    R = F.MPF(torch.from_numpy(Rm).cuda(), torch.from_numpy(Re).cuda())
    F.Chol(A).solve(R)                                    # warm (autotuner)
    torch.cuda.synchronize()
    tc = ts = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        C = F.Chol(A)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        C.solve_upper(C.solve_lower(R))
        torch.cuda.synchronize()
        tc += t1 - t0
        ts += time.perf_counter() - t1
    tc, ts = tc / reps, ts / reps
    return {"workload": f"Chol of A = B B^T + {n} I ({n}x{n}, L={L}) and L^-1, L^-T solves with {n} RHS",
            "chol_ms": round(tc * 1e3, 2), "solves_ms": round(ts * 1e3, 2),
            "chol_mp_mac_per_s": n ** 3 / 6 / tc, "trsm_mp_mac_per_s": n ** 3 / ts}


def ipm_leg(args, L=24):
    """seconds per IPM iteration, batched and per-block, on the polynomial SDP (see the docstring)"""
    import torch
    from sdpvibe_amd.ipm import Solver
    from sdpvibe_amd.sdp import SDP

    t0 = time.perf_counter()
    sdp = SDP.polynomial(n1=300, n3=60, degree=26, N=3000, L=L, seed=0)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0

    def timed(s, iters):
        solver = Solver(s, L)
        gen = solver.iterate(max_iter=iters)
        times = []
        for _ in range(iters):
            torch.cuda.synchronize()
            t = time.perf_counter()
            next(gen)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        return solver, times

    W, K = args.warmup, args.steps
    solver, times = timed(sdp, W + K)
    out = {"s_per_iteration": sum(times[W:]) / K, "iteration_s": [round(t, 4) for t in times],
           "setup_s": round(setup, 2), "groups": len(sdp.groups), "total_psd_rows": sdp.total_psd_rows,
           "last_record": {k: solver.log[-1][k][:12] for k in ("mu", "P-obj", "D-obj", "P-step", "D-step")}}
    if not args.no_plain and args.plain_iters > 0:
        # iterations 1 (autotuner) and 2 (graph capture) untimed in both forms
        P = args.plain_iters
        plain = sdp.relayout("plain")
        psolver, ptimes = timed(plain, 2 + P)
        same = all(ra[k] == rb[k] for ra, rb in zip(solver.log, psolver.log)
                   for k in ("mu", "P-obj", "D-obj", "P-err", "p-err", "D-err", "R-err", "P-step", "D-step", "beta"))
        bt = sum(times[2:2 + P]) / P
        pt = sum(ptimes[2:]) / P
        out["plain"] = {"groups": len(plain.groups), "s_per_iteration": round(pt, 4),
                        "batched_s_per_iteration_same_iterations": round(bt, 4), "speedup": round(pt / bt, 2),
                        "iterations": f"3..{2 + P}", "records_bitwise_equal": bool(same)}
        del plain, psolver
    if args.stream_iters > 0:
        # the same iterations with every B and P = L_S^-1 B in page-locked host memory, copied in per
        # chunk (host streaming for SDPs larger than HBM); iterations 1 and 2 untimed
        from sdpvibe_amd.layout import Layout
        S = args.stream_iters
        streamed = sdp.placed(Layout(sdp.shapes, sdp.N, sdp.L, stream="all"))
        st = streamed.store
        h0, d0 = st.bytes_h2d, st.bytes_d2h                   # the placement's own copies
        ssolver, stimes = timed(streamed, 2 + S)
        same = all(ra[k] == rb[k] for ra, rb in zip(solver.log, ssolver.log)
                   for k in ("mu", "P-obj", "D-obj", "P-err", "p-err", "D-err", "R-err", "P-step", "D-step", "beta"))
        out["streamed"] = {"plan": streamed.layout_plan.describe(), "s_per_iteration": round(sum(stimes[2:]) / S, 4),
                           "resident_s_per_iteration_same_iterations": round(sum(times[2:2 + S]) / S, 4),
                           "host_to_device_GB_per_iteration": round((st.bytes_h2d - h0) / (2 + S) / 2 ** 30, 1),
                           "device_to_host_GB_per_iteration": round((st.bytes_d2h - d0) / (2 + S) / 2 ** 30, 1),
                           "iterations": f"3..{2 + S}", "records_bitwise_equal": bool(same)}
        del ssolver
        st.close()
    return out


def gemm_leg(args):
    """the 512^3 exact GEMM at L = 24: % of INT8 peak, HIP-graph replay, per-phase times, roofline"""
    import torch
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd import mpmat
    from sdpvibe_amd.synthetic import gaussian_mpf

    n, L = args.n, args.L
    M = N = K = n
    Am, Ae = gaussian_mpf((1, n, n), L, seed=0)
    Bm, Be = gaussian_mpf((1, n, n), L, seed=1)
    A = F.MPF(torch.from_numpy(Am).cuda(), torch.from_numpy(Ae).cuda())
    B = F.MPF(torch.from_numpy(Bm).cuda(), torch.from_numpy(Be).cuda())
    R = mpmat.ctx(L)
    steps = args.gemm_steps
    for _ in range(5):
        F.gemm(A, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        C = F.gemm(A, B)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del C

    # ---- the same step replayed from a HIP graph (one capture, no host work per step)
    from sdpvibe_amd.graph import Captured
    cap = Captured(F.gemm, A, B)
    for _ in range(5):
        cap.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cap.replay()
    torch.cuda.synchronize()
    dt_graph = (time.perf_counter() - t0) / steps
    del cap

    # ---- per-phase times with HIP events on the launch stream (torch's current stream): every
    # buffer allocated first, then each phase launched `reps` times back to back between two
    # events, so the average is the phase's own kernel time (the rocprofv3 trace agrees)
    st = torch.cuda.current_stream()
    nb = 1
    kp = R.kp(K)
    phases = ["row_exp", "split_A", "split_B", "product", "crt"]
    RhA = torch.empty((R.n, kp // 16, M, 16), dtype=torch.int8, device="cuda")
    RlA = torch.empty_like(RhA)
    RhB = torch.empty((R.n, kp // 16, N, 16), dtype=torch.int8, device="cuda")
    RlB = torch.empty_like(RhB)
    acc = torch.empty((R.n, M, N), dtype=torch.int16, device="cuda")
    om = torch.empty((L, 1, M, N), dtype=torch.int32, device="cuda")
    oe = torch.empty((1, M, N), dtype=torch.int64, device="cuda")
    EA = F._row_exp(A.m, A.e, L, A.e.numel(), 1, M, K, M * K, K, 1)
    EB = F._row_exp(B.m, B.e, L, B.e.numel(), 1, N, K, K * N, 1, N)
    run = {
        "row_exp": lambda: (F._row_exp(A.m, A.e, L, A.e.numel(), 1, M, K, M * K, K, 1),
                            F._row_exp(B.m, B.e, L, B.e.numel(), 1, N, K, K * N, 1, N)),
        "split_A": lambda: R.split(A.m, A.e, EA, L, A.e.numel(), 1, M, M * K, K, 1, 0, K, kp, RhA, RlA),
        "split_B": lambda: R.split(B.m, B.e, EB, L, B.e.numel(), 1, N, K * N, 1, N, 0, K, kp, RhB, RlB),
        "product": lambda: R.product(RhA, RlA, RhB, RlB, acc, 1, M, N, M * N, scaled=True),
        "crt": lambda: R.crt(acc, M, N, 0, N, 1, outm=om, oute=oe, L=L, plane=M * N, ldo=N, jo=0, Mb=M,
                             erow=EA, ecol=EB, scaled=True),
    }
    reps = max(10, steps)
    acc_ms = {}
    for p in phases:
        run[p]()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            run[p]()
        e1.record(st)
        torch.cuda.synchronize()
        acc_ms[p] = e0.elapsed_time(e1)
    ph = {p: acc_ms[p] / reps for p in phases}

    # ---- algorithmic work (this algorithm: 3 INT8 products per prime; padding excluded)
    nP, E = R.n, R.E
    ops_split = 2 * (2 * nP * 4 * L * (M + N) * K)       # T_hi and T_lo byte dots per entry and prime
    ops_prod = 2 * (3 * nP * M * N * K)
    ops_crt = 2 * (2 * E * nP * M * N)                     # hi and lo digit products
    ops_total = ops_split + ops_prod + ops_crt
    ref_ops = 2 * (4 * 112 * M * N * K + 2 * 112 * 4 * L * (M + N) * K + 2 * 216 * 112 * M * N)
    value = 100.0 * ref_ops / dt / (PEAK_INT8_TOPS * 1e12)

    # algorithmic HBM bytes per launch of the bandwidth-bound phases (limbs + exponents in, residue
    # bytes out for a split; residue coefficients in, limbs + exponent out for the reconstruction)
    hbm = {"split_A": (4 * L + 8 + 2 * nP) * M * K, "split_B": (4 * L + 8 + 2 * nP) * N * K,
           "crt": (2 * nP + 4 * L + 8) * M * N}
    dom = max(phases, key=lambda p: ph[p])
    t_dom = ph[dom] * 1e-3
    if dom == "product":
        ach = ops_prod / t_dom / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                "frac": ach / PEAK_INT8_TOPS, "traffic": None}
    else:
        per_entry = {"split_A": (4 * L + 8 + 2 * nP) * M * K, "split_B": (4 * L + 8 + 2 * nP) * N * K,
                     "crt": (2 * nP + 4 * L + 8) * M * N, "row_exp": 16 * (M + N) * K}[dom]
        ach = per_entry / t_dom / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None}
    roof["kernel"] = {"product": "k_rns_gemm", "split_A": "k_rns_split", "split_B": "k_rns_split",
                      "crt": "k_rns_crt", "row_exp": "k_row_exp_*"}[dom]
    roof["avg_launch_ms"] = ph[dom]
    from sdpvibe_amd import rns as RN
    cfg = RN.tuned_config(R, 1, M, N, kp) if dom == "product" else None
    if cfg is not None:
        roof["config"] = cfg
    roof["traffic"] = measured_traffic(roof["kernel"], cfg if cfg is not None else -1, n, L)

    return {"workload": f"mpf.gemm {M}x{K}x{N} at {32 * L}-bit (L={L}), plain-kernel configuration",
            "pct_int8_peak": round(value, 3), "ms_per_gemm": round(dt * 1e3, 4),
            "ms_per_gemm_graph": round(dt_graph * 1e3, 4),
            "pct_peak_graph": round(100.0 * ref_ops / dt_graph / (PEAK_INT8_TOPS * 1e12), 3),
            "mp_mac_per_s": M * N * K / dt, "effective_int8_ops": ref_ops, "executed_int8_ops": ops_total,
            "pct_peak_executed_ops": round(100.0 * ops_total / dt / (PEAK_INT8_TOPS * 1e12), 3),
            "primes": nP, "crt_digits": E, "phase_ms": {p: round(v, 4) for p, v in ph.items()},
            "phase_hbm_frac": {p: round(hbm[p] / (ph[p] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) for p in hbm}}, roof


def _guard(f, *a):
    """an auxiliary measurement never loses the line"""
    try:
        return f(*a)
    except Exception as exc:                              # noqa: BLE001
        return {"error": f"{type(exc).__name__}: {exc}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on one GPU; "
              f"the path would shard by prime and by block, see the design notes)", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    L = args.L
    if args.chol_only:
        print(json.dumps({"chol_only": chol_only(args.chol_only, L)}))
        return
    if args.gemm_only:
        args.no_ipm = args.no_cpu_baseline = args.no_batched = args.no_gram = True
    ipm = None if args.no_ipm else ipm_leg(args, L)
    gemm, roof = gemm_leg(args)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.n, args.cpu_threads, L)
    line = {
        "metric": METRIC, "value": round(ipm["s_per_iteration"], 4) if ipm else None, "unit": "s/IPM iteration",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ipm["s_per_iteration"] * 1e3 if ipm else None,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int8 (exact 768-bit)",
        "data": "synthetic (strictly feasible polynomial-matrix SDP from host PCG64 data, seed 0)",
        "config": {"workload": "Mehrotra predictor-corrector IPM iteration, polynomial-matrix SDP: 300 dim-1 + 60 "
                               "dim-3 positivity constraints, degree 26, N=3000, 768-bit (L=24), batched layout",
                   "L": L, "N": 3000, "blocks": 360},
        "ipm": ipm, "mp_gemm": gemm, "roofline": roof, "cpu_baseline": cpu,
        "plain_chol": None if args.gemm_only else _guard(plain_chol, args.n, L),
        "batched_chol": None if args.no_batched else _guard(batched_chol, L, args.blocks),
        "gram_acc": None if args.no_gram else _guard(gram_acc, L),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
