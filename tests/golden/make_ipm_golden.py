"""Writes tests/golden/ipm_small.json: a small strictly feasible polynomial-matrix SDP (two dimension-1
and one dimension-3 positivity constraints, degree 4, N = 5, entries rounded to 768-bit MPF, L = 24)
and the per-iteration records of the oracle's interior-point iteration on it (oracle/ipm_ref.py, at
1600 bits), for tests/test_gpu_ipm.py.

The SDP follows the recipe of sdpvibe_amd.sdp.SDP.polynomial, computed here exactly in mpmath on the
host (rescaled Laguerre points, orthonormal bilinear bases, B = Phi A with Laguerre polynomials and
coefficients g 10^u, b = B^T x*, c = Tr(A_p) + B y*), then rounded to NB = 766 bits.

    python tests/golden/make_ipm_golden.py
"""
import json
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ipm_ref as IR  # noqa: E402

L = 24
NB = 32 * L - 2


def rnd(v):
    """round to NB bits, magnitude half up (the MPF rounding); returns (mantissa, exponent)"""
    v = mp.mpf(v)
    if v == 0:
        return 0, None
    sign, man, ex, bc = v._mpf_
    man = int(man)
    if bc > NB:
        sh = bc - NB
        man = (man + (1 << (sh - 1))) >> sh
        ex += sh
    return (-man if sign else man), int(ex)


def val(me):
    m, e = me
    return mp.mpf(0) if m == 0 else mp.ldexp(mp.mpf(m), e)


def polynomial_sdp(dims, degree, N, seed, coef_pow=8):
    mp.mp.prec = 2 * NB + 64
    K = degree + 1
    he, ho = degree // 2 + 1, (degree - 1) // 2 + 1
    rho = 3 - 2 * mp.sqrt(2)
    xs = [mp.pi ** 2 * (4 * k - 1) ** 2 / (-64 * K * mp.log(rho)) for k in range(K)]
    sk = [mp.exp(-x) for x in xs]

    def orthonormal_rows(h, weight):
        out = []
        for m in range(h):
            v = [mp.sqrt(weight(k)) * xs[k] ** m for k in range(K)]
            for _ in range(2):
                for q in out:
                    dp = mp.fsum(a * b for a, b in zip(v, q))
                    v = [a - dp * b for a, b in zip(v, q)]
            nrm = mp.sqrt(mp.fsum(a * a for a in v))
            out.append([a / nrm for a in v])
        return out

    Ve = [[rnd(v) for v in r] for r in orthonormal_rows(he, lambda k: sk[k])]
    Vo = [[rnd(v) for v in r] for r in orthonormal_rows(ho, lambda k: xs[k] * sk[k])]
    phi = [[sk[k] * mp.laguerre(t, 0, xs[k]) for t in range(K)] for k in range(K)]
    rng = np.random.default_rng(seed)
    ystar = [rnd(mp.mpf(float(v))) for v in 0.1 * rng.standard_normal(N)]
    vn = [mp.fsum(val(V[m][k]) ** 2 for V in (Ve, Vo) for m in range(len(V))) for k in range(K)]
    blocks, btot = [], [mp.mpf(0)] * N
    for j, d in enumerate(dims):
        npair = d * (d + 1) // 2
        gr = np.random.default_rng([seed, j])
        g = gr.standard_normal((npair, N, K))
        u = gr.integers(-coef_pow, coef_pow + 1, (npair, N, K))
        B = []
        for s in range(npair):
            for k in range(K):
                B.append([rnd(mp.fsum(phi[k][t] * mp.mpf(float(g[s, n, t])) * mp.mpf(10) ** int(u[s, n, t])
                                      for t in range(K))) for n in range(N)])
        c = []
        for s, (cc, r) in enumerate(IR.segments(d)):
            for k in range(K):
                p = s * K + k
                tr = vn[k] if cc == r else mp.mpf(0)
                c.append(rnd(tr + mp.fsum(val(B[p][n]) * val(ystar[n]) for n in range(N))))
                if cc == r:
                    for n in range(N):
                        btot[n] += val(B[p][n])
        blocks.append(dict(dim=d, K=K, Ve=Ve, Vo=Vo, c=c, B=B))
    b = [rnd(v) for v in btot]
    return blocks, b


def to_json(me):
    return [str(me[0]), me[1]]


def main():
    dims, degree, N = [1, 1, 3], 4, 5
    blocks, b = polynomial_sdp(dims, degree, N, seed=0)
    oblocks = [IR.Block(bl["dim"], bl["K"], [[val(v) for v in r] for r in bl["Ve"]],
                        [[val(v) for v in r] for r in bl["Vo"]], [val(v) for v in bl["c"]],
                        [[val(v) for v in r] for r in bl["B"]]) for bl in blocks]
    log, res = IR.solve(oblocks, [val(v) for v in b], max_iter=20, prec=1600)
    keys = ["mu", "pobj", "dobj", "gap", "primal_err_P", "primal_err_p", "dual_err", "R_err", "alphaP", "alphaD",
            "beta"]
    out = dict(L=L, dims=dims, degree=degree, N=N, seed=0,
               blocks=[dict(dim=bl["dim"], K=bl["K"], Ve=[[to_json(v) for v in r] for r in bl["Ve"]],
                            Vo=[[to_json(v) for v in r] for r in bl["Vo"]], c=[to_json(v) for v in bl["c"]],
                            B=[[to_json(v) for v in r] for r in bl["B"]]) for bl in blocks],
               b=[to_json(v) for v in b],
               rows=[{k: mp.nstr(r[k], 40) for k in keys} for r in log],
               terminate=res["terminateReason"])
    with open(os.path.join(HERE, "ipm_small.json"), "w") as f:
        json.dump(out, f, indent=0)
    for r in log:
        print(r["iteration"], mp.nstr(r["mu"], 8), mp.nstr(r["pobj"], 12), mp.nstr(r["dobj"], 12),
              mp.nstr(r["alphaP"], 5), mp.nstr(r["alphaD"], 5), mp.nstr(r["beta"], 5))


if __name__ == "__main__":
    main()
