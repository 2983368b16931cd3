"""MI355X-native (gfx950) multiprecision linear-algebra core with the sdpvibe module API.

Modules with the reference's names: limbops (limb kernels), rns (residue context, split, products,
reconstruction), mpmat (fixed-point matrices, exact GEMM), mpf (floating-point arrays, GEMM,
Cholesky, Gram accumulation), batched (batched fixed point, recursive Cholesky and solves),
basecase (tile kernels), linalg (single-matrix Cholesky and solves), blockchol (tiled Cholesky),
solver (the interior-point driver and step length). Implementation modules: ipm, sdp, steplen,
blockops, synthetic, graph. Device work runs in libsdpvibe_hip.so (hand-written HIP kernels behind a
C ABI, include/sdpvibe_hip.h); build it with `python -m sdpvibe_amd.build`.
"""
__all__ = ["limbops", "rns", "mpmat", "mpf", "batched", "basecase", "linalg", "blockchol", "solver",
           "ipm", "sdp", "steplen", "blockops", "synthetic", "graph"]
