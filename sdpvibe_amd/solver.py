"""The solver module of the drop-in API (sdpvibe/solver.py): the interior-point driver (SDP, Solver,
DEFAULTS, fmt_big, Timers), the Cholesky condition estimates, panelled solves, the step length and
a random-MPF helper, gathered from the modules that implement them."""
import torch

from . import mpf as F
from .blockops import paste, window
from .ipm import DEFAULTS, Solver, Timers, chol_log2_cond, chol_log2_cond_parts, fmt_big
from .sdp import SDP
from .steplen import (EIG_TOL, _mpf_from_float, _resolved, min_eig, refine_min_eig, step_from_eig, step_length,
                      symeig, to_double)

__all__ = ["SDP", "Solver", "DEFAULTS", "Timers", "fmt_big", "chol_log2_cond", "chol_log2_cond_parts",
           "solve_panels", "min_eig", "refine_min_eig", "step_from_eig", "step_length", "EIG_TOL",
           "symeig", "to_double", "_mpf_from_float", "_rand_mpf"]


def solve_panels(cs, B, panels):
    """L^-1 B computed in `panels` column panels; identical to one solve, since every column is
    aligned and rounded on its own (sdpvibe/solver.py:985-993)"""
    N = B.shape[-1]
    if panels <= 1:
        return cs.solve_lower(B)
    width = -(-N // panels)
    out = F.zeros(B.shape, B.L)
    for j0 in range(0, N, width):
        paste(out, 0, j0, cs.solve_lower(window(B, 0, B.shape[-2], j0, min(N, j0 + width))))
    return out


def _rand_mpf(shape, L, gen, spread=4):
    """random normalised MPF entries of magnitude about 2^[-spread, spread] drawn from the torch
    generator `gen` (random limbs normalised on the GPU) (sdpvibe/solver.py:401-406)"""
    n = 1
    for s in shape:
        n *= s
    limbs = torch.randint(-2 ** 31, 2 ** 31, (L, n), dtype=torch.int32, device=gen.device, generator=gen)
    ex = torch.randint(-spread, spread + 1, (n,), dtype=torch.int64, device=gen.device, generator=gen)
    return F.normalize(limbs, ex - 32 * L, L).reshape(*shape)
