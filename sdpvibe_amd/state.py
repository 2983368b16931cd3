"""The interior-point iterate as plain per-block data (the reference's sdpvibe/state.py interface:
get_state / set_state / c_minus_By_blocks), for checkpoints and SDPB's solution files.

A state is a dict:
    x:  list over blocks of MPF (P_j, 1)
    X, Y: list over blocks of [even, odd] MPF (d_j h_par, d_j h_par), padding rows removed
    y:  MPF (N, 1)
    iteration: iterations completed (information only)
The solver's pending step (computed at the end of an iteration, applied at the start of the next)
is applied before the state is read, which is the addition the next iteration would make; a solver
given the state by set_state therefore computes the same next iterate bitwise. As SDPB does after
loading a checkpoint, set_state restarts the iteration count and the last step lengths at 0.
"""
import mpmath as mp
import torch

from .mpf import MPF, add, gemm


def _rows(M, rows):
    """rows/columns `rows` (bool (n,)) of the 2-d MPF M"""
    idx = torch.nonzero(rows).reshape(-1)
    return MPF(M.m[:, idx][:, :, idx].contiguous(), M.e[idx][:, idx].contiguous())


def get_state(solver):
    solver._apply_update()
    sdp = solver.sdp
    xs, Xs, Ys = [], [], []
    for g, st in zip(sdp.groups, solver.state):
        for q in range(g.nb):
            xs.append(MPF(st.x.m[:, q].contiguous(), st.x.e[q].contiguous()))
            pX, pY = [], []
            for par in range(2):
                z = par * g.nb + q
                keep = g.Xmask[z]
                pX.append(_rows(MPF(st.X.m[:, z], st.X.e[z]), keep))
                pY.append(_rows(MPF(st.Y.m[:, z], st.Y.e[z]), keep))
            Xs.append(pX)
            Ys.append(pY)
    return dict(x=xs, X=Xs, Y=Ys, y=solver.y, iteration=getattr(solver, "iterations_done", 0))


def set_state(solver, st):
    """load a state (as get_state returns it) into a solver of the same SDP"""
    sdp = solver.sdp
    if len(st["x"]) != sdp.nblocks:
        raise ValueError(f"checkpoint has {len(st['x'])} blocks, the SDP {sdp.nblocks}")
    j = 0
    for g, s in zip(sdp.groups, solver.state):
        for q in range(g.nb):
            if tuple(st["x"][j].shape) != (g.P, 1):
                raise ValueError(f"block {g.idx[q]}: x has shape {tuple(st['x'][j].shape)}, expected {(g.P, 1)}")
            s.x.m[:, q] = st["x"][j].m
            s.x.e[q] = st["x"][j].e
            for par in range(2):
                z = par * g.nb + q
                idx = torch.nonzero(g.Xmask[z]).reshape(-1)
                for tgt, src in ((s.X, st["X"][j][par]), (s.Y, st["Y"][j][par])):
                    if src.shape[0] != idx.numel():
                        raise ValueError(f"block {g.idx[q]}: matrix of size {src.shape[0]}, expected {idx.numel()}")
                    sub_m = tgt.m[:, z]
                    sub_e = tgt.e[z]
                    sub_m[:, idx[:, None], idx[None, :]] = src.m
                    sub_e[idx[:, None], idx[None, :]] = src.e
            j += 1
    if tuple(st["y"].shape) != (sdp.N, 1):
        raise ValueError(f"y has shape {tuple(st['y'].shape)}, the SDP has N={sdp.N}")
    solver.y = MPF(st["y"].m.clone(), st["y"].e.clone())
    solver._update = None
    solver.final_factors = None
    solver.iterations_done, solver.alphaP, solver.alphaD = 0, mp.mpf(0), mp.mpf(0)


def c_minus_By_blocks(solver):
    """c - B y per block (MPF (P_j, 1)), in block order"""
    out = []
    y = solver.y
    for g in solver.sdp.groups:
        ys = MPF(y.m[:, None].expand(-1, g.nb, -1, -1).contiguous(), y.e[None].expand(g.nb, -1, -1).contiguous())
        r = add(g.c, gemm(solver.sdp.B_of(g), ys), -1)
        out += [MPF(r.m[:, q].contiguous(), r.e[q].contiguous()) for q in range(g.nb)]
    return out
