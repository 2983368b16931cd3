"""Solver checkpoints with SDPB's conventions (SDP_Solver/save_checkpoint.cxx, load_checkpoint/; the
reference's sdpvibe/checkpoint.py:1-130 offers the same save / load):

  * binary generations: <dir>/checkpoint_<k>.npz holds the exact limb planes and exponents of y and of
    every block's x_j, X_j,p, Y_j,p (state.get_state's layout) and a JSON header (structure, L,
    precision, options); <dir>/checkpoint.json names the current generation and the backup. A save
    writes generation k+1 under a temporary name, renames it, rewrites the index (also via a
    temporary file), then deletes the generation older than the new backup: a reader sees the old
    pair or the new one, never a half-written file.
  * text: a directory with y.txt, x_<j>.txt, X_matrix_<2j+p>.txt, Y_matrix_<2j+p>.txt (what
    --writeSolution=x,y,X,Y writes, here and in SDPB) is read as a checkpoint too.
load tries the current generation, the backup, then the text form. Values saved at another limb
count are renormalised to the run's L (exact when the run has at least as many limbs).
As in SDPB, a checkpoint holds the iterate only: a resumed run counts its iterations from 1 and
starts with step lengths 0, and its iterates equal those of the uninterrupted run bitwise.
"""
import json
import os

import numpy as np
import torch

from .mpf import MPF

FORMAT = "sdpvibe_amd checkpoint"
INDEX = "checkpoint.json"


def _gen_path(ck_dir, k):
    return os.path.join(ck_dir, f"checkpoint_{k}.npz")


def structure(sdp):
    """what a checkpoint must match: N and each block's (dim, K, even rows, odd rows)"""
    blocks = [None] * sdp.nblocks
    for g in sdp.groups:
        for q, j in enumerate(g.idx):
            rows = [int(g.Xmask[par * g.nb + q].sum()) // g.dim for par in range(2)]
            blocks[j] = [g.dim, g.K] + rows
    return {"N": sdp.N, "blocks": blocks}


def _entries(state):
    """(name, MPF) for every matrix of a state, in a fixed order"""
    yield "y", state["y"]
    for j, x in enumerate(state["x"]):
        yield f"x_{j}", x
        for tag in ("X", "Y"):
            for par, M in enumerate(state[tag][j]):
                yield f"{tag}_{2 * j + par}", M


def _read_index(ck_dir):
    try:
        with open(os.path.join(ck_dir, INDEX)) as f:
            meta = json.load(f)
    except (OSError, ValueError):
        return None
    return meta if meta.get("format") == FORMAT else None


def save(ck_dir, state, sdp, L, precision, options=None):
    """write `state` (state.get_state) as the next generation in ck_dir"""
    os.makedirs(ck_dir, exist_ok=True)
    old = _read_index(ck_dir) or {}
    k = int(old.get("current") or 0) + 1
    header = dict(format=FORMAT, L=L, precision=precision, structure=structure(sdp),
                  iteration=int(state.get("iteration", 0)), options=options or {})
    arrays = {"header": np.array(json.dumps(header))}
    for name, M in _entries(state):
        arrays[name + ".m"] = M.m.detach().cpu().numpy()
        arrays[name + ".e"] = M.e.detach().cpu().numpy()
    path = _gen_path(ck_dir, k)
    with open(path + ".tmp", "wb") as f:
        np.savez(f, **arrays)
    os.replace(path + ".tmp", path)
    index = dict(format=FORMAT, current=k, backup=old.get("current"), options=options or {})
    tmp = os.path.join(ck_dir, INDEX + ".tmp")
    with open(tmp, "w") as f:
        json.dump(index, f, indent=1)
    os.replace(tmp, os.path.join(ck_dir, INDEX))
    stale = old.get("backup")
    if stale is not None and os.path.exists(_gen_path(ck_dir, stale)):
        os.remove(_gen_path(ck_dir, stale))
    return path


def _to_limbs(m, e, L, device):
    """host arrays of a saved MPF -> MPF with L limbs on `device`"""
    m, e = torch.from_numpy(m).to(device), torch.from_numpy(e).to(device)
    if m.shape[0] == L:
        return MPF(m, e)
    from .mpf import normalize
    shape = tuple(e.shape)
    return normalize(m.reshape(m.shape[0], -1), e.reshape(-1), L).reshape(*shape)


def _load_generation(path, sdp, L, device):
    with np.load(path, allow_pickle=False) as z:
        header = json.loads(str(z["header"]))
        if header.get("format") != FORMAT:
            return None
        if header["structure"] != structure(sdp):
            raise ValueError(f"checkpoint {path} was written for another SDP (block structure or N differs)")
        get = lambda name: _to_limbs(z[name + ".m"], z[name + ".e"], L, device)
        st = dict(y=get("y"), x=[], X=[], Y=[], iteration=header.get("iteration", 0), source=path)
        for j in range(sdp.nblocks):
            st["x"].append(get(f"x_{j}"))
            for tag in ("X", "Y"):
                st[tag].append([get(f"{tag}_{2 * j + par}") for par in range(2)])
    return st


def _load_text(ck_dir, sdp, L):
    from . import mpf as F
    from .sdpb_output import read_text_matrix
    if not all(os.path.exists(os.path.join(ck_dir, f)) for f in ("y.txt", "x_0.txt")):
        return None
    shapes = structure(sdp)["blocks"]
    st = dict(y=read_text_matrix(os.path.join(ck_dir, "y.txt"), L), x=[], X=[], Y=[], iteration=0,
              source=ck_dir + " (text)")
    for j, (dim, _, he, ho) in enumerate(shapes):
        st["x"].append(read_text_matrix(os.path.join(ck_dir, f"x_{j}.txt"), L))
        for tag in ("X", "Y"):
            mats = []
            for par, h in enumerate((he, ho)):
                f = os.path.join(ck_dir, f"{tag}_matrix_{2 * j + par}.txt")
                if dim * h == 0:
                    mats.append(F.zeros((0, 0), L))
                elif os.path.exists(f):
                    mats.append(read_text_matrix(f, L))
                else:
                    return None
            st[tag].append(mats)
    return st


def load(ck_dir, sdp, L):
    """the state saved in ck_dir (binary, else text), or None when it holds no checkpoint;
    ValueError when the checkpoint belongs to an SDP of another structure"""
    if not ck_dir or not os.path.isdir(ck_dir):
        return None
    from .rns import dev
    index = _read_index(ck_dir)
    if index is not None:
        for k in (index.get("current"), index.get("backup")):
            if k is not None and os.path.exists(_gen_path(ck_dir, k)):
                st = _load_generation(_gen_path(ck_dir, k), sdp, L, dev)
                if st is not None:
                    return st
    return _load_text(ck_dir, sdp, L)
