"""SDPB 3.1's output files (the reference's sdpvibe/sdpb_output.py interface; the formats are SDPB's,
sdpb/save_solution.cxx and SDP_Solver/run/print_iteration.cxx):

  out.txt             terminateReason, objectives, gap, errors and "Solver runtime" (s)
  iterations.json     one record per iteration, written as the run goes ('[' first, ']' at the end;
                      a file left by an earlier run is kept as iterations.<k>.json)
  c_minus_By/...json  c - B y per block
  y.txt, z.txt, x_<j>.txt, X_matrix_<2j+p>.txt, Y_matrix_<2j+p>.txt
                      "H W" on the first line, then one value per line, row-major
Numbers are printed as SDPB prints BigFloats (fmt_big: %g with the solver's digit count).
"""
import json
import os

import mpmath as mp

from .ipm import fmt_big


def _values(A):
    """MPF -> row-major list of exact mpmath numbers"""
    return list(A.to_mp().reshape(-1))


def _write_values(path, h, w, vals, digits):
    with open(path, "w") as f:
        f.write(f"{h} {w}\n")
        f.writelines(fmt_big(v, digits) + "\n" for v in vals)


def write_text_matrix(path, A, digits, shape=None):
    h, w = shape or (A.shape if len(A.shape) == 2 else (A.shape[0], 1))
    _write_values(path, h, w, _values(A), digits)


def read_text_matrix(path, L):
    """'H W' then H W values -> MPF (H, W) on the GPU"""
    from . import mpf as F
    with open(path) as f:
        h, w = (int(t) for t in f.readline().split())
        vals = f.read().split()
    if len(vals) != h * w:
        raise ValueError(f"{path}: expected {h * w} values, found {len(vals)}")
    return F.zeros((h, w), L) if h * w == 0 else F.from_strings(vals, L, (h, w))


def write_vec(path, vals, digits):
    _write_values(path, len(vals), 1, vals, digits)


def write_out_txt(out_dir, res, digits, runtime_s):
    def f(key):
        return fmt_big(res[key], digits)
    lines = [f'terminateReason = "{res["terminateReason"]}";',
             f"primalObjective = {f('primalObjective')};",
             f"dualObjective   = {f('dualObjective')};",
             f"dualityGap      = {f('dualityGap')};",
             f"primalError     = {f('primalError')};",
             f"dualError       = {f('dualError')};",
             f"Solver runtime  = {int(runtime_s)};"]
    with open(os.path.join(out_dir, "out.txt"), "w") as o:
        o.write("\n".join(lines) + "\n")


def z_from_y(y, normalization):
    """SDPB's z: y with one entry inserted at the first largest |normalization| position, chosen so
    that normalization . z = 1"""
    n = [mp.mpf(v) for v in normalization]
    if len(n) != len(y) + 1:
        raise ValueError(f"normalization has {len(n)} entries, y has {len(y)}")
    k = max(range(len(n)), key=lambda i: abs(n[i]))
    z = list(y[:k]) + [mp.mpf(0)] + list(y[k:])
    z[k] = (1 - mp.fsum(a * b for a, b in zip(n, z))) / n[k]
    return z


def write_solution(out_dir, state, parts, digits, normalization=None):
    """the files of SDPB's --writeSolution: parts a subset of {'x', 'y', 'z', 'X', 'Y'}; state holds
    y (MPF (N, 1)), x (list of per-block MPF (P_j, 1)), X and Y (lists of per-block [even, odd] MPF)"""
    os.makedirs(out_dir, exist_ok=True)
    if {"y", "z"} & set(parts):
        y = _values(state["y"])
        if "y" in parts:
            write_vec(os.path.join(out_dir, "y.txt"), y, digits)
        if "z" in parts:
            if normalization is None:
                raise ValueError("Please provide SDP with valid normalization.json or exclude z from "
                                 "--writeSolution arguments.")
            write_vec(os.path.join(out_dir, "z.txt"), z_from_y(y, normalization), digits)
    for j, xj in enumerate(state["x"]):
        if "x" in parts:
            write_text_matrix(os.path.join(out_dir, f"x_{j}.txt"), xj, digits, (xj.shape[0], 1))
        for tag in ("X", "Y"):
            if tag not in parts:
                continue
            for p, M in enumerate(state[tag][j]):
                if M.shape[0]:
                    write_text_matrix(os.path.join(out_dir, f"{tag}_matrix_{2 * j + p}.txt"), M, digits)


def write_c_minus_By(path, blocks, digits):
    """blocks: per block (SDP order) the MPF vector c - B y"""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"c_minus_By": [[fmt_big(v, digits) for v in _values(b)] for b in blocks]}, f)


class IterationsJson:
    """iterations.json written as SDPB writes it"""
    KEYS = ["mu", "P-obj", "D-obj", "gap", "P-err", "p-err", "D-err", "R-err", "P-step", "D-step", "beta",
            "Q_cond_number", "max_block_cond_number", "block_name"]

    def __init__(self, out_dir):
        os.makedirs(out_dir, exist_ok=True)
        self.path = os.path.join(out_dir, "iterations.json")
        if os.path.exists(self.path):                       # an earlier run's file is kept
            k = 0
            while os.path.exists(os.path.join(out_dir, f"iterations.{k}.json")):
                k += 1
            os.rename(self.path, os.path.join(out_dir, f"iterations.{k}.json"))
            cm = os.path.join(out_dir, "c_minus_By")
            if os.path.exists(cm):
                os.rename(cm, f"{cm}.{k}")
        with open(self.path, "w") as f:
            f.write("[")
        self.n = 0

    def append(self, row):
        fields = [f'"iteration":{row["iteration"]}', f'"total_time": {row["total_time"]:.3f}',
                  f'"iter_time": {row["iter_time"]:.3f}'] + [f'"{k}": "{row[k]}"' for k in self.KEYS]
        with open(self.path, "a") as f:
            f.write(("," if self.n else "") + "\n{ " + ", ".join(fields) + " }")
        self.n += 1

    def close(self):
        with open(self.path, "a") as f:
            f.write("\n]")
