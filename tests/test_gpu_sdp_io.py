"""An SDP written in pmp2sdp's binary format and read back (sdp_io, SDP(path, L)) drives the GPU solver
to bitwise the same iterations as the same SDP built in memory; the output writers produce SDPB's
out.txt and iterations.json layouts."""
import json
import os

import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd import sdp_io as IO
from sdpvibe_amd.ipm import Solver
from sdpvibe_amd.sdp import SDP
from sdpvibe_amd.sdpb_output import IterationsJson, write_out_txt
from tests import ipm_cases as C
from tests.test_gpu_ipm import FIELDS, gpu_sdp

pytestmark = pytest.mark.gpu


def _host(X):
    L = X.L
    return X.m.reshape(L, -1).t().contiguous().cpu().numpy(), X.e.reshape(-1).cpu().numpy()


def write_golden_sdp(path, normalization=None):
    """the golden polynomial SDP written in pmp2sdp's binary format at `path`; -> (L, blocks, b)"""
    g = C.golden()
    L = g["L"]
    with mp.workprec(1700):
        blocks, b = C.golden_blocks(g)
    host_blocks = []
    for bl in blocks:
        K = bl["K"]
        P, N = len(bl["c"]), len(bl["B"][0])
        ent = {"dim": bl["dim"], "K": K}
        for key, V in (("bilinear_bases_even", bl["Ve"]), ("bilinear_bases_odd", bl["Vo"])):
            ent[key] = (len(V), _host(F.from_mp([x for r in V for x in r], L, (len(V), K)))) if V else (0, None)
        ent["c"] = _host(F.from_mp(bl["c"], L, (P, 1)))
        ent["B"] = _host(F.from_mp([x for r in bl["B"] for x in r], L, (P, N)))
        host_blocks.append(ent)
    with mp.workprec(32 * L + 64):
        b_strs = [mp.nstr(v, 240, strip_zeros=False) for v in b]
    IO.write_sdp(str(path), host_blocks, b_strs, 32 * L, L, normalization=normalization)
    return L, blocks, b


def test_binary_sdp_round_trip(tmp_path):
    L, blocks, b = write_golden_sdp(tmp_path / "sdp")
    mem = gpu_sdp(blocks, b, L)
    disk = SDP(str(tmp_path / "sdp"), L)
    assert torch.equal(disk.b.m, mem.b.m) and torch.equal(disk.b.e, mem.b.e)
    A, B = Solver(mem, L), Solver(disk, L)
    A.run(max_iter=4)
    B.run(max_iter=4)
    for ra, rb in zip(A.log, B.log):
        for k in FIELDS.values():
            assert ra[k] == rb[k], (ra["iteration"], k)
    out = tmp_path / "out"
    it = IterationsJson(str(out))
    for row in B.log:
        it.append(row)
    it.close()
    rows = json.load(open(out / "iterations.json"))
    assert [r["iteration"] for r in rows] == [1, 2, 3, 4] and rows[0]["mu"] == B.log[0]["mu"]
    write_out_txt(str(out), B.result, B.digits, 1.0)
    txt = open(out / "out.txt").read()
    assert txt.startswith('terminateReason = "maxIterations exceeded";') and "Solver runtime  = 1;" in txt


@pytest.mark.parametrize("stream,chunk", [("all", 1), ("none", 1)])
def test_streamed_load_iterates_like_the_resident_load(tmp_path, monkeypatch, stream, chunk):
    """SDP(path) under a memory plan of one-block chunks, with every B and P in page-locked host
    memory or all resident: the iterations equal those of the default load bit for bit"""
    L, blocks, b = write_golden_sdp(tmp_path / "sdp")
    base = SDP(str(tmp_path / "sdp"), L)
    monkeypatch.setenv("SDPB_GPU_STREAM", stream)
    placed = SDP(str(tmp_path / "sdp"), L, chunk=chunk)
    assert len(placed.groups) == placed.nblocks > len(base.groups)
    assert (placed.store is not None) == (stream == "all")
    if stream == "all":
        assert all(g.B is None for g in placed.groups)
    assert placed.read_info.endswith(placed.layout_plan.describe()) and "chunks" in placed.read_info
    A, B = Solver(base, L), Solver(placed, L)
    A.run(max_iter=4)
    B.run(max_iter=4)
    for ra, rb in zip(A.log, B.log):
        for k in FIELDS.values():
            assert ra[k] == rb[k], (ra["iteration"], k)
    assert torch.equal(A.y.m, B.y.m) and torch.equal(A.y.e, B.y.e)
