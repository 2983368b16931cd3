"""The sdpb command line's option handling and the checkpoint files, on the CPU (the solver runs in
test_gpu_cli.py): option forms and precedence as sdpb 3.1.0's (boost program_options: --name value,
--name=value, -s value, -svalue, bare and valued switches, parameter file below the command line),
derived defaults, the precision -> limb count rule of the reference (sdpvibe/cli.py:179) and exact
checkpoint round trips with backup generations.
"""
import json
import os
import types

import numpy as np
import pytest
import torch

from sdpvibe_amd import checkpoint as CK
from sdpvibe_amd import cli
from sdpvibe_amd import rns
from sdpvibe_amd.mpf import MPF


def _resolved(argv):
    params, flags, given = cli.parse_args(argv)
    return cli.resolve(params, given)


def test_option_forms_and_defaults():
    p = _resolved(["-s", "a/sdp", "--precision=768", "--maxIterations", "7", "--findDualFeasible",
                   "--detectPrimalFeasibleJump=false", "-oOUT", "--writeSolution", "x,y,z"])
    assert p["sdpDir"] == "a/sdp" and p["outDir"] == "OUT" and p["precision"] == 768 and p["maxIterations"] == 7
    assert p["findDualFeasible"] is True and p["detectPrimalFeasibleJump"] is False
    assert p["checkpointDir"] == "a/sdp.ck" and p["initialCheckpointDir"] == "a/sdp.ck"
    assert p["requireInitialCheckpoint"] is False and p["writeSolutionParts"] == {"x", "y", "z"}
    assert p["dualityGapThreshold"] == "1e-30" and p["verbosity"] == 1 and p["noFinalCheckpoint"] is False
    # a trailing slash or "." does not change the derived directories
    assert _resolved(["--sdpDir", "a/sdp/"])["outDir"] == "a/sdp_out"
    assert _resolved(["-s", "a/sdp/."])["checkpointDir"] == "a/sdp.ck"
    q = _resolved(["-s", "x", "-i", "ck0", "--verbosity", "debug"])
    assert q["requireInitialCheckpoint"] and q["initialCheckpointDir"] == "ck0" and q["verbosity"] == 2


def test_param_file_is_below_the_command_line(tmp_path):
    f = tmp_path / "params.txt"
    f.write_text("precision = 512\nmaxIterations = 7   # trailing comment\n\n# line comment\nnoFinalCheckpoint = true\n")
    p = _resolved(["-s", "x", "-p", str(f), "--maxIterations", "9"])
    assert p["precision"] == 512 and p["maxIterations"] == 9 and p["noFinalCheckpoint"] is True
    f.write_text("notAnOption = 1\n")
    with pytest.raises(cli.UsageError):
        cli.parse_args(["-s", "x", "-p", str(f)])


@pytest.mark.parametrize("argv", [["--noSuchOption", "1"], ["-s"], ["stray"], ["-s", "x", "--verbosity", "loud"],
                                  ["-s", "x", "--writeSolution", "x,w"], ["-s", "x", "--precision", "many"],
                                  ["--precision", "400"]])
def test_usage_errors_exit_1(argv, capsys):
    assert cli.main(argv) == 1
    assert "Error:" in capsys.readouterr().err


def test_help_and_version(capsys):
    assert cli.main(["--help"]) == 0
    text = capsys.readouterr().out
    assert all(f"--{o.name}" in text for o in cli.OPTION_LIST)
    assert cli.main(["--version"]) == 0
    assert cli.SDPB_VERSION in capsys.readouterr().out


def test_limbs_for_precision():
    # the bits asked for plus two, rounded up to 32-bit limbs, and one guard limb
    assert [cli.limbs_for(p) for p in (400, 512, 734, 768, 1024)] == [14, 18, 24, 26, 34]
    for p in range(64, 2048, 7):
        L = cli.limbs_for(p)
        assert 32 * (L - 2) < p + 2 <= 32 * (L - 1)


def test_solver_params_cover_the_solver_defaults():
    from sdpvibe_amd.ipm import DEFAULTS
    params = _resolved(["-s", "x", "--stepLengthReduction", "0.5", "--maxIterations", "3"])
    sp = cli.solver_params(params)
    assert set(sp) == set(DEFAULTS) and sp["stepLengthReduction"] == "0.5" and sp["maxIterations"] == 3
    rec = cli.options_record(params)
    assert rec["noFinalCheckpoint"] == "false" and "paramFile" not in rec


# ------------------------------------------------------------------ checkpoint files
def _fake_sdp(shapes, N):
    """the attributes checkpoint.structure reads: one group per block of (dim, K, even rows, odd rows)"""
    groups = []
    for j, (dim, K, he, ho) in enumerate(shapes):
        h = max(he, ho)
        mask = torch.zeros((2, dim * h), dtype=torch.bool)
        for c in range(dim):
            mask[0, c * h:c * h + he] = True
            mask[1, c * h:c * h + ho] = True
        groups.append(types.SimpleNamespace(idx=[j], nb=1, dim=dim, K=K, Xmask=mask))
    return types.SimpleNamespace(groups=groups, nblocks=len(shapes), N=N)


def _rand_mpf(rng, L, shape):
    m = torch.from_numpy(rng.integers(-2 ** 31, 2 ** 31, size=(L, *shape), dtype=np.int64).astype(np.int32))
    e = torch.from_numpy(rng.integers(-3000, 3000, size=shape, dtype=np.int64))
    return MPF(m, e)


def _state(rng, shapes, N, L):
    st = dict(y=_rand_mpf(rng, L, (N, 1)), x=[], X=[], Y=[], iteration=4)
    for dim, K, he, ho in shapes:
        st["x"].append(_rand_mpf(rng, L, (K * dim * (dim + 1) // 2, 1)))
        for tag in ("X", "Y"):
            st[tag].append([_rand_mpf(rng, L, (dim * h, dim * h)) for h in (he, ho)])
    return st


def _equal(a, b):
    return torch.equal(a.m, b.m) and torch.equal(a.e, b.e)


def test_checkpoint_round_trip_and_generations(tmp_path, monkeypatch):
    monkeypatch.setattr(rns, "dev", "cpu")
    shapes, N, L = [(1, 5, 3, 2), (2, 4, 2, 2), (3, 2, 1, 0)], 6, 5
    sdp = _fake_sdp(shapes, N)
    rng = np.random.default_rng(7)
    ck = str(tmp_path / "run.ck")
    assert CK.load(ck, sdp, L) is None
    first = _state(rng, shapes, N, L)
    CK.save(ck, first, sdp, L, 94, {"precision": "94"})
    second = _state(rng, shapes, N, L)
    CK.save(ck, second, sdp, L, 94)
    third = _state(rng, shapes, N, L)
    CK.save(ck, third, sdp, L, 94)
    # the current generation and one backup remain
    assert sorted(os.listdir(ck)) == ["checkpoint.json", "checkpoint_2.npz", "checkpoint_3.npz"]
    assert json.load(open(os.path.join(ck, "checkpoint.json")))["current"] == 3
    got = CK.load(ck, sdp, L)
    assert got["source"].endswith("checkpoint_3.npz") and got["iteration"] == 4
    assert _equal(got["y"], third["y"])
    for j in range(len(shapes)):
        assert _equal(got["x"][j], third["x"][j])
        for tag in ("X", "Y"):
            for par in range(2):
                assert _equal(got[tag][j][par], third[tag][j][par])
    # a lost current generation falls back to the backup
    os.remove(os.path.join(ck, "checkpoint_3.npz"))
    assert _equal(CK.load(ck, sdp, L)["y"], second["y"])
    # another SDP structure is refused
    with pytest.raises(ValueError):
        CK.load(ck, _fake_sdp(shapes[:2], N), L)
    with pytest.raises(ValueError):
        CK.load(ck, _fake_sdp(shapes, N + 1), L)


def test_text_checkpoint_needs_y_and_x0(tmp_path, monkeypatch):
    monkeypatch.setattr(rns, "dev", "cpu")
    sdp = _fake_sdp([(1, 2, 1, 1)], 2)
    d = tmp_path / "solution"
    d.mkdir()
    (d / "y.txt").write_text("2 1\n1\n2\n")
    assert CK.load(str(d), sdp, 4) is None          # no x_0.txt: not a checkpoint
