"""The sdpb command line (sdpvibe_amd/cli.py) end to end on the golden polynomial SDP written in
pmp2sdp's binary format, with the feature checks of the reference's drop-in suite
(tests/dropin_suite.py:150-205): a run resumed from a binary checkpoint repeats the uninterrupted
run's iterations bitwise; a text checkpoint (--writeSolution=x,y,X,Y) resumes it; a zip archive of
the SDP runs bitwise as the directory; --maxRuntime, SIGTERM (exit 15, checkpoint, resume) and
--paramFile behave as sdpb's.
"""
import json
import os
import signal
import zipfile

import mpmath as mp
import pytest

from sdpvibe_amd import cli
from sdpvibe_amd import sdpb_output as O
from tests.test_gpu_sdp_io import write_golden_sdp

pytestmark = pytest.mark.gpu
KEYS = ["mu", "P-obj", "D-obj", "gap", "P-err", "p-err", "D-err", "P-step", "D-step", "beta"]


@pytest.fixture(scope="module")
def golden(tmp_path_factory):
    root = tmp_path_factory.mktemp("cli")
    L, _, _ = write_golden_sdp(root / "sdp")
    precision = 32 * (L - 1) - 2
    assert cli.limbs_for(precision) == L
    return root, str(root / "sdp"), str(precision)


def _rows(out):
    return json.load(open(os.path.join(out, "iterations.json")))


def _out_txt(out):
    res = {}
    for line in open(os.path.join(out, "out.txt")):
        k, _, v = line.partition("=")
        res[k.strip()] = v.strip().rstrip(";").strip('"')
    return res


def _same(a, b):
    return len(a) == len(b) and all(x[k] == y[k] for x, y in zip(a, b) for k in KEYS)


def _run(root, sdp, prec, name, *extra):
    out = str(root / name)
    rc = cli.main(["-s", sdp, "-o", out, "--precision", prec, "--verbosity", "0", *extra])
    return rc, out


def test_resume_from_binary_checkpoint_is_bitwise(golden):
    root, sdp, prec = golden
    rc, full = _run(root, sdp, prec, "full", "-c", str(root / "full.ck"), "--maxIterations", "6", "--noFinalCheckpoint")
    assert rc == 0 and _out_txt(full)["terminateReason"] == "maxIterations exceeded"
    assert not os.path.exists(root / "full.ck")
    rc, part = _run(root, sdp, prec, "part", "-c", str(root / "res.ck"), "--maxIterations", "3")
    assert rc == 0 and os.path.exists(root / "res.ck" / "checkpoint.json")
    rc, rest = _run(root, sdp, prec, "rest", "-c", str(root / "res.ck"), "--maxIterations", "3")
    assert rc == 0
    assert [r["iteration"] for r in _rows(rest)] == [1, 2, 3]          # sdpb counts from 1 after a resume
    assert _same(_rows(full), _rows(part) + _rows(rest))
    a, b = _out_txt(full), _out_txt(rest)
    for k in ("terminateReason", "primalObjective", "dualObjective", "dualityGap", "primalError", "dualError"):
        assert a[k] == b[k], k
    # the final checkpoint of `rest` is generation 2; generation 1 stays as the backup
    idx = json.load(open(root / "res.ck" / "checkpoint.json"))
    assert (idx["current"], idx["backup"]) == (2, 1)
    assert sorted(os.listdir(root / "res.ck")) == ["checkpoint.json", "checkpoint_1.npz", "checkpoint_2.npz"]
    # c - B y, x and y of the two runs are the same text
    for f in ("y.txt", "x_0.txt", "c_minus_By/c_minus_By.json"):
        assert open(os.path.join(full, f)).read() == open(os.path.join(rest, f)).read(), f


def test_text_checkpoint_resume(golden):
    root, sdp, prec = golden
    rc, full = _run(root, sdp, prec, "tfull", "-c", str(root / "tfull.ck"), "--maxIterations", "6", "--noFinalCheckpoint")
    rc, t3 = _run(root, sdp, prec, "t3", "-c", str(root / "t3.ck"), "--maxIterations", "3", "--noFinalCheckpoint",
                  "--writeSolution=x,y,X,Y")
    assert rc == 0 and os.path.exists(os.path.join(t3, "X_matrix_0.txt"))
    rc, rest = _run(root, sdp, prec, "trest", "-c", str(root / "trest.ck"), "-i", t3, "--maxIterations", "3",
                    "--noFinalCheckpoint")
    assert rc == 0
    # the text files hold the iterate to `digits` decimal digits: the resumed run follows the full run
    # to about that many digits, not bitwise
    mp.mp.prec = 800
    for ra, rb in zip(_rows(full)[3:], _rows(rest)):
        for k in ("P-obj", "D-obj", "mu"):
            a, b = mp.mpf(ra[k]), mp.mpf(rb[k])
            assert abs(a - b) <= mp.mpf(10) ** -150 * abs(a), (ra["iteration"], k)


def test_zip_archive_runs_bitwise(golden):
    root, sdp, prec = golden
    arc = str(root / "sdp.zip")
    with zipfile.ZipFile(arc, "w") as z:
        for f in sorted(os.listdir(sdp)):
            z.write(os.path.join(sdp, f), f)
    rc, a = _run(root, sdp, prec, "zdir", "-c", str(root / "zdir.ck"), "--maxIterations", "3", "--noFinalCheckpoint")
    rc2, b = _run(root, arc, prec, "zarc", "-c", str(root / "zarc.ck"), "--maxIterations", "3", "--noFinalCheckpoint")
    assert rc == rc2 == 0 and _same(_rows(a), _rows(b))


def test_max_runtime_and_param_file(golden):
    root, sdp, prec = golden
    rc, out = _run(root, sdp, prec, "rt", "-c", str(root / "rt.ck"), "--maxRuntime", "0", "--noFinalCheckpoint")
    assert rc == 0 and _out_txt(out)["terminateReason"] == "maxRuntime exceeded" and _rows(out) == []
    pf = root / "params.txt"
    pf.write_text(f"precision = {prec}\nmaxIterations = 2\n# a comment\nverbosity = 0\n")
    out = str(root / "pf")
    rc = cli.main(["-s", sdp, "-o", out, "-c", str(root / "pf.ck"), "-p", str(pf), "--noFinalCheckpoint"])
    assert rc == 0 and len(_rows(out)) == 2 and _out_txt(out)["terminateReason"] == "maxIterations exceeded"


def test_sigterm_stops_checkpoints_and_resumes(golden, monkeypatch):
    root, sdp, prec = golden
    rc, full = _run(root, sdp, prec, "sfull", "-c", str(root / "sfull.ck"), "--maxIterations", "5", "--noFinalCheckpoint")
    append = O.IterationsJson.append
    before = signal.getsignal(signal.SIGTERM)

    def append_then_signal(self, row):
        append(self, row)
        if row["iteration"] == 2:
            handler = signal.getsignal(signal.SIGTERM)
            assert callable(handler) and handler is not before, "the run installed no SIGTERM handler"
            os.kill(os.getpid(), signal.SIGTERM)          # the run's handler only sets a flag
    monkeypatch.setattr(O.IterationsJson, "append", append_then_signal)
    rc, sig = _run(root, sdp, prec, "sig", "-c", str(root / "sig.ck"), "--maxIterations", "5")
    monkeypatch.setattr(O.IterationsJson, "append", append)
    assert rc == cli.SIGTERM_EXIT
    assert _out_txt(sig)["terminateReason"] == "SIGTERM signal received" and len(_rows(sig)) == 2
    assert os.path.exists(root / "sig.ck" / "checkpoint.json")
    assert signal.getsignal(signal.SIGTERM) is before                # restored after the run
    rc, rest = _run(root, sdp, prec, "sig2", "-c", str(root / "sig.ck"), "--maxIterations", "3")
    assert rc == 0 and _same(_rows(full), _rows(sig) + _rows(rest))


def test_input_errors(golden, capsys):
    root, sdp, prec = golden
    rc, _ = _run(root, sdp, prec, "e1", "-i", str(root / "no_such.ck"))
    assert rc == 1 and "Unable to load checkpoint" in capsys.readouterr().err
    rc, _ = _run(root, sdp, prec, "e2", "-c", str(root / "e2.ck"), "--writeSolution=x,z")
    assert rc == 1 and "normalization.json" in capsys.readouterr().err
    rc, _ = _run(root, str(root / "missing_sdp"), prec, "e3")
    assert rc == 1
