"""The SDPB input format and the library's BigFloat codec (sdpv_decode_bigfloats /
sdpv_encode_bigfloats, the counterparts of sdpvibe/_bigfloat.c:16-111), pinned by:
- the closed form of a record (sdpvibe/sdp_io.py:119-140 states it): value = sign(size) D 2^(64 (exp - |size|)),
  D the first |size| little-endian 64-bit limbs, rounded half up on the magnitude to 32L - 2 bits;
- exact round trips MPF -> records -> MPF and SDP directory / zip archive -> blocks;
- JSON blocks read through the decimal conversion.
CPU only (the codec is host code in the library)."""
import io
import json
import os
import random
import struct
import zipfile

import numpy as np
import pytest

from sdpvibe_amd import sdp_io as IO
from tests.intutil import norm_closed_form, to_planes

EZ = IO.EZ


def record(size, exp, limbs, nl):
    d = list(limbs) + [random.getrandbits(64) for _ in range(nl - len(limbs))]   # meaningless tail
    return b"\0" + struct.pack("<iq", size, exp) + struct.pack(f"<{nl}Q", *d)


@pytest.mark.parametrize("L,nl", [(4, 3), (8, 5), (24, 14), (5, 1)])
def test_decode_closed_form(L, nl):
    rng = random.Random(L * 31 + nl)
    NB = 32 * L - 2
    recs, want = [], []
    for i in range(200):
        asz = rng.randint(1, nl)
        limbs = [rng.getrandbits(64) for _ in range(asz)]
        if i % 7 == 0:
            limbs[-1] = 0                                  # top limb zero: fewer significant bits
        if i % 11 == 0:
            limbs = [0] * asz                              # a zero stored with a nonzero size
        neg = rng.random() < 0.5
        ex = rng.randint(-50, 50)
        recs.append(record(-asz if neg else asz, ex, limbs, nl))
        D = sum(v << (64 * k) for k, v in enumerate(limbs))
        if D == 0:
            want.append((0, EZ))
            continue
        s = max(0, D.bit_length() - NB)
        m = (D + (1 << (s - 1))) >> s if s else D
        want.append((-m if neg else m, 64 * (ex - asz) + s))
        if i % 5 == 0:
            recs.append(b"\1")                             # an is_zero record
            want.append((0, EZ))
    buf = b"junk" + b"".join(recs)
    limbs, exps, pos = IO.decode_bigfloats(buf, 4, len(want), nl, L)
    assert pos == len(buf)
    got = [int.from_bytes(limbs[i].tobytes(), "little", signed=True) for i in range(len(want))]
    for g, e_, (wm, we) in zip(got, exps, want):
        assert (g, int(e_)) == (wm, we)


def test_decode_rejects_short_and_oversized():
    with pytest.raises(ValueError):
        IO.decode_bigfloats(record(2, 0, [1, 2], 3)[:-1], 0, 1, 3, 4)
    with pytest.raises(ValueError):                       # |size| larger than the limb count
        IO.decode_bigfloats(b"\0" + struct.pack("<iq", 4, 0) + struct.pack("<3Q", 1, 2, 3), 0, 1, 3, 4)


@pytest.mark.parametrize("L", [2, 7, 24])
def test_round_trip_exact(L):
    """normalised MPF entries survive encode -> decode bit for bit (encoding is exact)"""
    rng = random.Random(L)
    NB = 32 * L - 2
    vals, exps = [], []
    for i in range(300):
        v = rng.randrange(1 << (NB - 1), 1 << NB) * rng.choice((1, -1))
        vals.append(v if i % 13 else 0)
        exps.append(rng.randint(-3000, 3000) if i % 13 else EZ)
    limbs = to_planes(vals, L).T.copy()
    e = np.array(exps, np.int64)
    nl = IO.gmp_limbs(32 * L) + 1
    buf = IO.encode_bigfloats(limbs, e, nl, L)
    back, be, pos = IO.decode_bigfloats(buf, 0, len(vals), nl, L)
    assert pos == len(buf)
    assert np.array_equal(back, limbs) and np.array_equal(be, e)


def test_encode_too_small_precision():
    L = 8
    v = [(1 << 200) + 1]
    with pytest.raises(ValueError):
        IO.encode_bigfloats(to_planes(v, L).T.copy(), np.array([0], np.int64), 2, L)


def _block(rng, dim, K, he, ho, N, L):
    NB = 32 * L - 2
    P = K * dim * (dim + 1) // 2

    def arr(n):
        vals = [rng.randrange(1 << (NB - 1), 1 << NB) * rng.choice((1, -1)) for _ in range(n)]
        return to_planes(vals, L).T.copy(), np.array([rng.randint(-900, 900) for _ in range(n)], np.int64)
    return dict(dim=dim, K=K, bilinear_bases_even=(he, arr(he * K) if he else None),
                bilinear_bases_odd=(ho, arr(ho * K) if ho else None), c=arr(P), B=arr(P * N))


def _same_block(a, b):
    assert a["dim"] == b["dim"] and a["K"] == b["K"]
    for key in ("bilinear_bases_even", "bilinear_bases_odd"):
        assert a[key][0] == b[key][0]
        if a[key][1] is not None:
            assert np.array_equal(a[key][1][0], b[key][1][0]) and np.array_equal(a[key][1][1], b[key][1][1])
    for key in ("c", "B"):
        assert np.array_equal(a[key][0], b[key][0]) and np.array_equal(a[key][1], b[key][1])


def test_sdp_directory_and_zip_round_trip(tmp_path):
    L, N = 6, 4
    rng = random.Random(3)
    blocks = [_block(rng, 1, 3, 2, 1, N, L), _block(rng, 2, 2, 1, 0, N, L), _block(rng, 3, 1, 1, 1, N, L)]
    d = tmp_path / "sdp"
    IO.write_sdp(str(d), blocks, ["1.5", "-2", "3e-5", "0"], 32 * L, L, constant="0.25")
    for src in (IO.SdpSource(str(d)), _zipped(d, tmp_path / "sdp.zip")):
        assert IO.read_objectives(src) == ("0.25", ["1.5", "-2", "3e-5", "0"])
        assert IO.read_normalization(src) is None
        for j, blk in enumerate(blocks):
            _same_block(blk, IO.read_block(src, j, L))


def _zipped(d, zpath):
    with zipfile.ZipFile(zpath, "w") as z:
        for name in os.listdir(d):
            z.write(os.path.join(d, name), arcname="inner/" + name)
    return IO.SdpSource(str(zpath))


def test_json_block(tmp_path):
    from sdpvibe_amd.mpf import strings_to_host
    L = 4
    d = tmp_path / "sdpj"
    d.mkdir()
    (d / "control.json").write_text(json.dumps({"num_blocks": 1}))
    (d / "block_info_0.json").write_text(json.dumps({"dim": 1, "num_points": 2}))
    data = {"bilinear_bases_even": [["1", "0.5"]], "bilinear_bases_odd": [], "c": ["2", "-3.25"],
            "B": [["1e-3", "7"], ["-1", "0"]]}
    (d / "block_data_0.json").write_text(json.dumps(data))
    r = IO.read_block(IO.SdpSource(str(d)), 0, L)
    assert r["dim"] == 1 and r["K"] == 2 and r["bilinear_bases_odd"] == (0, None)
    for got, want in ((r["c"], strings_to_host(data["c"], L)),
                      (r["B"], strings_to_host(["1e-3", "7", "-1", "0"], L))):
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
