"""SDPB's SDP input format (the output of pmp2sdp) and its binary BigFloat records (the reference's
sdpvibe/sdp_io.py interface: SdpSource, read_block, read_objectives, read_normalization,
write_block_bin, write_sdp).

An SDP is a directory or an archive (zip, tar, tar.gz, ...) holding control.json, objectives.json,
[normalization.json], block_info_<j>.json and block_data_<j>.bin or block_data_<j>.json. A binary
block is a boost::archive::binary_oarchive of, in order: the matrix B (El::Matrix<BigFloat>,
column-major with a leading dimension), the vector c, and the two bilinear-basis matrices. BigFloat
records are decoded and encoded by the library's host C++ codec (sdpv_decode_bigfloats /
sdpv_encode_bigfloats); decimal strings go through mpf.strings_to_host.

Blocks come back in the host format of the solver: (limbs (n, L) int32 entry-major, exponents (n,))
pairs, matrices row-major.
"""
import ctypes
import json
import os
import struct
import tarfile
import zipfile

import numpy as np

from ._lib import lib

EZ = -(1 << 40)
_SIG = b"serialization::archive"


class SdpSource:
    """read access to an SDP directory or archive (archives may hold the files in one top-level
    directory)"""

    def __init__(self, path):
        self.path = os.fspath(path)
        if os.path.isdir(self.path):
            self.kind = "dir"
        elif os.path.isfile(self.path) and zipfile.is_zipfile(self.path):
            self.kind = "zip"
        elif os.path.isfile(self.path) and tarfile.is_tarfile(self.path):
            self.kind = "tar"
        else:
            raise FileNotFoundError(f"SDP not found or not a directory/zip/tar archive: {self.path}")
        self._arch = None
        self._index = None

    def _members(self):
        """name within the SDP -> member name in the archive"""
        if self._index is None:
            if self.kind == "dir":
                self._index = {n: n for n in os.listdir(self.path)}
            else:
                if self._arch is None:
                    self._arch = zipfile.ZipFile(self.path) if self.kind == "zip" else tarfile.open(self.path)
                names = self._arch.namelist() if self.kind == "zip" else self._arch.getnames()
                ctl = [n for n in names if n.endswith("control.json")]
                prefix = ctl[0][:-len("control.json")] if ctl else ""
                self._index = {n[len(prefix):]: n for n in names
                               if n.startswith(prefix) and "/" not in n[len(prefix):]}
        return self._index

    def names(self):
        return set(self._members())

    def exists(self, name):
        return name in self._members()

    def read(self, name):
        member = self._members()[name]
        if self.kind == "dir":
            with open(os.path.join(self.path, member), "rb") as f:
                return f.read()
        if self.kind == "zip":
            return self._arch.read(member)
        return self._arch.extractfile(member).read()

    def json(self, name):
        return json.loads(self.read(name))

    def mtime(self, name="control.json"):
        return os.path.getmtime(os.path.join(self.path, name) if self.kind == "dir" else self.path)


# ------------------------------------------------------------------ BigFloat records
def _codec():
    h = lib()
    return h.sdpv_decode_bigfloats, h.sdpv_encode_bigfloats


def decode_bigfloats(buf, pos, n, nl, L):
    """n records from byte pos of buf -> (limbs (n, L), exponents (n,), position after them)"""
    dec, _ = _codec()
    limbs = np.empty((n, L), dtype=np.int32)
    exps = np.empty(n, dtype=np.int64)
    arr = np.frombuffer(buf, dtype=np.uint8)
    q = dec(arr.ctypes.data_as(ctypes.c_void_p), len(buf), pos, n, nl, L, limbs.ctypes.data_as(ctypes.c_void_p),
            exps.ctypes.data_as(ctypes.c_void_p), EZ)
    if q < 0:
        raise ValueError("truncated or corrupt BigFloat records")
    return limbs, exps, q


def encode_bigfloats(limbs, exps, nl, L):
    """(limbs (n, L), exponents (n,)) -> bytes of n records with nl GMP limbs each"""
    _, enc = _codec()
    limbs = np.ascontiguousarray(limbs, dtype=np.int32)
    exps = np.ascontiguousarray(exps, dtype=np.int64)
    n = len(exps)
    out = np.empty(n * (13 + 8 * nl) + 16, dtype=np.uint8)
    k = enc(limbs.ctypes.data_as(ctypes.c_void_p), exps.ctypes.data_as(ctypes.c_void_p), n, L, nl,
            out.ctypes.data_as(ctypes.c_void_p), len(out), EZ)
    if k < 0:
        raise ValueError("cannot encode: a mantissa needs more limbs than the archive precision gives")
    return out[:k].tobytes()


def gmp_limbs(precision):
    """GMP mpf_t limbs of a `precision`-bit BigFloat (mpf_set_prec rounds up and adds one)"""
    return (precision + 127) // 64 + 1


class _Archive:
    """cursor over a binary block archive; class-info headers appear before the first object of
    each class"""

    def __init__(self, buf, L):
        self.buf, self.L = buf, L
        (n,) = struct.unpack_from("<Q", buf, 0)
        if buf[8:8 + n] != _SIG:
            raise ValueError("not a boost binary archive")
        p = 8 + n + 2 + 4 + 4                              # signature, library version, type sizes, endianness
        (self.precision,) = struct.unpack_from("<Q", buf, p)
        self.pos = p + 8
        self.nl = gmp_limbs(self.precision)
        self.seen = set()

    def _class_info(self, kind):
        if kind not in self.seen:
            self.pos += 5                                  # tracking flag + class version
            self.seen.add(kind)

    def floats(self, n):
        if n == 0:
            return np.zeros((0, self.L), np.int32), np.zeros(0, np.int64)
        self._class_info("bigfloat")
        limbs, exps, self.pos = decode_bigfloats(self.buf, self.pos, n, self.nl, self.L)
        return limbs, exps

    def matrix(self):
        """El::Matrix (column-major, leading dimension ld) -> (h, w, limbs (h w, L) row-major, exps)"""
        self._class_info("matrix")
        h, w, ld = struct.unpack_from("<iii", self.buf, self.pos)
        self.pos += 12
        limbs, exps = self.floats(ld * w)
        limbs = limbs.reshape(w, ld, self.L)[:, :h].transpose(1, 0, 2).reshape(h * w, self.L)
        exps = exps.reshape(w, ld)[:, :h].T.reshape(-1)
        return h, w, np.ascontiguousarray(limbs), np.ascontiguousarray(exps)

    def vector(self):
        self._class_info("vector")
        (n,) = struct.unpack_from("<Q", self.buf, self.pos)
        self.pos += 8 + 4                                  # length, item version
        return self.floats(n)


# ------------------------------------------------------------------ blocks and objectives
def read_block(src, j, L):
    """block j -> dict(dim, K, bilinear_bases_even=(h, (limbs, exps) or None), bilinear_bases_odd, c, B)
    with B row-major ((K dim(dim+1)/2) x N)"""
    from .mpf import strings_to_host
    info = src.json(f"block_info_{j}.json")
    out = {"dim": info["dim"], "K": info["num_points"]}
    if src.exists(f"block_data_{j}.bin"):
        ar = _Archive(src.read(f"block_data_{j}.bin"), L)
        _, _, Bl, Be = ar.matrix()
        out["c"] = ar.vector()
        for key in ("bilinear_bases_even", "bilinear_bases_odd"):
            h, _, bl, be = ar.matrix()
            out[key] = (h, (bl, be) if h else None)
        if ar.pos != len(ar.buf):
            raise ValueError(f"block_data_{j}.bin: {len(ar.buf) - ar.pos} unexpected trailing bytes")
        out["B"] = (Bl, Be)
        out["precision"] = ar.precision
        return out
    data = src.json(f"block_data_{j}.json")
    for key in ("bilinear_bases_even", "bilinear_bases_odd"):
        rows = data[key]
        out[key] = (len(rows), strings_to_host([x for r in rows for x in r], L) if rows else None)
    out["c"] = strings_to_host(data["c"], L)
    out["B"] = strings_to_host([x for r in data["B"] for x in r], L)
    return out


def read_objectives(src):
    """(constant, b) as decimal strings"""
    obj = src.json("objectives.json")
    return obj.get("constant", "0"), obj["b"]


def read_normalization(src):
    """the normalization vector (decimal strings), or None"""
    return src.json("normalization.json")["normalization"] if src.exists("normalization.json") else None


def write_block_bin(path, blk, precision, L, N):
    """a block in read_block's form -> block_data_<j>.bin (pmp2sdp's binary layout)"""
    nl = gmp_limbs(precision)
    head = struct.pack("<Q", len(_SIG)) + _SIG + struct.pack("<H", 19) + bytes([4, 8, 4, 8]) + struct.pack("<i", 1)
    chunks = [head, struct.pack("<Q", precision)]
    first_float = [True]

    def floats(limbs, exps):
        pre = b""
        if first_float[0]:
            pre, first_float[0] = b"\0\x01\0\0\0", False      # BigFloat class info, version 1
        return pre + encode_bigfloats(limbs, exps, nl, L)

    def matrix(h, w, limbs, exps, class_info):
        body = (b"\0" * 5 if class_info else b"") + struct.pack("<iii", h, w, max(h, 1) if w == 0 else h)
        if h * w:
            cm = limbs.reshape(h, w, L).transpose(1, 0, 2).reshape(h * w, L)     # column-major, ld = h
            body += floats(cm, exps.reshape(h, w).T.reshape(-1))
        return body

    Bl, Be = blk["B"]
    chunks.append(matrix(len(Be) // N, N, Bl, Be, True))
    cl, ce = blk["c"]
    vec = b"\0" * 5 + struct.pack("<Q", len(ce)) + struct.pack("<I", 1)
    if len(ce):
        vec += floats(cl, ce)
    chunks.append(vec)
    for key in ("bilinear_bases_even", "bilinear_bases_odd"):
        h, arr = blk[key]
        chunks.append(struct.pack("<iii", 0, blk["K"], 0) if arr is None else matrix(h, blk["K"], arr[0], arr[1], False))
    with open(path, "wb") as f:
        f.write(b"".join(chunks))


def write_sdp(out_dir, blocks, b_strings, precision, L, constant="0", normalization=None):
    """an SDP directory in pmp2sdp's binary format"""
    os.makedirs(out_dir, exist_ok=True)

    def dump(name, obj):
        with open(os.path.join(out_dir, name), "w") as f:
            json.dump(obj, f)
    dump("control.json", {"num_blocks": len(blocks), "command": "sdpvibe_amd.sdp_io.write_sdp"})
    dump("objectives.json", {"constant": constant, "b": list(b_strings)})
    if normalization is not None:
        dump("normalization.json", {"normalization": list(normalization)})
    N = len(b_strings)
    for j, blk in enumerate(blocks):
        dump(f"block_info_{j}.json", {"dim": blk["dim"], "num_points": blk["K"]})
        write_block_bin(os.path.join(out_dir, f"block_data_{j}.bin"), blk, precision, L, N)
