"""The memory plan (sdpvibe_amd/layout.py; the reference's Layout, sdpvibe/solver.py:227-398) on
the CPU: which chunks it forms and which arrays it places in host memory, for given budgets."""
import pytest

from sdpvibe_amd.layout import GB, Layout

SHAPES = [(1, 27, 14)] * 300 + [(3, 27, 14)] * 60         # the N = 3000 benchmark program


def plan(monkeypatch, mode, budget, chunk=None, shapes=SHAPES, N=3000):
    monkeypatch.setenv("SDPB_GPU_STREAM", mode)
    monkeypatch.delenv("SDPB_GPU_MEM_GB", raising=False)
    monkeypatch.delenv("SDPB_GPU_CHUNK_GB", raising=False)
    return Layout(shapes, N, 24, budget=budget, chunk=chunk, n_primes=117)


def covers(lay, shapes=SHAPES):
    flat = [j for c in lay.chunks for j in c]
    assert flat == list(range(len(shapes)))                 # every block once, in order
    for c in lay.chunks:
        assert c == list(range(c[0], c[-1] + 1)) and len({shapes[j] for j in c}) == 1


def test_resident_when_it_fits(monkeypatch):
    lay = plan(monkeypatch, "auto", 250 * GB)
    covers(lay)
    assert [len(c) for c in lay.chunks] == [300, 60] and not lay.host_keys and not lay.streamed()


def test_streams_what_does_not_fit(monkeypatch):
    lay = plan(monkeypatch, "auto", 20 * GB)
    covers(lay)
    assert lay.streamed() and len(lay.chunks) > 2
    assert {n for n, _ in lay.host_keys} <= {"B", "P"}
    # P (read five times an iteration) is kept resident before B (read twice)
    kept = {(n, gi) for gi in range(len(lay.chunks)) for n in ("B", "P")} - lay.host_keys
    assert all(n == "P" for n, _ in kept) or not kept
    resident = sum(lay.items(lay.chunks[gi])[n] for n, gi in kept)
    assert lay.fixed + resident <= lay.budget


def test_all_and_none(monkeypatch):
    lay = plan(monkeypatch, "all", 250 * GB)
    assert lay.host_keys == {(n, gi) for gi in range(len(lay.chunks)) for n in ("B", "P")}
    lay = plan(monkeypatch, "none", 250 * GB, chunk=GB)
    covers(lay)
    assert not lay.host_keys and len(lay.chunks) > 2
    assert all(lay.ws(c) + sum(lay.items(c).values()) <= GB or len(c) == 1 for c in lay.chunks)


def test_chunk_cap_and_env(monkeypatch):
    monkeypatch.setenv("SDPB_GPU_CHUNK_GB", "0.5")
    monkeypatch.setenv("SDPB_GPU_STREAM", "auto")
    lay = Layout(SHAPES, 3000, 24, budget=250 * GB, n_primes=117)
    covers(lay)
    assert all(lay.ws(c) + sum(lay.items(c).values()) <= 0.5 * GB or len(c) == 1 for c in lay.chunks)
    monkeypatch.setenv("SDPB_GPU_STREAM", "sometimes")
    with pytest.raises(ValueError):
        Layout(SHAPES, 3000, 24, budget=250 * GB, n_primes=117)


def test_mixed_shapes_never_share_a_chunk(monkeypatch):
    shapes = [(1, 5, 3)] * 4 + [(2, 5, 3)] * 3 + [(1, 5, 3)] * 2
    lay = plan(monkeypatch, "all", 1e12, shapes=shapes, N=10)
    covers(lay, shapes)
    assert [len(c) for c in lay.chunks] == [4, 3, 2]


def test_store_tree_utilities():
    """tmap / tleaves / nbytes walk MPF slots, lists, tuples, dicts and plain objects (CPU tensors)"""
    import torch

    from sdpvibe_amd.mpf import MPF
    from sdpvibe_amd.store import nbytes, tleaves, tmap

    class Holder:
        def __init__(self):
            self.a = torch.zeros(3, dtype=torch.int64)
            self.name = "x"
    A = MPF(torch.zeros((2, 4, 5), dtype=torch.int32), torch.zeros((4, 5), dtype=torch.int64))
    tree = {"A": A, "l": [torch.ones(7, dtype=torch.int16), (None, 3, Holder())]}
    leaves = tleaves(tree)
    assert [tuple(t.shape) for t in leaves] == [(2, 4, 5), (4, 5), (7,), (3,)]
    assert nbytes(tree) == 2 * 4 * 5 * 4 + 4 * 5 * 8 + 7 * 2 + 3 * 8
    doubled = tmap(lambda t: t + 1, tree)
    assert isinstance(doubled["A"], MPF) and int(doubled["A"].m.sum()) == 40
    assert doubled["l"][1][0] is None and doubled["l"][1][1] == 3 and doubled["l"][1][2].name == "x"
    assert int(doubled["l"][1][2].a.sum()) == 3 and int(tree["l"][1][2].a.sum()) == 0
