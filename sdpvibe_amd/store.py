"""Placement of the large per-chunk arrays of an SDP: resident in HBM, or kept in page-locked host
memory and copied in when a phase needs them (the counterpart of sdpvibe/store.py).

Objects are trees (MPF, lists, tuples, dicts, objects with __slots__ or __dict__) whose leaves are
device tensors.
  put(key, obj)    store: a resident key keeps the reference; a host key is copied out
                   asynchronously on the device-to-host stream
  prefetch(key)    start the host-to-device copy on its own stream
  get(key)         the device object (waits for its copy on the calling stream)
  stream(items)    iterate over device objects, prefetching `ahead` items in advance
Each host key owns one page-locked buffer of exactly its size (sdpv_host_alloc), carved into
the leaves; shapes do not change between iterations, so it is allocated once. Copies run on two
side streams of default priority. Per-key events order the two directions against each other and
against the compute stream, and record_stream keeps the caching allocator from reusing memory
that a copy still reads or writes.
"""
import ctypes

import torch

from ._lib import HipLibraryError, call, lib

ALIGN = 256


def tmap(f, obj):
    """obj with f applied to every tensor leaf (a new tree of the same types)"""
    if torch.is_tensor(obj):
        return f(obj)
    if obj is None or isinstance(obj, (int, float, bool, str)):
        return obj
    if isinstance(obj, (list, tuple)):
        return type(obj)(tmap(f, o) for o in obj)
    if isinstance(obj, dict):
        return {k: tmap(f, v) for k, v in obj.items()}
    new = object.__new__(type(obj))
    names = list(getattr(obj, "__dict__", {}).keys())
    names += [s for c in type(obj).__mro__ for s in getattr(c, "__slots__", ())]
    for k in names:
        if hasattr(obj, k):
            setattr(new, k, tmap(f, getattr(obj, k)))
    return new


def tleaves(obj):
    """the tensor leaves of obj, in tmap order"""
    out = []
    tmap(lambda t: out.append(t) or t, obj)
    return out


def nbytes(obj):
    return sum(t.numel() * t.element_size() for t in tleaves(obj))


class HostBuffer:
    """page-locked host memory of exactly n bytes, viewed as a uint8 CPU tensor"""

    def __init__(self, n):
        self.n = max(1, int(n))
        self.ptr = lib().sdpv_host_alloc(self.n)
        if not self.ptr:
            raise HipLibraryError(f"sdpv_host_alloc: no page-locked host memory for {self.n} bytes")
        self.t = torch.frombuffer((ctypes.c_uint8 * self.n).from_address(self.ptr), dtype=torch.uint8)

    def free(self):
        if self.ptr:
            self.t = None
            call("sdpv_host_free", ctypes.c_void_p(self.ptr))
            self.ptr = None


class Store:
    """keys in host_keys live in host memory between uses; every other key stays on the device"""

    def __init__(self, host_keys=(), ahead=1, device=None):
        self.host_keys = set(host_keys)
        self.ahead = ahead
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.dev = {}                        # resident keys: the device object
        self.host = {}                       # host keys: a skeleton of host views
        self.buffers = {}                    # host keys: their HostBuffer
        self.ev_d2h, self.ev_h2d = {}, {}    # last copy event of each direction, per key
        self.pending = {}                    # prefetched: key -> (device object, ready event)
        self.h2d = torch.cuda.Stream(device=self.device)
        self.d2h = torch.cuda.Stream(device=self.device)
        self.bytes_h2d = self.bytes_d2h = 0

    @property
    def host_bytes(self):
        return sum(b.n for b in self.buffers.values())

    def is_host(self, key):
        return key in self.host_keys

    def has(self, key):
        return key in self.dev or key in self.host

    def _alloc_host(self, key, obj):
        sizes = [(t.numel() * t.element_size() + ALIGN - 1) // ALIGN * ALIGN for t in tleaves(obj)]
        buf = HostBuffer(sum(sizes))
        self.buffers[key] = buf
        off = [0]

        def view(t):
            n = t.numel() * t.element_size()
            v = buf.t[off[0]:off[0] + n].view(t.dtype).view(t.shape)
            off[0] += (n + ALIGN - 1) // ALIGN * ALIGN
            return v
        return tmap(view, obj)

    def put(self, key, obj):
        if key not in self.host_keys:
            self.dev[key] = obj
            return
        self.pending.pop(key, None)                      # a prefetched copy is stale now
        if key not in self.host:
            self.host[key] = self._alloc_host(key, obj)
        elif [tuple(h.shape) for h in tleaves(self.host[key])] != [tuple(t.shape) for t in tleaves(obj)]:
            raise ValueError(f"Store.put: {key} changed shape")
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.d2h):
            self.d2h.wait_event(ready)
            if key in self.ev_h2d:                       # the host buffer may still be read
                self.d2h.wait_event(self.ev_h2d[key])
            for h, t in zip(tleaves(self.host[key]), tleaves(obj)):
                t = t.contiguous()
                t.record_stream(self.d2h)
                h.copy_(t, non_blocking=True)
                self.bytes_d2h += h.numel() * h.element_size()
            done = torch.cuda.Event()
            done.record(self.d2h)
        self.ev_d2h[key] = done

    def prefetch(self, key):
        if key not in self.host_keys or key in self.pending or key not in self.host:
            return
        with torch.cuda.stream(self.h2d):
            if key in self.ev_d2h:
                self.h2d.wait_event(self.ev_d2h[key])

            def up(h):
                self.bytes_h2d += h.numel() * h.element_size()
                return torch.empty(h.shape, dtype=h.dtype, device=self.device).copy_(h, non_blocking=True)
            d = tmap(up, self.host[key])
            ready = torch.cuda.Event()
            ready.record(self.h2d)
        self.ev_h2d[key] = ready
        self.pending[key] = (d, ready)

    def get(self, key):
        if key not in self.host_keys:
            return self.dev[key]
        if key not in self.host:
            raise KeyError(key)
        self.prefetch(key)
        d, ready = self.pending.pop(key)
        cur = torch.cuda.current_stream(self.device)
        cur.wait_event(ready)
        for t in tleaves(d):
            t.record_stream(cur)
        return d

    def stream(self, items, nxt=()):
        """items: a sequence of key tuples -> yields tuples of device objects, prefetching the next
        `ahead` items; nxt: keys the following pass needs first, prefetched during the last item"""
        items = [tuple(it) for it in items]
        for i, it in enumerate(items):
            for j in range(i, min(len(items), i + 1 + self.ahead)):
                for k in items[j]:
                    self.prefetch(k)
            if i == len(items) - 1:
                for k in nxt:
                    self.prefetch(k)
            yield tuple(self.get(k) for k in it)

    def drop(self, key):
        self.dev.pop(key, None)
        self.pending.pop(key, None)

    def close(self):
        """wait for every copy, then release the host buffers"""
        self.d2h.synchronize()
        self.h2d.synchronize()
        self.pending.clear()
        self.host.clear()
        for b in self.buffers.values():
            b.free()
        self.buffers.clear()

    def __del__(self):
        try:
            if self.buffers:
                self.close()
        except Exception:
            pass
