"""Python face of the serving engine (fn2Serve* in include/fn2_gpu.h):
continuous batching of chunk requests against resident node trees."""
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import GenSpec, SERVE_PLAIN, SERVE_BATCHED, SERVE_GRAPH, SERVE_PERSISTENT, last_error

__all__ = ["ServeEngine", "SERVE_PLAIN", "SERVE_BATCHED", "SERVE_GRAPH", "SERVE_PERSISTENT"]


class ServeEngine:
    def __init__(self, hbm_budget_bytes=16 << 30, max_batch_points=1 << 23, mode=SERVE_BATCHED):
        self.lib = _lib.load()
        self.h = self.lib.fn2ServeCreate(hbm_budget_bytes, max_batch_points, mode)
        if not self.h:
            raise RuntimeError("fn2ServeCreate failed: " + last_error())
        self.mode = mode
        self._outs = {}
        self._clients = weakref.WeakSet()   # worlds built on this engine: closed first

    def close(self):
        if self.h:
            for c in list(self._clients):
                c.close()
            self.lib.fn2ServeDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_tree(self, encoded):
        if not isinstance(encoded, str):
            encoded = encoded.encode()
        t = self.lib.fn2ServeAddTree(self.h, encoded.encode())
        if t < 0:
            raise ValueError("fn2ServeAddTree failed: " + last_error())
        return t

    def submit(self, tree, spec, out=None):
        """Queues one request; `out` receives the values: a float32 numpy array
        of the spec's point count (default: a new one), or a DeviceArray /
        PinnedArray of at least that many floats."""
        n = spec_points(spec)
        if out is None:
            out = np.empty(n, dtype=np.float32)
        if isinstance(out, np.ndarray):
            assert out.dtype == np.float32 and out.size >= n and out.flags.c_contiguous
            ptr = out.ctypes.data_as(_lib.c_float_p)
        else:
            assert out.count >= n
            ptr = ctypes.cast(out.ptr, _lib.c_float_p)
        rid = self.lib.fn2ServeSubmit(self.h, tree, ctypes.byref(spec), ptr)
        if rid < 0:
            raise RuntimeError("fn2ServeSubmit failed: " + last_error())
        self._outs[rid] = out
        return rid

    def wait(self, rid):
        """(values, (min, max), latency in ms) of a submitted request."""
        mm = np.zeros(2, dtype=np.float32)
        lat = ctypes.c_double()
        rc = self.lib.fn2ServeWait(self.h, rid, mm.ctypes.data_as(_lib.c_float_p), ctypes.byref(lat))
        out = self._outs.pop(rid)
        if rc:
            raise RuntimeError("request %d failed: %s" % (rid, last_error()))
        return out, (float(mm[0]), float(mm[1])), lat.value

    def drain(self):
        self.lib.fn2ServeDrain(self.h)

    def replay(self, tree_ids, specs, arrivals=None, max_outstanding=64, device_outputs=False):
        """Native trace replay (fn2ServeReplay): returns (latencies in ms, wall seconds).
        Outputs go to pinned host buffers, or to device buffers with device_outputs."""
        n = len(specs)
        arr_specs = (GenSpec * n)(*specs)
        tids = (ctypes.c_int * n)(*tree_ids)
        lat = np.zeros(n, dtype=np.float64)
        wall = ctypes.c_double()
        arr = None
        if arrivals is not None:
            arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        rc = self.lib.fn2ServeReplay(self.h, n, tids, arr_specs,
                                     arr.ctypes.data_as(_lib.c_double_p) if arr is not None else None,
                                     max_outstanding, 1 if device_outputs else 0,
                                     lat.ctypes.data_as(_lib.c_double_p), ctypes.byref(wall))
        if rc:
            raise RuntimeError("fn2ServeReplay failed: " + last_error())
        return lat, wall.value

    def tree_tuning(self, tree, dims):
        """Points per lane chosen for the tree's batched launches (0 = not tuned yet)."""
        return self.lib.fn2ServeTreeTuning(self.h, tree, dims)

    def tune_octave_unroll(self, enable=True):
        """Lets the tuner time the octave-unroll variants of trees with fused
        fractals (off by default)."""
        self.lib.fn2ServeTuneOctaveUnroll(self.h, 1 if enable else 0)

    def queue_stats(self):
        """Persistent mode: (work-queue launches, launches whose waves pulled
        items beyond their first, pulls made)."""
        v = [ctypes.c_int64() for _ in range(3)]
        self.lib.fn2ServeQueueStats(self.h, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def tree_tile(self, tree, dims):
        """(threads per block, points per lane, octave unroll) the tuner chose
        for the tree in `dims` dimensions, or None when not tuned yet."""
        t, p, u = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = self.lib.fn2ServeTreeTile(self.h, tree, dims, ctypes.byref(t), ctypes.byref(p), ctypes.byref(u))
        if rc < 0:
            raise ValueError("bad tree or dimensionality")
        return None if rc else (t.value, p.value, u.value)

    def stats(self):
        b, c, p = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.lib.fn2ServeStats(self.h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(p))
        return {"batches": b.value, "completed": c.value, "device_bytes": p.value}


def spec_points(spec):
    if spec.kind == _lib.SOURCE_TILEABLE:
        return spec.count[0] * spec.count[1]
    n = 1
    for a in range(spec.dims):
        n *= spec.count[a]
    return n
