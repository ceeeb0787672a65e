"""Chunked-world client (fn2World* of fn2_gpu.h): streams (size + 2)^3
density blocks around a moving camera through a serving engine, nearest
chunks first, and consumes them on the GPU (range culling against an iso
surface, exposed-face counts). Mirrors the chunk pipeline of the reference's
node editor (tools/NodeEditor/MeshNoisePreview.cpp:260-470)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import last_error

CHUNK_AIR, CHUNK_SOLID, CHUNK_SURFACE = 0, 1, 2


class ChunkResult(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32), ("state", ctypes.c_int32),
                ("min", ctypes.c_float), ("max", ctypes.c_float), ("quads", ctypes.c_uint32),
                ("solid_voxels", ctypes.c_uint32), ("latency_ms", ctypes.c_double), ("version", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def chunks_in_range(camera, chunk_size, load_range):
    """Chunk origins within load_range of the camera, nearest first (n x 3 int32)."""
    lib = _lib.load()
    cam = (ctypes.c_float * 3)(*camera)
    n = lib.fn2WorldChunksInRange(cam, chunk_size, load_range, None, 0)
    if n < 0:
        raise ValueError("fn2WorldChunksInRange failed: " + last_error())
    out = np.zeros((max(n, 1), 3), dtype=np.int32)
    lib.fn2WorldChunksInRange(cam, chunk_size, load_range, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n)
    return out[:n]


class ChunkWorld:
    def __init__(self, engine, tree, chunk_size=128, iso_surface=0.0, load_range=512.0, seed=1337, max_in_flight=32):
        self.lib = _lib.load()
        self.engine = engine
        self.h = self.lib.fn2WorldCreate(engine.h, tree, chunk_size, iso_surface, load_range, seed, max_in_flight)
        if not self.h:
            raise RuntimeError("fn2WorldCreate failed: " + last_error())
        self.chunk_size = chunk_size
        self._buf = (ChunkResult * 256)()
        engine._clients.add(self)   # the engine closes its worlds before itself

    def close(self):
        if self.h:
            self.lib.fn2WorldDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, camera):
        """Requests the missing chunks in range (nearest first); returns how many."""
        n = self.lib.fn2WorldUpdate(self.h, *[float(c) for c in camera])
        if n < 0:
            raise RuntimeError("fn2WorldUpdate failed: " + last_error())
        return n

    def poll(self, wait_all=False):
        """Finished chunks (oldest first) as a list of ChunkResult dicts."""
        out = []
        while True:
            n = self.lib.fn2WorldPoll(self.h, self._buf, len(self._buf), 1 if wait_all else 0)
            if n < 0:
                raise RuntimeError("fn2WorldPoll failed: " + last_error())
            out += [self._buf[i].as_dict() for i in range(n)]
            if n < len(self._buf):
                return out

    def new_version(self):
        return self.lib.fn2WorldNewVersion(self.h)

    def stats(self):
        r, c, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        loaded, inflight = ctypes.c_int(), ctypes.c_int()
        self.lib.fn2WorldStats(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(d), ctypes.byref(loaded),
                               ctypes.byref(inflight))
        return {"requested": r.value, "completed": c.value, "dropped": d.value, "loaded": loaded.value,
                "in_flight": inflight.value}
