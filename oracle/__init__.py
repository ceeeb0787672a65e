"""Test infrastructure only: ctypes face of the CPU oracle
(oracle/_build/libfn2_oracle.so, built by ``make -C oracle``).

The oracle restates FastNoise2's generator path on the CPU (see
fn2_oracle.h for its modes and the reference files it follows). Only the
tests, the smoke check and the benchmark's host-baseline leg load it; the
product library and package never do.

Parity status of the oracle's parts (what pins each one) is listed in the
test files that check them: tests/test_oracle_closed_forms.py and
tests/test_oracle_golden.py. Parts whose semantics live in FastSIMD and that
no reference test, fixture or closed form fixes are "parity unpinned":
FS::Sin/FS::Cos (GenTileable2D, SineWave), FS::Pow (PowFloat, Minkowski
distance), FS::Modulus (Modulus node).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfn2_oracle.so")

FUSED, UNFUSED, AVX2 = 0, 1, 2

_lib = None
_fp = ctypes.POINTER(ctypes.c_float)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("oracle not built: run `make -C oracle` (expected %s)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, f, i = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
    L.fn2o_new_from_encoded.restype = vp
    L.fn2o_new_from_encoded.argtypes = [ctypes.c_char_p]
    L.fn2o_delete.argtypes = [vp]
    L.fn2o_root_id.restype = i
    L.fn2o_root_id.argtypes = [vp]
    tail = [i, i, i, i, _fp]
    L.fn2o_gen_grid2d.argtypes = [vp, _fp, f, f, i, i, f, f] + tail
    L.fn2o_gen_grid3d.argtypes = [vp, _fp, f, f, f, i, i, i, f, f, f] + tail
    L.fn2o_gen_grid4d.argtypes = [vp, _fp, f, f, f, f, i, i, i, i, f, f, f, f] + tail
    L.fn2o_gen_pos2d.argtypes = [vp, _fp, i, _fp, _fp, f, f] + tail
    L.fn2o_gen_pos3d.argtypes = [vp, _fp, i, _fp, _fp, _fp, f, f, f] + tail
    L.fn2o_gen_pos4d.argtypes = [vp, _fp, i, _fp, _fp, _fp, _fp, f, f, f, f] + tail
    L.fn2o_gen_tileable2d.argtypes = [vp, _fp, i, i, f, f] + tail
    for n in ("grid2d", "grid3d", "grid4d", "pos2d", "pos3d", "pos4d", "tileable2d"):
        getattr(L, "fn2o_gen_" + n).restype = i
    L.fn2o_gen_single2d.restype = f
    L.fn2o_gen_single2d.argtypes = [vp, f, f, i, i]
    L.fn2o_gen_single3d.restype = f
    L.fn2o_gen_single3d.argtypes = [vp, f, f, f, i, i]
    L.fn2o_gen_single4d.restype = f
    L.fn2o_gen_single4d.argtypes = [vp, f, f, f, f, i, i]
    ip = ctypes.POINTER(ctypes.c_int)
    L.fn2o_hash_primes.restype = i
    L.fn2o_hash_primes.argtypes = [i, i, ip]
    L.fn2o_hash_primes_hb.restype = i
    L.fn2o_hash_primes_hb.argtypes = [i, i, ip]
    L.fn2o_value_coord.restype = f
    L.fn2o_value_coord.argtypes = [i, i, ip]
    L.fn2o_grad_dot.restype = f
    L.fn2o_grad_dot.argtypes = [i, i, _fp, i]
    L.fn2o_interp.restype = f
    L.fn2o_interp.argtypes = [i, f, i]
    L.fn2o_warp_vec.restype = None
    L.fn2o_warp_vec.argtypes = [i, i, i, _fp, f, _fp, i]
    L.fn2o_base64_decode.restype = i
    L.fn2o_base64_decode.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_fp)


class Tree:
    """A decoded node tree in the oracle."""

    def __init__(self, encoded):
        if not isinstance(encoded, str):
            encoded = encoded.encode()  # NodeSpec
        self.encoded = encoded
        self.h = load().fn2o_new_from_encoded(encoded.encode())
        if not self.h:
            raise ValueError("oracle cannot decode %r" % encoded)

    def __del__(self):
        if getattr(self, "h", None):
            load().fn2o_delete(self.h)
            self.h = None

    @property
    def root_id(self):
        return load().fn2o_root_id(self.h)

    def grid(self, offset, count, step=None, seed=1337, mode=FUSED, lanes=1, threads=1):
        d = len(count)
        step = [1.0] * d if step is None else list(step)
        n = int(np.prod([max(0, c) for c in count]))
        out = np.empty(max(n, 1), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        L = load()
        fn = {2: L.fn2o_gen_grid2d, 3: L.fn2o_gen_grid3d, 4: L.fn2o_gen_grid4d}[d]
        rc = fn(self.h, _p(out), *[float(o) for o in offset], *[int(c) for c in count], *[float(s) for s in step],
                seed, mode, lanes, threads, _p(mm))
        if rc:
            raise ValueError("oracle grid call rejected")
        return out[:n], (float(mm[0]), float(mm[1]))

    def positions(self, pos, offset=None, seed=1337, mode=FUSED, lanes=1, threads=1):
        pos = [np.ascontiguousarray(p, dtype=np.float32) for p in pos]
        d = len(pos)
        off = [0.0] * d if offset is None else [float(o) for o in offset]
        n = pos[0].size
        out = np.empty(max(n, 1), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        L = load()
        fn = {2: L.fn2o_gen_pos2d, 3: L.fn2o_gen_pos3d, 4: L.fn2o_gen_pos4d}[d]
        rc = fn(self.h, _p(out), n, *[_p(p) for p in pos], *off, seed, mode, lanes, threads, _p(mm))
        if rc:
            raise ValueError("oracle position call rejected")
        return out[:n], (float(mm[0]), float(mm[1]))

    def tileable(self, x_size, y_size, x_step=1.0, y_step=1.0, seed=1337, mode=FUSED, lanes=1, threads=1):
        n = max(0, x_size) * max(0, y_size)
        out = np.empty(max(n, 1), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        rc = load().fn2o_gen_tileable2d(self.h, _p(out), x_size, y_size, x_step, y_step, seed, mode, lanes, threads, _p(mm))
        if rc:
            raise ValueError("oracle tileable call rejected")
        return out[:n], (float(mm[0]), float(mm[1]))

    def single(self, *coords, seed=1337, mode=FUSED):
        L = load()
        fn = {2: L.fn2o_gen_single2d, 3: L.fn2o_gen_single3d, 4: L.fn2o_gen_single4d}[len(coords)]
        return fn(self.h, *[float(c) for c in coords], seed, mode)


COUNT_LIB_PATH = os.path.join(HERE, "_build", "libfn2_oracle_count.so")
_count_lib = None


def op_counts_per_point(encoded, dims, offset, count, step=None, seed=1337, tileable=False):
    """Algorithmic work of one tree per point, from the op-counting build of
    the oracle: (FP32 FLOPs with fma = 2, vector ALU operations), averaged over
    the given grid (or tileable image when tileable=True, count = (x, y))."""
    global _count_lib
    if _count_lib is None:
        if not os.path.exists(COUNT_LIB_PATH):
            raise ImportError("counting oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(COUNT_LIB_PATH)
        L.fn2o_new_from_encoded.restype = ctypes.c_void_p
        L.fn2o_new_from_encoded.argtypes = [ctypes.c_char_p]
        L.fn2o_delete.argtypes = [ctypes.c_void_p]
        L.fn2o_reset_counters.argtypes = []
        L.fn2o_counters.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)] * 2
        _count_lib = L
    L = _count_lib
    if not isinstance(encoded, str):
        encoded = encoded.encode()
    h = L.fn2o_new_from_encoded(encoded.encode())
    if not h:
        raise ValueError("counting oracle cannot decode %r" % encoded)
    vp, f, i = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
    tail = [i, i, i, i, _fp]
    n = int(np.prod(count))
    out = np.empty(n, dtype=np.float32)
    mm = np.zeros(2, dtype=np.float32)
    step = [1.0] * len(count) if step is None else step
    L.fn2o_reset_counters()
    if tileable:
        L.fn2o_gen_tileable2d.argtypes = [vp, _fp, i, i, f, f] + tail
        L.fn2o_gen_tileable2d(h, _p(out), count[0], count[1], step[0], step[1], seed, 0, 1, 1, _p(mm))
    else:
        d = len(count)
        fn = {2: L.fn2o_gen_grid2d, 3: L.fn2o_gen_grid3d, 4: L.fn2o_gen_grid4d}[d]
        fn.argtypes = [vp, _fp] + [f] * d + [i] * d + [f] * d + tail
        fn(h, _p(out), *[float(o) for o in offset], *[int(c) for c in count], *[float(s) for s in step], seed, 0, 1, 1, _p(mm))
    flop, valu = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    L.fn2o_counters(ctypes.byref(flop), ctypes.byref(valu))
    L.fn2o_delete(h)
    return flop.value / n, valu.value / n


TRACE_LIB_PATH = os.path.join(HERE, "_build", "libfn2_oracle_trace.so")
_trace_lib = None


def _load_trace():
    global _trace_lib
    if _trace_lib is None:
        if not os.path.exists(TRACE_LIB_PATH):
            raise ImportError("index-replay oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(TRACE_LIB_PATH)
        vp, f, i = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
        L.fn2o_new_from_encoded.restype = vp
        L.fn2o_new_from_encoded.argtypes = [ctypes.c_char_p]
        L.fn2o_delete.argtypes = [vp]
        L.fn2o_set_trace.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        tail = [i, i, i, i, _fp]
        L.fn2o_gen_grid2d.argtypes = [vp, _fp, f, f, i, i, f, f] + tail
        L.fn2o_gen_grid3d.argtypes = [vp, _fp, f, f, f, i, i, i, f, f, f] + tail
        L.fn2o_gen_grid4d.argtypes = [vp, _fp, f, f, f, f, i, i, i, i, f, f, f, f] + tail
        L.fn2o_gen_pos2d.argtypes = [vp, _fp, i, _fp, _fp, f, f] + tail
        L.fn2o_gen_pos3d.argtypes = [vp, _fp, i, _fp, _fp, _fp, f, f, f] + tail
        L.fn2o_gen_pos4d.argtypes = [vp, _fp, i, _fp, _fp, _fp, _fp, f, f, f, f] + tail
        L.fn2o_gen_tileable2d.argtypes = [vp, _fp, i, i, f, f] + tail
        _trace_lib = L
    return _trace_lib


def index_trace(encoded, kind, mode=FUSED, seed=1337, offset=None, count=None, step=None, pos=None):
    """Values and per-point decision fingerprints of one call in the oracle's
    index-replay build: kind "grid" (offset, count, step), "positions" (pos
    arrays, offset) or "tileable" (count = (x, y), step). Two variants of the
    oracle agree on a point's fingerprint exactly when every float -> int
    conversion, float comparison, floor / round result and hashed position
    bit pattern on its path agrees: the same lattice cells, simplex vertices,
    cells and selections."""
    L = _load_trace()
    if not isinstance(encoded, str):
        encoded = encoded.encode()
    h = L.fn2o_new_from_encoded(encoded.encode())
    if not h:
        raise ValueError("index-replay oracle cannot decode %r" % encoded)
    try:
        if kind == "positions":
            pos = [np.ascontiguousarray(p, dtype=np.float32) for p in pos]
            n, d = pos[0].size, len(pos)
        else:
            d = len(count) if kind == "grid" else 2
            n = int(np.prod(count))
        out = np.empty(max(n, 1), np.float32)
        fp = np.zeros(max(n, 1), np.uint64)
        mm = np.zeros(2, np.float32)
        L.fn2o_set_trace(fp.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)))
        if kind == "grid":
            step = [1.0] * d if step is None else step
            fn = {2: L.fn2o_gen_grid2d, 3: L.fn2o_gen_grid3d, 4: L.fn2o_gen_grid4d}[d]
            rc = fn(h, _p(out), *[float(o) for o in offset], *[int(c) for c in count], *[float(s) for s in step],
                    seed, mode, 1, 1, _p(mm))
        elif kind == "positions":
            off = [0.0] * d if offset is None else offset
            fn = {2: L.fn2o_gen_pos2d, 3: L.fn2o_gen_pos3d, 4: L.fn2o_gen_pos4d}[d]
            rc = fn(h, _p(out), n, *[_p(p) for p in pos], *[float(o) for o in off], seed, mode, 1, 1, _p(mm))
        else:
            step = [1.0, 1.0] if step is None else step
            rc = L.fn2o_gen_tileable2d(h, _p(out), count[0], count[1], float(step[0]), float(step[1]), seed, mode, 1, 1, _p(mm))
        L.fn2o_set_trace(None)
        if rc:
            raise ValueError("index-replay oracle rejected the call")
        return out[:n], fp[:n]
    finally:
        L.fn2o_delete(h)
