"""FastNoise2's generator path on AMD Instinct MI355X (gfx950).

Python face of libfastnoise2_amd.so, mirroring FastNoise2's node API
(include/FastNoise/FastNoise.h, Generator.h and FastNoise_C.h of the
reference): nodes are created by name or from encoded node-tree strings,
configured through their metadata members, and generate into numpy arrays
(host) or device buffers. All generation runs on the GPU through the C ABI;
there is no CPU implementation behind this module, and a missing library or
device raises.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import GenSpec, SOURCE_GRID, SOURCE_POSITIONS, SOURCE_TILEABLE, last_error
from .tree import NodeSpec, meta, registry, b64encode, b64decode

__all__ = ["Node", "NodeSpec", "DeviceArray", "PinnedArray", "GenSpec", "lib", "last_error", "meta", "registry",
           "b64encode", "b64decode", "FEATURE_SET_GPU", "device_count"]

FEATURE_SET_GPU = 0x1000


def lib():
    return _lib.load()


def device_count():
    return lib().fn2DeviceCount()


JIT_OFF, JIT_ASYNC, JIT_SYNC = 0, 1, 2
# variant of a specialised kernel that position arrays of 3D CellularDistance trees run
JIT_VARIANT_SHARED_CELL_WALK = 4


def set_jit_mode(mode):
    """Run-time specialisation of trees: JIT_OFF, JIT_ASYNC (default) or JIT_SYNC.
    Returns the previous mode."""
    return lib().fn2SetJitMode(mode)


def jit_wait():
    """Waits for every queued specialisation to finish compiling."""
    lib().fn2JitWait()


def _fp(a):
    return a.ctypes.data_as(_lib.c_float_p)


def _check(what):
    err = last_error()
    if err:
        raise RuntimeError("%s failed: %s" % (what, err))


class DeviceArray:
    """A float32 buffer in device memory (fn2Malloc)."""

    def __init__(self, count):
        self.count = int(count)
        self.ptr = lib().fn2Malloc(max(1, self.count) * 4)
        if not self.ptr:
            raise RuntimeError("fn2Malloc failed: " + last_error())

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        d = cls(arr.size)
        if arr.size and lib().fn2Memcpy(d.ptr, arr.ctypes.data, arr.size * 4, 0):
            raise RuntimeError("copy to device failed: " + last_error())
        return d

    def to_host(self):
        out = np.empty(self.count, dtype=np.float32)
        if self.count and lib().fn2Memcpy(out.ctypes.data, self.ptr, self.count * 4, 1):
            raise RuntimeError("copy to host failed: " + last_error())
        return out

    def free(self):
        if self.ptr:
            lib().fn2Free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A float32 numpy view of pinned host memory (fn2MallocHost), which the
    device fills by DMA with no staging copy."""

    def __init__(self, count):
        self.count = int(count)
        self.ptr = lib().fn2MallocHost(max(1, self.count) * 4)
        if not self.ptr:
            raise RuntimeError("fn2MallocHost failed: " + last_error())
        buf = (ctypes.c_float * max(1, self.count)).from_address(self.ptr)
        self.array = np.ctypeslib.as_array(buf)[:self.count]

    def free(self):
        if self.ptr:
            self.array = None
            lib().fn2FreeHost(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Node:
    """A node handle of the C ABI (one reference, released on deletion)."""

    def __init__(self, name=None, feature_set=0, **members):
        if name is None:
            raise TypeError("Node needs a node name")
        m = meta(name)
        h = lib().fnNewFromMetadata(m.id, feature_set)
        if not h:
            raise RuntimeError("fnNewFromMetadata(%s) returned null" % m.name)
        self.handle = h
        for k, v in members.items():
            self.set(k, v)

    @classmethod
    def _adopt(cls, handle):
        if not handle:
            raise RuntimeError("null node handle")
        n = cls.__new__(cls)
        n.handle = handle
        return n

    @classmethod
    def from_encoded(cls, encoded, feature_set=0):
        """fnNewFromEncodedNodeTree; raises on a string that does not decode."""
        if isinstance(encoded, NodeSpec):
            encoded = encoded.encode()
        h = lib().fnNewFromEncodedNodeTree(encoded.encode() if isinstance(encoded, str) else encoded, feature_set)
        if not h:
            raise ValueError("encoded node tree does not decode: %r" % (encoded,))
        return cls._adopt(h)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().fnDeleteNodeRef(h)
            except Exception:
                pass
            self.handle = None

    @property
    def metadata_id(self):
        return lib().fnGetMetadataID(self.handle)

    @property
    def name(self):
        return meta(self.metadata_id).name

    @property
    def active_feature_set(self):
        return lib().fnGetActiveFeatureSet(self.handle)

    def set(self, key, value):
        m = meta(self.metadata_id)
        kind, idx = m.member(key)
        L = lib()
        if kind == "var":
            v = m.variables[idx]
            if v.type == 2 and isinstance(value, str):
                value = [e.lower().replace(" ", "") for e in v.enum_names].index(value.lower().replace(" ", "").replace("_", ""))
            ok = L.fnSetVariableFloat(self.handle, idx, float(value)) if v.type == 0 else \
                L.fnSetVariableIntEnum(self.handle, idx, int(value))
        elif kind == "lookup":
            ok = L.fnSetNodeLookup(self.handle, idx, value.handle)
        elif isinstance(value, Node):
            ok = L.fnSetHybridNodeLookup(self.handle, idx, value.handle)
        else:
            ok = L.fnSetHybridFloat(self.handle, idx, float(value))
        if not ok:
            raise ValueError("cannot set %s.%s" % (m.name, key))
        return self

    # --- generation into host arrays (FastNoise_C.h entry points) -----------
    def gen_uniform_grid_2d(self, x_offset, y_offset, x_count, y_count, x_step=1.0, y_step=1.0, seed=1337):
        out = np.empty(max(0, x_count) * max(0, y_count), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        lib().fnGenUniformGrid2D(self.handle, _fp(out), x_offset, y_offset, x_count, y_count, x_step, y_step, seed, _fp(mm))
        _check("fnGenUniformGrid2D")
        return out, (float(mm[0]), float(mm[1]))

    def gen_uniform_grid_3d(self, x_offset, y_offset, z_offset, x_count, y_count, z_count,
                            x_step=1.0, y_step=1.0, z_step=1.0, seed=1337):
        out = np.empty(max(0, x_count) * max(0, y_count) * max(0, z_count), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        lib().fnGenUniformGrid3D(self.handle, _fp(out), x_offset, y_offset, z_offset, x_count, y_count, z_count,
                                 x_step, y_step, z_step, seed, _fp(mm))
        _check("fnGenUniformGrid3D")
        return out, (float(mm[0]), float(mm[1]))

    def gen_uniform_grid_4d(self, x_offset, y_offset, z_offset, w_offset, x_count, y_count, z_count, w_count,
                            x_step=1.0, y_step=1.0, z_step=1.0, w_step=1.0, seed=1337):
        n = max(0, x_count) * max(0, y_count) * max(0, z_count) * max(0, w_count)
        out = np.empty(n, dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        lib().fnGenUniformGrid4D(self.handle, _fp(out), x_offset, y_offset, z_offset, w_offset,
                                 x_count, y_count, z_count, w_count, x_step, y_step, z_step, w_step, seed, _fp(mm))
        _check("fnGenUniformGrid4D")
        return out, (float(mm[0]), float(mm[1]))

    def gen_position_array(self, positions, offset=None, seed=1337):
        """positions: sequence of 2, 3 or 4 float32 arrays (x, y[, z[, w]])."""
        pos = [np.ascontiguousarray(p, dtype=np.float32) for p in positions]
        d = len(pos)
        n = pos[0].size
        off = list(offset) if offset is not None else [0.0] * d
        out = np.empty(n, dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        L = lib()
        if d == 2:
            L.fnGenPositionArray2D(self.handle, _fp(out), n, _fp(pos[0]), _fp(pos[1]), off[0], off[1], seed, _fp(mm))
        elif d == 3:
            L.fnGenPositionArray3D(self.handle, _fp(out), n, _fp(pos[0]), _fp(pos[1]), _fp(pos[2]),
                                   off[0], off[1], off[2], seed, _fp(mm))
        elif d == 4:
            L.fnGenPositionArray4D(self.handle, _fp(out), n, _fp(pos[0]), _fp(pos[1]), _fp(pos[2]), _fp(pos[3]),
                                   off[0], off[1], off[2], off[3], seed, _fp(mm))
        else:
            raise ValueError("2, 3 or 4 position arrays")
        _check("fnGenPositionArray%dD" % d)
        return out, (float(mm[0]), float(mm[1]))

    def gen_tileable_2d(self, x_size, y_size, x_step=1.0, y_step=1.0, seed=1337):
        out = np.empty(max(0, x_size) * max(0, y_size), dtype=np.float32)
        mm = np.zeros(2, dtype=np.float32)
        lib().fnGenTileable2D(self.handle, _fp(out), x_size, y_size, x_step, y_step, seed, _fp(mm))
        _check("fnGenTileable2D")
        return out, (float(mm[0]), float(mm[1]))

    def gen_single(self, *coords, seed=1337):
        L = lib()
        fn = {2: L.fnGenSingle2D, 3: L.fnGenSingle3D, 4: L.fnGenSingle4D}[len(coords)]
        v = fn(self.handle, *[float(c) for c in coords], seed)
        _check("fnGenSingle%dD" % len(coords))
        return v

    # --- device entry points (fn2_gpu.h) --------------------------------------
    def generate_device(self, spec, out, minmax=True, stream=None):
        """Generates into a DeviceArray (or raw device pointer) per a GenSpec."""
        mm = np.zeros(2, dtype=np.float32)
        ptr = out.ptr if isinstance(out, DeviceArray) else out
        rc = lib().fn2GenerateDevice(self.handle, ctypes.byref(spec), ptr, _fp(mm) if minmax else None, stream)
        if rc:
            raise RuntimeError("fn2GenerateDevice failed: " + last_error())
        return (float(mm[0]), float(mm[1])) if minmax else None

    def bench_device(self, spec, out, warmup=3, iters=10):
        """(event ms per launch, wall ms per launch) of back-to-back launches."""
        ms = ctypes.c_double()
        wall = ctypes.c_double()
        ptr = out.ptr if isinstance(out, DeviceArray) else out
        rc = lib().fn2BenchGenerate(self.handle, ctypes.byref(spec), ptr, warmup, iters, ctypes.byref(ms), ctypes.byref(wall))
        if rc:
            raise RuntimeError("fn2BenchGenerate failed: " + last_error())
        return ms.value, wall.value

    def jit_state(self, dims, variant=0):
        """-1 never requested, 0 queued, 1 ready, 2 failed, 3 kept on the interpreter (too long).
        variant: JIT_VARIANT_SHARED_CELL_WALK for the kernel position arrays of
        3D CellularDistance trees run, 1..3 for the octave-unroll variants."""
        return lib().fn2GetJitVariantState(self.handle, dims, variant)

    def set_jit_variant(self, dims, variant):
        """Octave-unroll variant (0 the compiler's choice, 1 none, 2 by two, 3 by
        four) that single calls run in `dims` dimensions; same results."""
        if lib().fn2SetJitVariant(self.handle, dims, variant):
            raise ValueError("fn2SetJitVariant failed: " + last_error())

    def jit_kernel_info(self, dims, variant=0):
        """(VGPRs, scratch bytes per lane, resident blocks per CU) of the loaded
        specialised kernel, or None when it is not loaded."""
        v, sc, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = lib().fn2GetJitKernelInfo(self.handle, dims, variant, ctypes.byref(v), ctypes.byref(sc), ctypes.byref(b))
        if rc < 0:
            raise ValueError("fn2GetJitKernelInfo failed: " + last_error())
        return None if rc else (v.value, sc.value, b.value)

    def jit_source(self, dims):
        """HIP source of the tree's specialised kernel."""
        n = lib().fn2JitSource(self.handle, dims, None, 0)
        if n < 0:
            raise RuntimeError("fn2JitSource failed: " + last_error())
        buf = ctypes.create_string_buffer(n + 1)
        lib().fn2JitSource(self.handle, dims, buf, n + 1)
        return buf.value.decode()

    def program_info(self, dims):
        ins, regs, dm, fm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        if lib().fn2GetProgramInfo(self.handle, dims, ctypes.byref(ins), ctypes.byref(regs), ctypes.byref(dm), ctypes.byref(fm)):
            raise RuntimeError("fn2GetProgramInfo failed: " + last_error())
        return {"instructions": ins.value, "registers": regs.value, "dim_mask": dm.value, "feature_mask": fm.value}


def grid_spec(dims, offset, count, step=None, seed=1337):
    s = GenSpec()
    s.kind = SOURCE_GRID
    s.dims = dims
    for a in range(dims):
        s.count[a] = count[a]
        s.offset[a] = offset[a]
        s.step[a] = 1.0 if step is None else step[a]
    s.seed = seed
    return s


def tileable_spec(x_size, y_size, x_step=1.0, y_step=1.0, seed=1337):
    s = GenSpec()
    s.kind = SOURCE_TILEABLE
    s.dims = 2
    s.count[0], s.count[1] = x_size, y_size
    s.step[0], s.step[1] = x_step, y_step
    s.seed = seed
    return s


def positions_spec(device_arrays, offset=None, seed=1337):
    s = GenSpec()
    s.kind = SOURCE_POSITIONS
    s.dims = len(device_arrays)
    s.count[0] = device_arrays[0].count
    for a, d in enumerate(device_arrays):
        s.pos[a] = d.ptr
        s.offset[a] = 0.0 if offset is None else offset[a]
    s.seed = seed
    return s
