"""ctypes binding of libfastnoise2_amd.so: FastNoise2's C ABI
(include/FastNoise/FastNoise_C.h of the reference, 45 symbols) plus the
device extensions declared in include/fn2_gpu.h.

The library is loaded from ``fastnoise2_amd/lib`` (built in-tree by
``make -C fastnoise2_amd``). There is no fallback: if the shared library is
missing, loading raises ``ImportError`` naming the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfastnoise2_amd.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_double_p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p
f = ctypes.c_float
i = ctypes.c_int
u = ctypes.c_uint
s = ctypes.c_char_p
b = ctypes.c_bool


class GenSpec(ctypes.Structure):
    """fn2GenSpec of include/fn2_gpu.h."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("dims", ctypes.c_int),
        ("count", ctypes.c_int * 4),
        ("offset", ctypes.c_float * 4),
        ("step", ctypes.c_float * 4),
        ("pos", ctypes.c_void_p * 4),
        ("seed", ctypes.c_int),
    ]


SOURCE_GRID, SOURCE_POSITIONS, SOURCE_TILEABLE = 0, 1, 2
SERVE_PLAIN, SERVE_BATCHED, SERVE_GRAPH, SERVE_PERSISTENT = 0, 1, 2, 3

# name -> (restype, argtypes); the FastNoise_C.h entry points first.
C_ABI = {
    "fnNewFromEncodedNodeTree": (vp, [s, u]),
    "fnDeleteNodeRef": (None, [vp]),
    "fnGetActiveFeatureSet": (u, [vp]),
    "fnGetMetadataID": (i, [vp]),
    "fnGenUniformGrid2D": (None, [vp, c_float_p, f, f, i, i, f, f, i, c_float_p]),
    "fnGenUniformGrid3D": (None, [vp, c_float_p, f, f, f, i, i, i, f, f, f, i, c_float_p]),
    "fnGenUniformGrid4D": (None, [vp, c_float_p, f, f, f, f, i, i, i, i, f, f, f, f, i, c_float_p]),
    "fnGenPositionArray2D": (None, [vp, c_float_p, i, c_float_p, c_float_p, f, f, i, c_float_p]),
    "fnGenPositionArray3D": (None, [vp, c_float_p, i, c_float_p, c_float_p, c_float_p, f, f, f, i, c_float_p]),
    "fnGenPositionArray4D": (None, [vp, c_float_p, i, c_float_p, c_float_p, c_float_p, c_float_p, f, f, f, f, i, c_float_p]),
    "fnGenTileable2D": (None, [vp, c_float_p, i, i, f, f, i, c_float_p]),
    "fnGenSingle2D": (f, [vp, f, f, i]),
    "fnGenSingle3D": (f, [vp, f, f, f, i]),
    "fnGenSingle4D": (f, [vp, f, f, f, f, i]),
    "fnGetMetadataCount": (i, []),
    "fnGetMetadataName": (s, [i]),
    "fnNewFromMetadata": (vp, [i, u]),
    "fnGetMetadataVariableCount": (i, [i]),
    "fnGetMetadataVariableName": (s, [i, i]),
    "fnGetMetadataVariableType": (i, [i, i]),
    "fnGetMetadataVariableDimensionIdx": (i, [i, i]),
    "fnGetMetadataEnumCount": (i, [i, i]),
    "fnGetMetadataEnumName": (s, [i, i, i]),
    "fnSetVariableFloat": (b, [vp, i, f]),
    "fnSetVariableIntEnum": (b, [vp, i, i]),
    "fnGetMetadataNodeLookupCount": (i, [i]),
    "fnGetMetadataNodeLookupName": (s, [i, i]),
    "fnGetMetadataNodeLookupDimensionIdx": (i, [i, i]),
    "fnSetNodeLookup": (b, [vp, i, vp]),
    "fnGetMetadataHybridCount": (i, [i]),
    "fnGetMetadataHybridName": (s, [i, i]),
    "fnGetMetadataHybridDimensionIdx": (i, [i, i]),
    "fnSetHybridNodeLookup": (b, [vp, i, vp]),
    "fnSetHybridFloat": (b, [vp, i, f]),
    "fnGetMetadataDescription": (s, [i]),
    "fnGetMetadataGroupCount": (i, [i]),
    "fnGetMetadataGroupName": (s, [i, i]),
    "fnGetMetadataVariableDescription": (s, [i, i]),
    "fnGetMetadataVariableDefaultFloat": (f, [i, i]),
    "fnGetMetadataVariableDefaultIntEnum": (i, [i, i]),
    "fnGetMetadataVariableMinFloat": (f, [i, i]),
    "fnGetMetadataVariableMaxFloat": (f, [i, i]),
    "fnGetMetadataNodeLookupDescription": (s, [i, i]),
    "fnGetMetadataHybridDescription": (s, [i, i]),
    "fnGetMetadataHybridDefault": (f, [i, i]),
}

GPU_EXT = {
    "fn2GenerateDevice": (i, [vp, ctypes.POINTER(GenSpec), vp, c_float_p, vp]),
    "fn2GenerateDeviceNoAlloc": (i, [vp, ctypes.POINTER(GenSpec), vp, c_float_p, vp]),
    "fn2NoAllocScratchBytes": (ctypes.c_size_t, []),
    "fn2BenchGenerate": (i, [vp, ctypes.POINTER(GenSpec), vp, i, i, c_double_p, c_double_p]),
    "fn2Malloc": (vp, [ctypes.c_size_t]),
    "fn2Free": (None, [vp]),
    "fn2MallocHost": (vp, [ctypes.c_size_t]),
    "fn2FreeHost": (None, [vp]),
    "fn2Memcpy": (i, [vp, vp, ctypes.c_size_t, i]),
    "fn2DeviceSynchronize": (i, []),
    "fn2DeviceCount": (i, []),
    "fn2GetLastError": (s, []),
    "fn2SelfTestMath": (i, [i, vp, vp, i]),
    "fn2SetJitMode": (i, [i]),
    "fn2JitWait": (i, []),
    "fn2JitCompileCheck": (i, [vp, i, c_int_p]),
    "fn2JitSource": (i, [vp, i, ctypes.c_char_p, i]),
    "fn2GetJitState": (i, [vp, i]),
    "fn2GetJitVariantState": (i, [vp, i, i]),
    "fn2SetJitVariant": (i, [vp, i, i]),
    "fn2GetJitKernelInfo": (i, [vp, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
    "fn2GetProgramInfo": (i, [vp, i, c_int_p, c_int_p, c_int_p, c_int_p]),
    "fn2ServeCreate": (vp, [ctypes.c_size_t, i, i]),
    "fn2ServeDestroy": (None, [vp]),
    "fn2ServeAddTree": (i, [vp, s]),
    "fn2ServeSubmit": (ctypes.c_int64, [vp, i, ctypes.POINTER(GenSpec), c_float_p]),
    "fn2ServeWait": (i, [vp, ctypes.c_int64, c_float_p, c_double_p]),
    "fn2ServeDrain": (i, [vp]),
    "fn2ServeStats": (i, [vp, c_int64_p, c_int64_p, c_int64_p]),
    "fn2ServeTreeTuning": (i, [vp, i, i]),
    "fn2ServeTuneOctaveUnroll": (i, [vp, i]),
    "fn2ServeQueueStats": (i, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "fn2ServeTreeTile": (i, [vp, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
    "fn2ServePoll": (i, [vp, ctypes.c_int64, c_float_p, c_double_p]),
    "fn2WorldChunksInRange": (i, [c_float_p, i, ctypes.c_float, ctypes.POINTER(ctypes.c_int32), i]),
    "fn2WorldCreate": (vp, [vp, i, i, ctypes.c_float, ctypes.c_float, i, i]),
    "fn2WorldDestroy": (None, [vp]),
    "fn2WorldUpdate": (i, [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "fn2WorldPoll": (i, [vp, vp, i, i]),
    "fn2WorldNewVersion": (ctypes.c_uint32, [vp]),
    "fn2WorldStats": (i, [vp, c_int64_p, c_int64_p, c_int64_p, c_int_p, c_int_p]),
    "fn2ServeReplay": (i, [vp, i, c_int_p, ctypes.POINTER(GenSpec), c_double_p, i, i, c_double_p, c_double_p]),
}

_lib = None


def load():
    """Loads the backend library once; raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libfastnoise2_amd.so is not built (expected at %s); run "
            "`make -C fastnoise2_amd` or __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for table in (C_ABI, GPU_EXT):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    msg = load().fn2GetLastError()
    return msg.decode() if msg else ""
