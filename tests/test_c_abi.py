"""The drop-in boundary on the CPU: the shared library loads, exports every
entry point its headers declare, reports the reference's node ids and member
orders, follows the C ABI's return conventions, decodes and encodes the
wire format, and refuses to generate without a GPU (no silent fallback)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import _lib
from fastnoise2_amd.tree import NodeSpec as S, b64encode, b64decode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Registration order of the reference (src/FastNoise/FastSIMD_Build.inl:86-143)
NODE_ORDER = [
    "Constant", "White", "Checkerboard", "SineWave", "Gradient", "DistanceToPoint",
    "Simplex", "SuperSimplex", "Perlin", "Value", "CellularValue", "CellularDistance", "CellularLookup",
    "FractalFBm", "PingPong", "FractalRidged",
    "DomainWarpSimplex", "DomainWarpSuperSimplex", "DomainWarpGradient",
    "DomainWarpFractalProgressive", "DomainWarpFractalIndependent",
    "Add", "Subtract", "Multiply", "Divide", "Abs", "Min", "Max", "MinSmooth", "MaxSmooth",
    "SignedSquareRoot", "PowFloat", "PowInt", "DomainScale", "DomainOffset", "DomainRotate", "DomainAxisScale",
    "SeedOffset", "ConvertRGBA8", "GeneratorCache", "Fade", "Remap", "Terrace", "AddDimension",
    "RemoveDimension", "Modulus", "DomainRotatePlane",
]


def declared_symbols():
    syms = []
    for sub, _, files in os.walk(os.path.join(ROOT, "include")):
        for f in files:
            if not f.endswith(".h"):
                continue
            text = open(os.path.join(sub, f)).read()
            if 'extern "C"' not in text:
                continue
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms += re.findall(r"FASTNOISE_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    c_abi = [s for s in syms if s.startswith("fn") and not s.startswith("fn2")]
    assert len(c_abi) == 45, c_abi
    for s in syms:
        assert hasattr(L, s), "missing export " + s
    # the Python binding declares the same set
    assert sorted(_lib.C_ABI) == sorted(c_abi)
    assert set(_lib.GPU_EXT) == set(syms) - set(c_abi)


def test_node_ids_follow_registration_order():
    L = _lib.load()
    assert L.fnGetMetadataCount() == 47
    names = [L.fnGetMetadataName(i).decode() for i in range(47)]
    assert names == NODE_ORDER


def test_member_orders_match_reference():
    m = fn.meta("CellularDistance")
    assert [v.name for v in m.variables] == ["Feature Scale", "Seed Offset", "Output Min", "Output Max",
                                             "Distance Function", "Distance Index 0", "Distance Index 1", "Return Type"]
    assert [h[0] for h in m.hybrids] == ["Minkowski P", "Grid Jitter", "Size Jitter"]
    m = fn.meta("FractalFBm")
    assert [v.name for v in m.variables] == ["Octaves", "Lacunarity"]
    assert [l[0] for l in m.lookups] == ["Source"]
    assert [h[0] for h in m.hybrids] == ["Gain", "Weighted Strength"]
    m = fn.meta("DomainWarpSimplex")
    assert [(v.name, v.dim) for v in m.variables] == [("Feature Scale", -1), ("Seed Offset", -1), ("Amplitude Scaling", 0),
                                                      ("Amplitude Scaling", 1), ("Amplitude Scaling", 2),
                                                      ("Amplitude Scaling", 3), ("Vectorization Scheme", -1)]
    m = fn.meta("Remap")
    assert [h[0] for h in m.hybrids] == ["From Min", "From Max", "To Min", "To Max"]
    for n in ("Perlin", "Simplex", "SuperSimplex", "Value"):
        assert [v.name for v in fn.meta(n).variables] == ["Feature Scale", "Seed Offset", "Output Min", "Output Max"]


def test_invalid_ids_and_indices():
    L = _lib.load()
    assert L.fnGetMetadataName(47) == b"INVALID NODE ID"
    assert L.fnGetMetadataName(-1) == b"INVALID NODE ID"  # narrowed to 255
    assert L.fnGetMetadataVariableName(8, 99) == b"INVALID VARIABLE INDEX"
    assert L.fnGetMetadataEnumName(11, 4, 99) == b"INVALID ENUM INDEX"
    assert L.fnGetMetadataNodeLookupName(8, 0) == b"INVALID NODE LOOKUP INDEX"
    assert L.fnGetMetadataHybridName(8, 0) == b"INVALID HYBRID INDEX"
    assert L.fnGetMetadataGroupName(8, 9) == b"INVALID GROUP INDEX"
    assert L.fnGetMetadataVariableCount(200) == -1
    assert L.fnGetMetadataVariableType(8, 99) == -1
    assert L.fnGetMetadataEnumCount(8, 99) == -1
    assert L.fnGetMetadataVariableDefaultFloat(8, 99) == 0.0
    assert L.fnNewFromMetadata(200, 0) is None
    # ids are narrowed to uint8_t as in the reference: 256 + 8 is Perlin
    assert L.fnGetMetadataName(256 + 8) == b"Perlin"


def test_defaults_and_ranges():
    L = _lib.load()
    perlin = NODE_ORDER.index("Perlin")
    assert L.fnGetMetadataVariableDefaultFloat(perlin, 0) == 100.0
    assert L.fnGetMetadataVariableDefaultFloat(perlin, 2) == -1.0
    assert L.fnGetMetadataVariableDefaultFloat(perlin, 3) == 1.0
    fbm = NODE_ORDER.index("FractalFBm")
    assert L.fnGetMetadataVariableDefaultIntEnum(fbm, 0) == 3
    assert L.fnGetMetadataHybridDefault(fbm, 0) == 0.5
    cd = NODE_ORDER.index("CellularDistance")
    assert L.fnGetMetadataEnumCount(cd, 7) == 5
    assert L.fnGetMetadataEnumName(cd, 7, 1) == b"Index0Add1"


def test_lifetime_and_setters():
    L = _lib.load()
    h = L.fnNewFromMetadata(NODE_ORDER.index("FractalFBm"), 0)
    assert h
    assert L.fnGetMetadataID(h) == NODE_ORDER.index("FractalFBm")
    assert L.fnGetActiveFeatureSet(h) == fn.FEATURE_SET_GPU
    src = L.fnNewFromMetadata(NODE_ORDER.index("Simplex"), 0)
    assert L.fnSetNodeLookup(h, 0, src)
    assert not L.fnSetNodeLookup(h, 1, src)
    assert L.fnSetHybridFloat(h, 0, 0.25)
    assert L.fnSetHybridNodeLookup(h, 1, src)
    assert not L.fnSetHybridFloat(h, 2, 1.0)
    assert L.fnSetVariableIntEnum(h, 0, 6)
    assert L.fnSetVariableFloat(h, 1, 2.5)
    assert not L.fnSetVariableFloat(h, 2, 1.0)
    L.fnDeleteNodeRef(src)   # the fractal still holds the source
    info = fn.Node._adopt(h).program_info(2)
    assert info["instructions"] > 0


def test_feature_set_mapping():
    L = _lib.load()
    for fs in (0, fn.FEATURE_SET_GPU, 0xFFFFFFFF):
        h = L.fnNewFromMetadata(8, fs)
        assert h, fs
        L.fnDeleteNodeRef(h)
    # CPU SIMD levels below the GPU value are not provided by this library
    assert L.fnNewFromMetadata(8, 1) is None
    assert L.fnNewFromEncodedNodeTree(b"DQkGDA==", 1) is None


def test_decodes_reference_strings():
    L = _lib.load()
    # FastNoise.h:53 of the reference: FBm over Simplex
    h = L.fnNewFromEncodedNodeTree(b"DQkGDA==", 0)
    assert h and L.fnGetMetadataID(h) == NODE_ORDER.index("FractalFBm")
    L.fnDeleteNodeRef(h)
    assert S("FractalFBm", source=S("Simplex")).encode() == "DQkGDA=="
    for bad in (b"", b"!!!!", b"DQkG", b"/w==", b"DQkGDA"):
        assert L.fnNewFromEncodedNodeTree(bad, 0) is None, bad


def test_base64_run_length():
    for data in (b"", b"\x00", b"\x00\x00\x00", bytes(300), bytes(range(256)), b"\x01" + bytes(200) + b"\x02"):
        enc = b64encode(data)
        assert b64decode(enc) == data
    # runs of 3 or more 'A' collapse to '@' + digit(run - 3)
    assert b64encode(bytes(3)) == "@B"
    assert b64encode(bytes(6)) == "@F"


def test_shared_nodes_encode_as_references():
    shared = S("Simplex", feature_scale=45.0)
    tree = S("Add", lhs=shared, rhs=S("Multiply", lhs=shared, rhs=0.5))
    raw = b64decode(tree.encode())
    assert b"\xff" in raw
    n = fn.Node.from_encoded(tree.encode())
    assert n.name == "Add"


def test_generation_without_gpu_fails_loudly():
    if fn.device_count() > 0:
        pytest.skip("a GPU is visible; this checks the no-device error path")
    n = fn.Node.from_encoded("DQkGDA==")
    with pytest.raises(RuntimeError, match="no HIP device"):
        n.gen_uniform_grid_2d(0, 0, 8, 8)
    # through the raw C ABI: output is NaN-filled and the range is NaN
    L = _lib.load()
    out = np.zeros(64, dtype=np.float32)
    mm = np.zeros(2, dtype=np.float32)
    L.fnGenUniformGrid2D(n.handle, out.ctypes.data_as(_lib.c_float_p), 0, 0, 8, 8, 1, 1, 1337,
                         mm.ctypes.data_as(_lib.c_float_p))
    assert np.isnan(out).all() and np.isnan(mm).all()
    assert "no HIP device" in fn.last_error()
    assert math.isnan(L.fnGenSingle2D(n.handle, 0.5, 0.5, 1))


def test_program_compiles_for_every_workload():
    from fastnoise2_amd import workloads as W
    for tree, dims in ((W.perlin_3d(), 3), (W.fbm_simplex_2d(), 2), (W.fbm_simplex_2d(), 4),
                       (W.warped_cellular_3d(), 3), (W.serving_tree(), 3), (W.serving_tree(), 2)):
        info = fn.Node.from_encoded(tree.encode()).program_info(dims)
        assert 0 < info["instructions"] < 4096
        assert 0 < info["registers"] <= 32
    # the FBm(Simplex) tree compiles to one fused fractal instruction
    assert fn.Node.from_encoded(W.fbm_simplex_2d().encode()).program_info(2)["instructions"] == 1
