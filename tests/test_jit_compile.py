"""The run-time specialisation compiles (hipRTC, for gfx950) for every node
family and dimension count. Compiling needs no GPU; the GPU tests check that
the specialised kernels equal the interpreter and the oracle bit for bit."""
import ctypes

import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
from fastnoise2_amd.tree import NodeSpec as S

TREES = [
    ("fbm_simplex", W.fbm_simplex_2d(), 2),
    ("warped_cellular", W.warped_cellular_3d(), 3),
    ("serving", W.serving_tree(), 3),
    ("perlin_4d", S("Perlin"), 4),
    ("lookup", S("CellularLookup", lookup=S("Value")), 3),
    ("ridged_ping_pong", S("FractalRidged", source=S("PingPong", source=S("SuperSimplex"), ping_pong_strength=1.5), octaves=3), 2),
    ("warp_progressive", S("DomainWarpFractalProgressive", domain_warp_source=S("DomainWarpSimplex", source=S("Perlin")),
                           octaves=2), 3),
    ("blend_ops", S("Max", lhs=S("Abs", source=S("SineWave")), rhs=S("Modulus", lhs=S("Gradient"), rhs=3.0)), 2),
]


@pytest.mark.parametrize("name,spec,dims", TREES, ids=[t[0] for t in TREES])
def test_specialised_source_compiles(name, spec, dims):
    node = fn.Node.from_encoded(spec.encode())
    size = ctypes.c_int(0)
    rc = fn.lib().fn2JitCompileCheck(node.handle, dims, ctypes.byref(size))
    assert rc == 0, fn.last_error()
    assert size.value > 1000


def test_compile_check_rejects_null():
    assert fn.lib().fn2JitCompileCheck(None, 3, None) == 1
    assert "null" in fn.last_error()


def test_long_programs_get_register_fences():
    # programs of 128 instructions or more get a fence between instructions
    # (one fewer than instructions); shorter ones none
    long_tree = W.mixed_trees()[7]
    node = fn.Node.from_encoded(long_tree)
    n = node.program_info(3)["instructions"]
    assert n >= 128
    assert node.jit_source(3).count("registerFence<") == n - 1
    for name, spec, dims in TREES:
        short = fn.Node.from_encoded(spec.encode())
        assert short.program_info(dims)["instructions"] < 128
        assert "registerFence<" not in short.jit_source(dims)
