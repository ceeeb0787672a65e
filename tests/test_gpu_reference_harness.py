"""The reference's own test harness shapes, run on the GPU and checked
against the oracle (the reference harness checks no values; here every value
is compared):

- tests/FastNoiseBenchmark.cpp of the reference: every registered node kind,
  its node lookups set to Constant (or, where Constant is not a valid source,
  to the first node kind in registration order that is, with its own lookups
  set to Constant) (:14-46), plus the demo encoded trees (:152-166), through
  GenPositionArray2D/3D/4D on 8192 positions x = y = z = w = idx * 0.6
  (:11, 131-134), seed 0, 1, 2 as the harness's seed++ gives them (:56-62).
- tests/FastNoiseCpp11Include.cpp of the reference: FractalFBm(Simplex) with a
  Value node as gain on a 4x4 grid at seed 1337 (:8-24)."""
import json
import os

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd.tree import NodeSpec as S, registry, meta
import oracle
from conftest import assert_bitexact, assert_close

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DEMO_TREES = [t for t in json.load(open(os.path.join(GOLDEN, "reference_trees.json")))
              if "DemoNodeTrees" in t["source"]]
POSITIONS = (np.arange(8192, dtype=np.float32) * np.float32(0.6)).astype(np.float32)
# nodes whose result goes through FS::Sin / FS::Cos / FS::Pow (FastSIMD-defined,
# no reference fixture): compared within 1e-5 relative above magnitude 1
TRANSCENDENTAL = {"SineWave", "PowFloat"}


def harness_tree(m):
    """The harness's BuildGenerator as a node spec, with the validity of each
    source checked through the C ABI (fnSetNodeLookup)."""
    lib = fn.lib()
    constant = meta("Constant")
    spec = S(m.name)
    for idx, (name, dim) in enumerate(m.lookups):
        key = name + ("xyzw"[dim] if dim >= 0 else "")
        gen = fn.Node._adopt(lib.fnNewFromMetadata(m.id, 0))
        src = fn.Node._adopt(lib.fnNewFromMetadata(constant.id, 0))
        if lib.fnSetNodeLookup(gen.handle, idx, src.handle):
            spec.set(key, S("Constant"))
            continue
        for t in registry()[0]:
            trial = fn.Node._adopt(lib.fnNewFromMetadata(t.id, 0))
            if lib.fnSetNodeLookup(gen.handle, idx, trial.handle):
                sub = S(t.name)
                for tidx, (tname, tdim) in enumerate(t.lookups):
                    inner = fn.Node._adopt(lib.fnNewFromMetadata(constant.id, 0))
                    assert lib.fnSetNodeLookup(trial.handle, tidx, inner.handle)
                    sub.set(tname + ("xyzw"[tdim] if tdim >= 0 else ""), S("Constant"))
                spec.set(key, sub)
                break
        else:
            pytest.fail("no valid source for %s.%s" % (m.name, name))
    return spec


def check(encoded, name, dims):
    node = fn.Node.from_encoded(encoded)
    t = oracle.Tree(encoded)
    for seed in (0, 1, 2):
        out, mm = node.gen_position_array([POSITIONS] * dims, seed=seed)
        ref, ref_mm = t.positions([POSITIONS] * dims, seed=seed)
        what = "%s %dD seed %d" % (name, dims, seed)
        if name in TRANSCENDENTAL:
            assert_close(out, ref, 1e-5, what, rel=True)
        else:
            assert_bitexact(out, ref, what)
            assert mm == ref_mm or (np.isnan(mm).all() and np.isnan(ref_mm).all())


@pytest.mark.parametrize("node_id", range(len(registry()[0])), ids=[m.name for m in registry()[0]])
def test_every_node_on_the_harness_positions(gpu, node_id):
    m = registry()[0][node_id]
    encoded = harness_tree(m).encode()
    for dims in (2, 3, 4):
        check(encoded, m.name, dims)


@pytest.mark.parametrize("tree", DEMO_TREES, ids=[t["name"] for t in DEMO_TREES])
def test_demo_trees_on_the_harness_positions(gpu, tree):
    for dims in (2, 3, 4):
        check(tree["encoded"], tree["name"], dims)


def test_cpp11_include_grid(gpu):
    spec = S("FractalFBm", source=S("Simplex"), gain=S("Value"))
    node = fn.Node.from_encoded(spec.encode())
    out, mm = node.gen_uniform_grid_2d(0.0, 0.0, 4, 4, 1.0, 1.0, seed=1337)
    ref, ref_mm = oracle.Tree(spec.encode()).grid([0.0, 0.0], [4, 4], [1.0, 1.0], seed=1337)
    assert_bitexact(out, ref, "FBm(Simplex, gain = Value) 4x4")
    assert mm == ref_mm
