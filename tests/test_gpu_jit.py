"""Run-time specialised kernels (hipRTC) on the GPU: for each tree the
specialised kernel's output equals the interpreter's and the oracle's bit
for bit, on the single-call path and in batched serving."""
import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
from fastnoise2_amd.serve import ServeEngine, SERVE_BATCHED, SERVE_PERSISTENT
from fastnoise2_amd.tree import NodeSpec as S
import oracle
from conftest import assert_bitexact

pytestmark = pytest.mark.gpu

TREES = [
    ("fbm_simplex", W.fbm_simplex_2d(), 2),
    ("warped_cellular", W.warped_cellular_3d(), 3),
    ("serving", W.serving_tree(), 3),
    ("serving_2d", W.serving_tree(), 2),
    ("lookup_4d", S("CellularLookup", feature_scale=9.0, lookup=S("Perlin", feature_scale=4.0)), 4),
    ("blend_rotate", S("Fade", a=S("DomainRotate", source=S("Value"), yaw=0.4, pitch=0.3),
                       b=S("Terrace", source=S("SuperSimplex"), step_count=3.0, smoothness=0.5),
                       fade=S("Checkerboard", feature_scale=30.0)), 3),
]


@pytest.fixture
def jit_sync():
    old = fn.set_jit_mode(fn.JIT_SYNC)
    yield
    fn.set_jit_mode(old)


@pytest.mark.parametrize("name,spec,dims", TREES, ids=[t[0] for t in TREES])
def test_specialised_equals_interpreter_and_oracle(gpu, name, spec, dims):
    enc = spec.encode()
    node = fn.Node.from_encoded(enc)
    off = [-37.5, 12.25, 8.0, -3.0][:dims]
    cnt = [40, 33, 17, 9][:dims]
    gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[dims]
    old = fn.set_jit_mode(fn.JIT_OFF)
    interp, imm = gen(*off, *cnt, seed=21)
    fn.set_jit_mode(fn.JIT_SYNC)
    try:
        spec_out, smm = gen(*off, *cnt, seed=21)
        assert node.jit_state(dims) == 1, fn.last_error()
    finally:
        fn.set_jit_mode(old)
    ref, ref_mm = oracle.Tree(enc).grid(off, cnt, seed=21)
    assert_bitexact(interp, ref, name + " interpreter")
    assert_bitexact(spec_out, ref, name + " specialised")
    assert imm == smm == ref_mm


@pytest.mark.parametrize("mode", [SERVE_BATCHED, SERVE_PERSISTENT])
def test_specialised_batches_in_serving(gpu, jit_sync, mode):
    # batched: the specialised batch kernels; persistent: their work-queue
    # entry points (tiles pulled from a counter), one launch per tree
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=mode)
    tids = [eng.add_tree(t[1].encode()) for t in TREES[:3]]
    reqs = []
    for k in range(18):
        t = k % 3
        d = TREES[t][2]
        c = 24 if k < 15 else 48      # the last ones have hundreds of tiles
        reqs.append((t, fn.grid_spec(d, [k * 17.0] * d, [c] * d, seed=k)))
    ids = [eng.submit(tids[t], spec) for t, spec in reqs]
    for (t, spec), rid in zip(reqs, ids):
        out, mm, _ = eng.wait(rid)
        d = spec.dims
        ref, ref_mm = oracle.Tree(TREES[t][1].encode()).grid(list(spec.offset)[:d], list(spec.count)[:d], seed=spec.seed)
        assert_bitexact(out, ref, "specialised batch")
        assert mm == ref_mm
    eng.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_long_fenced_program_equals_interpreter_and_oracle(gpu, dims):
    # the deepest mixed-serving tree (about 355 instructions): specialised
    # with register fences between its instructions
    enc = W.mixed_trees()[7]
    node = fn.Node.from_encoded(enc)
    assert node.program_info(dims)["instructions"] >= 128
    off = [-101.0, 57.5, 13.0][:dims]
    cnt = [48, 37, 11][:dims]
    gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d}[dims]
    old = fn.set_jit_mode(fn.JIT_OFF)
    try:
        interp, imm = gen(*off, *cnt, seed=-7)
        fn.set_jit_mode(fn.JIT_SYNC)
        spec_out, smm = gen(*off, *cnt, seed=-7)
        assert node.jit_state(dims) == 1, fn.last_error()
    finally:
        fn.set_jit_mode(old)
    ref, ref_mm = oracle.Tree(enc).grid(off, cnt, seed=-7)
    assert_bitexact(interp, ref, "long tree interpreter")
    assert_bitexact(spec_out, ref, "long tree specialised")
    assert imm == smm == ref_mm


# every fifth node case: all node families, dimensions mixed
from cases import CASES, REGION  # noqa: E402

JIT_CASES = CASES[::5]


@pytest.mark.parametrize("name,spec,dims,tol,seed", JIT_CASES, ids=[c[0] for c in JIT_CASES])
def test_node_cases_specialised_equal_interpreter(gpu, name, spec, dims, tol, seed):
    node = fn.Node.from_encoded(spec.encode())
    r = REGION[dims]
    gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[dims]
    args = list(r["offset"]) + list(r["count"]) + list(r["step"])
    old = fn.set_jit_mode(fn.JIT_OFF)
    try:
        interp, imm = gen(*args, seed=seed)
        fn.set_jit_mode(fn.JIT_SYNC)
        spec_out, smm = gen(*args, seed=seed)
        state = node.jit_state(dims)
    finally:
        fn.set_jit_mode(old)
    assert state in (1, 3), fn.last_error()
    # same device arithmetic on both paths: identical bits, NaNs included
    assert_bitexact(spec_out, interp, name)
    assert imm == smm or (np.isnan(imm).all() and np.isnan(smm).all())
