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


@pytest.mark.parametrize("name,spec,dims", [TREES[0], ("mixed_fractal", None, 3)], ids=["fbm_simplex", "mixed_fractal"])
def test_octave_unroll_variants_equal_the_oracle(gpu, jit_sync, name, spec, dims):
    # every octave-unroll variant of a tree with fused fractals (no unroll,
    # by two, by four) runs on the single-call path and gives the oracle's bits
    enc = spec.encode() if spec is not None else W.mixed_trees()[6]
    node = fn.Node.from_encoded(enc)
    off = [-37.5, 12.25, 8.0][:dims]
    cnt = [40, 33, 17][:dims]
    gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d}[dims]
    ref, ref_mm = oracle.Tree(enc).grid(off, cnt, seed=21)
    for v in (1, 2, 3):
        node.set_jit_variant(dims, v)
        out, mm = gen(*off, *cnt, seed=21)
        assert node.jit_state(dims, v) == 1, fn.last_error()
        assert_bitexact(out, ref, "%s unroll variant %d" % (name, v))
        assert mm == ref_mm
    node.set_jit_variant(dims, 0)
    with pytest.raises(ValueError):
        node.set_jit_variant(dims, 4)


def test_tuned_tiles_in_serving(gpu, jit_sync):
    # the tuner times threads per block (64/128/256) and points per lane on
    # each tree's specialised batch kernel, then its octave-unroll variants;
    # whatever it picks, every request equals the oracle
    encs = [TREES[0][1].encode(), W.mixed_trees()[6], TREES[2][1].encode()]
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 21, mode=SERVE_BATCHED)
    eng.tune_octave_unroll()
    tids = [eng.add_tree(e) for e in encs]
    dims = [2, 3, 3]
    for rnd in range(3):   # tile tuning, then the unroll variants, then their choice in use
        reqs = [(t, fn.grid_spec(dims[t], [rnd * 31.0 + t] * dims[t], [37, 29, 23][:dims[t]], seed=rnd * 7 + t))
                for t in range(3) for _ in range(3)]
        ids = [eng.submit(tids[t], spec) for t, spec in reqs]
        for (t, spec), rid in zip(reqs, ids):
            out, mm, _ = eng.wait(rid)
            d = spec.dims
            ref, ref_mm = oracle.Tree(encs[t]).grid(list(spec.offset)[:d], list(spec.count)[:d], seed=spec.seed)
            assert_bitexact(out, ref, "tree %d round %d" % (t, rnd))
            assert mm == ref_mm
    for t in range(3):
        tile = eng.tree_tile(tids[t], dims[t])
        assert tile is not None
        threads, ppt, unroll = tile
        assert threads in (64, 128, 256) and ppt in (1, 2, 4, 8) and unroll in (0, 1, 2, 4), tile
        assert eng.tree_tuning(tids[t], dims[t]) == ppt
    assert eng.tree_tile(tids[0], 4) is None   # never requested in 4D
    eng.close()


def test_persistent_specialised_queue_pulls(gpu, jit_sync):
    # the specialised work-queue kernels (fn2_jit_queue) with far more items
    # than resident waves: one request of 2^24 points (at least 8192 tiles at
    # any points per lane, against at most 2048 resident blocks) and one of
    # 200^3, each equal to its single call; the counters read back show that
    # waves pulled items
    trees = [TREES[2][1].encode(), TREES[1][1].encode()]
    nodes = [fn.Node.from_encoded(e) for e in trees]
    eng = ServeEngine(hbm_budget_bytes=1 << 31, max_batch_points=1 << 24, mode=SERVE_PERSISTENT)
    tids = [eng.add_tree(e) for e in trees]
    reqs = [(0, fn.grid_spec(3, [-7.5, 3.25, 100.0], [256, 256, 256], seed=5)),
            (1, fn.grid_spec(3, [12.0, -40.5, 3.0], [200, 200, 200], seed=6))]
    ids = [eng.submit(tids[t], spec) for t, spec in reqs]
    for (t, spec), rid in zip(reqs, ids):
        out, mm, _ = eng.wait(rid)
        ref, ref_mm = nodes[t].gen_uniform_grid_3d(*list(spec.offset)[:3], *list(spec.count)[:3], seed=spec.seed)
        assert nodes[t].jit_state(3) == 1
        assert_bitexact(out, ref, "specialised persistent tree %d" % t)
        assert mm == ref_mm
    launches, pulled, pulls = eng.queue_stats()
    print("work-queue launches %d, with pulls %d, pulls %d" % (launches, pulled, pulls))
    assert pulled >= 1 and pulls > 0, (launches, pulled, pulls)
    eng.close()


def test_jit_kernel_info(gpu, jit_sync):
    # the runtime's view of a loaded specialised kernel: registers, scratch
    # and residency; none before it is compiled
    node = fn.Node.from_encoded(W.fbm_simplex_2d().encode())
    assert node.jit_kernel_info(3) is None
    node.gen_uniform_grid_2d(0.0, 0.0, 64, 64, seed=1)
    regs, scratch, blocks = node.jit_kernel_info(2)
    print("FBm specialised kernel: %d VGPRs, %d scratch bytes, %d blocks per CU" % (regs, scratch, blocks))
    assert 0 < regs <= 512 and scratch >= 0 and 1 <= blocks <= 8
    with pytest.raises(ValueError):
        node.jit_kernel_info(2, variant=99)
