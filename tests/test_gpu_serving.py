"""Serving engine on the GPU: every request's result is the same alone (one
plain API call) and in the mix (continuously batched with requests of other
trees, sizes, dimensions and seeds, submitted from several threads), bit for
bit, in all three engine modes; batched results also match the oracle."""
import threading

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
from fastnoise2_amd.serve import ServeEngine, SERVE_PLAIN, SERVE_BATCHED, SERVE_GRAPH, SERVE_PERSISTENT
from fastnoise2_amd.tree import NodeSpec as S
import oracle
from conftest import assert_bitexact

pytestmark = pytest.mark.gpu

TREES = [
    W.serving_tree(),
    W.fbm_simplex_2d(),
    W.warped_cellular_3d(),
    S("FractalRidged", source=S("Perlin", feature_scale=40.0), octaves=4),
    S("Terrace", source=S("CellularValue", feature_scale=20.0), step_count=6.0),
]


def make_requests(n, seed=0):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        t = int(rng.integers(len(TREES)))
        kind = rng.choice(["grid3", "grid2", "tile", "grid4"], p=[0.45, 0.35, 0.1, 0.1])
        s = int(rng.integers(-2**31, 2**31))
        if kind == "grid3":
            c = int(rng.choice([8, 16, 24, 32]))
            o = [float(v) for v in rng.integers(-1000, 1000, 3) * c]
            spec = fn.grid_spec(3, o, [c, c, c], seed=s)
        elif kind == "grid2":
            c = int(rng.choice([33, 64, 128, 200]))
            o = [float(v) for v in rng.integers(-1000, 1000, 2) * c]
            spec = fn.grid_spec(2, o, [c, c + 7], seed=s)
        elif kind == "grid4":
            spec = fn.grid_spec(4, [1.5, -3.0, 7.0, 2.5], [6, 5, 4, 3], seed=s)
        else:
            spec = fn.tileable_spec(int(rng.choice([32, 64])), 48, seed=s)
        reqs.append((t, spec))
    return reqs


def alone(nodes, t, spec):
    n = nodes[t]
    if spec.kind == fn.SOURCE_TILEABLE:
        return n.gen_tileable_2d(spec.count[0], spec.count[1], spec.step[0], spec.step[1], spec.seed)
    d = spec.dims
    gen = {2: n.gen_uniform_grid_2d, 3: n.gen_uniform_grid_3d, 4: n.gen_uniform_grid_4d}[d]
    return gen(*list(spec.offset)[:d], *list(spec.count)[:d], *list(spec.step)[:d], seed=spec.seed)


@pytest.mark.parametrize("mode", [SERVE_BATCHED, SERVE_GRAPH, SERVE_PLAIN, SERVE_PERSISTENT])
def test_alone_equals_in_mix(gpu, mode):
    nodes = [fn.Node.from_encoded(t.encode()) for t in TREES]
    reqs = make_requests(120, seed=mode)
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=mode)
    tids = [eng.add_tree(t.encode()) for t in TREES]
    ids = [None] * len(reqs)

    def submitter(k):
        for i in range(k, len(reqs), 4):
            t, spec = reqs[i]
            ids[i] = eng.submit(tids[t], spec)

    threads = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    results = [eng.wait(i) for i in ids]
    st = eng.stats()
    assert st["completed"] == len(reqs)
    if mode != SERVE_PLAIN:
        for t in tids:
            assert all(eng.tree_tuning(t, d) in (0, 1, 2, 4, 8) for d in (2, 3, 4))
        assert any(eng.tree_tuning(t, 3) in (1, 2, 4, 8) for t in tids)
    if mode != SERVE_PLAIN:
        assert st["batches"] < len(reqs)
    for (t, spec), (out, mm, lat) in zip(reqs, results):
        ref, ref_mm = alone(nodes, t, spec)
        assert_bitexact(out, ref, "request alone vs in the mix")
        assert mm == ref_mm
        assert lat > 0
    eng.close()


def test_batched_matches_oracle(gpu):
    reqs = make_requests(24, seed=11)
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=SERVE_BATCHED)
    tids = [eng.add_tree(t.encode()) for t in TREES]
    ids = [eng.submit(tids[t], spec) for t, spec in reqs]
    trees = [oracle.Tree(t.encode()) for t in TREES]
    for (t, spec), rid in zip(reqs, ids):
        out, mm, _ = eng.wait(rid)
        if spec.kind == fn.SOURCE_TILEABLE:
            ref, ref_mm = trees[t].tileable(spec.count[0], spec.count[1], seed=spec.seed)
        else:
            d = spec.dims
            ref, ref_mm = trees[t].grid(list(spec.offset)[:d], list(spec.count)[:d], list(spec.step)[:d], seed=spec.seed)
        assert_bitexact(out, ref, "batched request vs oracle")
        assert mm == ref_mm
    eng.close()


def test_graph_replays_stay_correct(gpu):
    # the same batch shape replayed many times, with eager work in between
    node = fn.Node.from_encoded(TREES[3].encode())
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 16, mode=SERVE_GRAPH)
    tid = eng.add_tree(TREES[3].encode())
    for rep in range(6):
        spec = fn.grid_spec(3, [rep * 16.0, 0, 0], [16, 16, 16], seed=rep)
        out, mm, _ = eng.wait(eng.submit(tid, spec))
        ref, ref_mm = node.gen_uniform_grid_3d(rep * 16.0, 0, 0, 16, 16, 16, seed=rep)
        assert_bitexact(out, ref, "graph replay %d" % rep)
        assert mm == ref_mm
    eng.close()


def test_admission_splits_oversized_and_rejects_bad_requests(gpu):
    eng = ServeEngine(hbm_budget_bytes=1 << 28, max_batch_points=1 << 12, mode=SERVE_BATCHED)
    tid = eng.add_tree(TREES[1].encode())
    # four times the batch capacity: four parts, one request
    spec = fn.grid_spec(2, [0, 0], [128, 128])
    out, mm, _ = eng.wait(eng.submit(tid, spec))
    ref, ref_mm = alone([fn.Node.from_encoded(TREES[1].encode())], 0, spec)
    assert_bitexact(out, ref, "oversized request")
    assert mm == ref_mm
    with pytest.raises(RuntimeError, match="unknown tree"):
        eng.submit(tid + 5, fn.grid_spec(2, [0, 0], [8, 8]))
    with pytest.raises(ValueError):
        eng.add_tree("not a tree")
    out, mm, _ = eng.wait(eng.submit(tid, fn.grid_spec(2, [0, 0], [64, 64])))
    assert np.isfinite(out).all()
    eng.close()
    with pytest.raises(RuntimeError, match="budget"):
        ServeEngine(hbm_budget_bytes=1 << 20, max_batch_points=1 << 22, mode=SERVE_BATCHED)


@pytest.mark.parametrize("mode", [SERVE_BATCHED, SERVE_GRAPH, SERVE_PLAIN, SERVE_PERSISTENT])
def test_outputs_to_pageable_pinned_and_device_memory(gpu, mode):
    # one batch holding all three destination kinds: staged through the slot's
    # pinned buffer, DMA straight into pinned memory, device-to-device
    nodes = [fn.Node.from_encoded(t.encode()) for t in TREES]
    reqs = make_requests(30, seed=40 + mode)
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=mode)
    tids = [eng.add_tree(t.encode()) for t in TREES]
    outs = []
    for k, (t, spec) in enumerate(reqs):
        n = int(np.prod(list(spec.count)[:spec.dims])) if spec.kind != fn.SOURCE_TILEABLE else spec.count[0] * spec.count[1]
        out = [None, fn.PinnedArray(n), fn.DeviceArray(n)][k % 3]
        outs.append(out)
    ids = [eng.submit(tids[t], spec, out) for (t, spec), out in zip(reqs, outs)]
    for (t, spec), rid, out in zip(reqs, ids, outs):
        got, mm, _ = eng.wait(rid)
        if isinstance(got, fn.PinnedArray):
            got = got.array.copy()
        elif isinstance(got, fn.DeviceArray):
            got = got.to_host()
        ref, ref_mm = alone(nodes, t, spec)
        assert_bitexact(got, ref, "output kind %s" % type(out).__name__)
        assert mm == ref_mm
    eng.close()


def test_graph_replays_with_varying_request_sizes(gpu):
    # batches whose launch shape repeats while their point counts differ: the
    # result copies are issued per batch, outside the recorded graph
    node = fn.Node.from_encoded(TREES[1].encode())
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=SERVE_GRAPH)
    tid = eng.add_tree(TREES[1].encode())
    sizes = [(548, 547), (550, 546), (200, 1), (256, 1), (100, 2), (16, 16), (511, 513), (513, 511)]
    for rep in range(2):
        for k, (x, y) in enumerate(sizes):
            spec = fn.grid_spec(2, [k * 1000.0, rep * 500.0], [x, y], seed=k + rep)
            out, mm, _ = eng.wait(eng.submit(tid, spec))
            ref, ref_mm = node.gen_uniform_grid_2d(k * 1000.0, rep * 500.0, x, y, seed=k + rep)
            assert_bitexact(out, ref, "graph replay size %dx%d" % (x, y))
            assert mm == ref_mm
    eng.close()


def test_persistent_queue_with_more_tiles_than_resident_blocks(gpu):
    # the interpreter's work-queue kernel over batches of large requests of
    # every tree: a batch of the four or more that follow the first holds
    # at least 8.4 M points, at least 4096 tiles at any points per lane,
    # while at most 2048 blocks are resident (8 per CU), so waves pull items
    # beyond their first from the launch's counters; the engine's counter
    # read-back shows the pulls happened
    nodes = [fn.Node.from_encoded(t.encode()) for t in TREES]
    old = fn.set_jit_mode(fn.JIT_OFF)
    try:
        eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 24, mode=SERVE_PERSISTENT)
        tids = [eng.add_tree(t.encode()) for t in TREES]
        reqs = []
        for k in range(2):
            for t in range(len(TREES)):
                c = 128 if k == 0 else 57
                reqs.append((t, fn.grid_spec(3, [k * 3000.0, -float(t) * 100, 7.0], [c, c, c], seed=11 * t + k)))
        ids = [eng.submit(tids[t], spec) for t, spec in reqs]
        for (t, spec), rid in zip(reqs, ids):
            out, mm, _ = eng.wait(rid)
            ref, ref_mm = alone(nodes, t, spec)
            assert_bitexact(out, ref, "persistent tree %d count %d" % (t, spec.count[0]))
            assert mm == ref_mm
        assert eng.stats()["batches"] < len(reqs)
        launches, pulled, pulls = eng.queue_stats()
        print("work-queue launches %d, with pulls %d, pulls %d" % (launches, pulled, pulls))
        assert launches >= 1 and pulled >= 1 and pulls > 0, (launches, pulled, pulls)
        eng.close()
    finally:
        fn.set_jit_mode(old)


@pytest.mark.parametrize("mode", [SERVE_BATCHED, SERVE_GRAPH, SERVE_PLAIN, SERVE_PERSISTENT])
def test_mixed_trees_alone_in_mix_and_oracle(gpu, mode):
    # the 8 seeded random trees of the mixed serving workload (depth 3..12):
    # requests of 2^10..2^16 points in 2D/3D/4D from four threads, each equal
    # alone and in the mix; one small request per tree against the oracle
    encs = W.mixed_trees()
    nodes = [fn.Node.from_encoded(e) for e in encs]
    rng = np.random.default_rng(90 + mode)
    reqs = []
    for i in range(48):
        t = i % len(encs) if i < len(encs) else int(rng.integers(len(encs)))
        d = int(rng.integers(2, 5))
        pts = 2 ** rng.uniform(10, 12) if i < len(encs) else 2 ** rng.uniform(10, 16)
        side = max(2, int(round(pts ** (1.0 / d))))
        origin = [float(v) for v in rng.integers(-100000, 100000, d)]
        reqs.append((t, fn.grid_spec(d, origin, [side] * d, seed=int(rng.integers(-2**31, 2**31)))))
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 20, mode=mode)
    tids = [eng.add_tree(e) for e in encs]
    ids = [None] * len(reqs)

    def submitter(k):
        for i in range(k, len(reqs), 4):
            t, spec = reqs[i]
            ids[i] = eng.submit(tids[t], spec)

    threads = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    results = [eng.wait(i) for i in ids]
    trees = [oracle.Tree(e) for e in encs]
    for i, ((t, spec), (out, mm, _)) in enumerate(zip(reqs, results)):
        ref, ref_mm = alone(nodes, t, spec)
        assert_bitexact(out, ref, "mixed tree %d alone vs in the mix" % t)
        assert mm == ref_mm
        if i < len(encs):
            d = spec.dims
            oref, oref_mm = trees[t].grid(list(spec.offset)[:d], list(spec.count)[:d], list(spec.step)[:d], seed=spec.seed)
            assert_bitexact(out, oref, "mixed tree %d vs oracle" % t)
            assert mm == oref_mm
    eng.close()


@pytest.mark.parametrize("mode", [SERVE_BATCHED, SERVE_GRAPH, SERVE_PLAIN, SERVE_PERSISTENT])
def test_requests_beyond_the_batch_capacity_are_split(gpu, mode):
    # a slot holds 2^18 points; larger requests run as parts over points
    # base + [0, part) of the whole grid or image, in successive batches with
    # other requests between them, and each equals its single call bit for
    # bit, range included
    nodes = [fn.Node.from_encoded(t.encode()) for t in TREES]
    eng = ServeEngine(hbm_budget_bytes=1 << 30, max_batch_points=1 << 18, mode=mode)
    tids = [eng.add_tree(t.encode()) for t in TREES]
    reqs = [(0, fn.grid_spec(3, [-17.5, 40.25, 3.0], [80, 81, 79], seed=5)),           # 511,920 points: 2 parts
            (1, fn.grid_spec(2, [100.0, -2000.0], [700, 701], seed=-3)),                # 490,700: 2 parts
            (2, fn.grid_spec(3, [1.0, 2.0, 3.0], [16, 16, 16], seed=9)),                # small, in between
            (3, fn.tileable_spec(1024, 600, seed=77)),                                  # 614,400: 3 parts
            (4, fn.grid_spec(4, [1.5, -3.0, 7.0, 2.5], [24, 23, 22, 21], seed=1)),     # 255,024: 1 part
            (1, fn.grid_spec(2, [0.5, 0.25], [1 << 18, 3], seed=11))]                   # exactly 3 parts of 2^18
    ids = [eng.submit(tids[t], spec) for t, spec in reqs]
    for (t, spec), rid in zip(reqs, ids):
        out, mm, _ = eng.wait(rid)
        ref, ref_mm = alone(nodes, t, spec)
        assert_bitexact(out, ref, "split request tree %d, %d points" % (t, out.size))
        assert mm == ref_mm
    assert eng.stats()["completed"] == len(reqs)
    eng.close()
