"""CellularDistance with squared distances, in a tree's specialised kernel,
visits cells near side first and leaves out, per wave, cells that cannot
reach the kept distance slots; grid requests of such trees are walked in
bricks. Position arrays of 3D trees with one or two kept slots run a variant
of the kernel in which a wave whose lanes need many different far cells
shares them out through an LDS queue instead.
These cases stress that walk against the oracle's full walk, bit for bit:
coherent and ragged grids (where whole waves skip cells) and scattered
positions (where they rarely do), pairs of distance indices, jitter values
that are negative, zero, small, large and per-point (a node), in 2D, 3D and
4D. Every tree is run with its specialised kernel compiled first."""
import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd.tree import NodeSpec as S
import oracle
from conftest import assert_bitexact

pytestmark = pytest.mark.gpu

JITTERS = [1.0, -1.5, 0.0, 0.3, 2.5, S("Perlin", feature_scale=7.0)]


@pytest.fixture(autouse=True)
def specialised():
    old = fn.set_jit_mode(fn.JIT_SYNC)
    yield
    fn.set_jit_mode(old)


def tree(jitter, i0, i1, func, ret="Index0Add1"):
    return S("CellularDistance", feature_scale=9.0, distance_function=func, return_type=ret,
             distance_index_0=i0, distance_index_1=i1, grid_jitter=jitter)


@pytest.mark.parametrize("dims", [2, 3, 4])
@pytest.mark.parametrize("jitter", range(len(JITTERS)), ids=["j1", "jneg", "j0", "jsmall", "jbig", "jnode"])
def test_grid_walk_matches_oracle(gpu, dims, jitter):
    for i0, i1, func in [(0, 1, "Euclidean"), (3, 1, "Euclidean Squared")]:
        spec = tree(JITTERS[jitter], i0, i1, func, ret="Index0Sub1")
        node = fn.Node.from_encoded(spec.encode())
        cnt = {2: [97, 41], 3: [61, 13, 10], 4: [20, 8, 6, 5]}[dims]
        off = [-31.25, 7.5, 100.0, -3.0][:dims]
        gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[dims]
        out, mm = gen(*off, *cnt, *([0.37] * dims), seed=12345)
        assert node.jit_state(dims) == 1, fn.last_error()
        ref, ref_mm = oracle.Tree(spec.encode()).grid(off, cnt, [0.37] * dims, seed=12345)
        assert_bitexact(out, ref, "%dD jitter %d indices %d,%d %s" % (dims, jitter, i0, i1, func))
        assert mm == ref_mm


@pytest.mark.parametrize("dims", [2, 3, 4])
def test_scattered_positions_match_oracle(gpu, dims):
    rng = np.random.default_rng(dims)
    n = 5000
    # mostly scattered, some clustered, some near lattice points and half-way planes
    pos = rng.uniform(-400.0, 400.0, (dims, n)).astype(np.float32)
    pos[:, :1000] = rng.normal(0.0, 0.3, (dims, 1000)).astype(np.float32) * 9.0
    pos[:, 1000:1500] = (np.round(pos[:, 1000:1500] / 9.0) * 9.0).astype(np.float32)
    pos[:, 1500:2000] = ((np.floor(pos[:, 1500:2000] / 9.0) + 0.5) * 9.0).astype(np.float32)
    for jitter, i0, i1 in ((1.0, 0, 1), (-0.7, 0, 1), (2.5, 0, 1), (1.0, 0, 0), (0.8, 1, 0)):
        spec = tree(jitter, i0, i1, "Euclidean")
        node = fn.Node.from_encoded(spec.encode())
        out, mm = node.gen_position_array(list(pos), seed=77)
        assert node.jit_state(dims) == 1, fn.last_error()
        # 3D position arrays run the shared far-cell walk (one or two slots)
        shared = node.jit_state(dims, fn.JIT_VARIANT_SHARED_CELL_WALK)
        assert shared == (1 if dims == 3 else -1), (shared, fn.last_error())
        ref, ref_mm = oracle.Tree(spec.encode()).positions(list(pos), seed=77)
        assert_bitexact(out, ref, "%dD scattered jitter %g indices %d,%d" % (dims, jitter, i0, i1))
        assert mm == ref_mm


def test_huge_and_special_coordinates(gpu):
    spec = tree(1.0, 0, 1, "Euclidean")
    node = fn.Node.from_encoded(spec.encode())
    xs = np.array([1e7, -1e7, 3.0e9, 1e20, np.inf, np.nan, 0.0, -0.0, 4.5, 1e-30] * 7, dtype=np.float32)
    ys = np.roll(xs, 3)
    zs = np.roll(xs, 5)
    out, _ = node.gen_position_array([xs, ys, zs], seed=5)
    ref, _ = oracle.Tree(spec.encode()).positions([xs, ys, zs], seed=5)
    assert_bitexact(out, ref, "huge and special coordinates")


TABLE_NODES = [
    ("CellularValue", [dict(value_index=0, distance_function="Euclidean"), dict(value_index=2, distance_function="Manhattan"),
                       dict(value_index=1, distance_function="Hybrid")]),
    ("CellularDistance", [dict(distance_function="Manhattan", return_type="Index0Add1", distance_index_0=0, distance_index_1=1),
                          dict(distance_function="Hybrid", return_type="Index0Sub1", distance_index_0=2, distance_index_1=0),
                          dict(distance_function="Euclidean", return_type="Index0Div1", distance_index_0=0, distance_index_1=3)]),
    ("CellularLookup", [dict(distance_function="Euclidean", lookup=S("Simplex", feature_scale=3.0)),
                        dict(distance_function="Manhattan", lookup=S("White"))]),
]


@pytest.mark.parametrize("kind", range(len(TABLE_NODES)), ids=[k for k, _ in TABLE_NODES])
@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("jitter", [0, 2, 5], ids=["j1", "j0", "jnode"])
def test_table_walks_match_oracle(gpu, kind, dims, jitter):
    # In a specialised kernel, CellularValue, CellularLookup and the
    # CellularDistance cases the pruned walk leaves (other distance
    # functions, size jitter) read each cell's hash, jitter vector and
    # inverse length from the wave's LDS table when the wave's cells lie
    # close together (grids, clustered positions), walking in the
    # reference's order, which decides between tied distances (jitter 0);
    # scattered positions take the per-lane walk. Per-point jitter and size
    # jitter come from nodes, so lanes of one wave scale the same table
    # entries differently.
    name, variants = TABLE_NODES[kind]
    size_jitter = S("Value", feature_scale=5.0) if jitter == 5 else 0.25
    for members in variants:
        spec = S(name, feature_scale=9.0, grid_jitter=JITTERS[jitter], size_jitter=size_jitter, **members)
        node = fn.Node.from_encoded(spec.encode())
        cnt = {2: [97, 41], 3: [61, 13, 10]}[dims]
        off = [-31.25, 7.5, 100.0][:dims]
        gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d}[dims]
        out, mm = gen(*off, *cnt, *([0.37] * dims), seed=777)
        assert node.jit_state(dims) == 1, fn.last_error()
        ref, ref_mm = oracle.Tree(spec.encode()).grid(off, cnt, [0.37] * dims, seed=777)
        what = "%s %dD jitter %d %s" % (name, dims, jitter, members.get("distance_function"))
        assert_bitexact(out, ref, what + " grid")
        assert mm == ref_mm
        rng = np.random.default_rng(dims + 10 * jitter)
        pos = rng.uniform(-300.0, 300.0, (dims, 3000)).astype(np.float32)
        pos[:, :2000] = (rng.normal(0.0, 1.0, (dims, 2000)) * 2.0 + 40.0).astype(np.float32)   # clustered
        pos[:, 500:600] = (np.round(pos[:, 500:600] / 9.0) * 9.0).astype(np.float32)          # on lattice points
        out, mm = node.gen_position_array(list(pos), seed=777)
        ref, ref_mm = oracle.Tree(spec.encode()).positions(list(pos), seed=777)
        assert_bitexact(out, ref, what + " positions")
        assert mm == ref_mm
