"""Edge cases of the batch entry points on the GPU, through the C ABI:
empty and single-point calls, ragged sizes that do not fill a wave or a
block, NaN / infinite / huge positions, extreme seeds, and output buffers
that are device pointers. Each result is compared with the oracle."""
import ctypes

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import _lib
from fastnoise2_amd.tree import NodeSpec as S
import oracle
from conftest import assert_bitexact

pytestmark = pytest.mark.gpu

TREE = S("FractalFBm", source=S("CellularDistance", feature_scale=13.0, distance_function="Euclidean"), octaves=2)


def test_empty_calls_write_nothing(gpu):
    node = fn.Node.from_encoded(TREE.encode())
    L = _lib.load()
    sentinel = np.full(4, 7.0, dtype=np.float32)
    mm = np.zeros(2, dtype=np.float32)
    fp = _lib.c_float_p
    L.fnGenUniformGrid2D(node.handle, sentinel.ctypes.data_as(fp), 0, 0, 0, 5, 1, 1, 1, mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all() and mm[0] == np.inf and mm[1] == -np.inf
    L.fnGenUniformGrid3D(node.handle, sentinel.ctypes.data_as(fp), 0, 0, 0, 3, 0, 3, 1, 1, 1, 1, mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all() and mm[0] == np.inf
    L.fnGenPositionArray2D(node.handle, sentinel.ctypes.data_as(fp), 0, sentinel.ctypes.data_as(fp),
                           sentinel.ctypes.data_as(fp), 0, 0, 1, mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all() and mm[0] == np.inf
    L.fnGenTileable2D(node.handle, sentinel.ctypes.data_as(fp), 0, 0, 1, 1, 1, mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all() and mm[0] == np.inf
    assert fn.last_error() == ""


@pytest.mark.parametrize("count", [[1, 1], [1, 63], [65, 1], [3, 257], [255, 3], [17, 31]])
def test_ragged_2d_grids(gpu, count):
    node = fn.Node.from_encoded(TREE.encode())
    t = oracle.Tree(TREE.encode())
    out, mm = node.gen_uniform_grid_2d(-3.5, 11.25, *count, 0.75, 1.25, seed=7)
    ref, ref_mm = t.grid([-3.5, 11.25], count, [0.75, 1.25], seed=7)
    assert_bitexact(out, ref, "ragged %r" % (count,))
    assert mm == ref_mm


@pytest.mark.parametrize("count", [[1, 1, 1], [7, 1, 33], [2, 129, 2], [5, 5, 5, 5], [1, 2, 3, 4]])
def test_ragged_3d_4d_grids(gpu, count):
    node = fn.Node.from_encoded(TREE.encode())
    t = oracle.Tree(TREE.encode())
    d = len(count)
    off = [1.5, -2.25, 100.0, -7.0][:d]
    gen = {3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[d]
    out, mm = gen(*off, *count, seed=-9)
    ref, ref_mm = t.grid(off, count, seed=-9)
    assert_bitexact(out, ref, "ragged %r" % (count,))
    assert mm == ref_mm


def test_special_positions(gpu):
    node = fn.Node.from_encoded(TREE.encode())
    t = oracle.Tree(TREE.encode())
    xs = np.array([0.0, -0.0, 1e-30, -1e-30, 1e6, -1e6, 3.0e7, np.inf, -np.inf, np.nan, 2.5, 1e38], dtype=np.float32)
    ys = np.array([0.0, 0.0, 1.0, -1.0, 1e6, 5.5, -3.0e7, 0.0, 1.0, 2.0, np.nan, -1e38], dtype=np.float32)
    out, mm = node.gen_position_array([xs, ys], seed=3)
    ref, ref_mm = t.positions([xs, ys], seed=3)
    assert_bitexact(out, ref, "special positions")
    assert mm == ref_mm or (np.isnan(mm[0]) and np.isnan(ref_mm[0]))


@pytest.mark.parametrize("seed", [0, 1, -1, 2**31 - 1, -2**31])
def test_extreme_seeds(gpu, seed):
    for spec in (TREE, S("DomainWarpSimplex", source=S("Value"), warp_amplitude=20.0, seed_offset=-5)):
        node = fn.Node.from_encoded(spec.encode())
        out, _ = node.gen_uniform_grid_3d(-10, 20, 30, 9, 8, 7, seed=seed)
        ref, _ = oracle.Tree(spec.encode()).grid([-10, 20, 30], [9, 8, 7], seed=seed)
        assert_bitexact(out, ref, "seed %d" % seed)


def test_device_output_pointer_through_c_abi(gpu):
    node = fn.Node.from_encoded(TREE.encode())
    n = 37 * 41
    d = fn.DeviceArray(n)
    mm = np.zeros(2, dtype=np.float32)
    _lib.load().fnGenUniformGrid2D(node.handle, ctypes.cast(d.ptr, _lib.c_float_p), -5, 5, 37, 41, 1, 1, 11,
                                   mm.ctypes.data_as(_lib.c_float_p))
    ref, ref_mm = oracle.Tree(TREE.encode()).grid([-5, 5], [37, 41], seed=11)
    assert_bitexact(d.to_host(), ref, "device output")
    assert tuple(mm) == ref_mm


def test_concurrent_host_calls(gpu):
    import threading
    node = fn.Node.from_encoded(TREE.encode())
    ref, _ = oracle.Tree(TREE.encode()).grid([0, 0], [96, 80], seed=5)
    errors = []

    def worker():
        for _ in range(5):
            out, _ = node.gen_uniform_grid_2d(0, 0, 96, 80, seed=5)
            if not np.array_equal(out, ref):
                errors.append("mismatch")

    ths = [threading.Thread(target=worker) for _ in range(6)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors


def test_oversized_request_is_rejected(gpu):
    # 2^32 points: rejected before anything is written (the reference's int
    # product would overflow); the small sentinel buffer stays untouched
    node = fn.Node.from_encoded(TREE.encode())
    L = _lib.load()
    sentinel = np.full(4, 7.0, dtype=np.float32)
    mm = np.zeros(2, dtype=np.float32)
    fp = _lib.c_float_p
    L.fnGenUniformGrid3D(node.handle, sentinel.ctypes.data_as(fp), 0, 0, 0, 2048, 2048, 1024, 1, 1, 1, 1,
                         mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all()
    assert np.isnan(mm).all()
    assert "2^31" in fn.last_error()
    L.fnGenTileable2D(node.handle, sentinel.ctypes.data_as(fp), 1 << 30, 1 << 30, 1, 1, 1, mm.ctypes.data_as(fp))
    assert (sentinel == 7.0).all()
    assert "2^31" in fn.last_error()


@pytest.mark.parametrize("leaf", ["Simplex", "SuperSimplex", "Perlin", "Value"])
@pytest.mark.parametrize("dims", [2, 3, 4])
def test_fused_fractal_coordinate_bound(gpu, leaf, dims):
    # a specialised fused fractal drops the out-of-range select of its float
    # -> int conversions for waves whose every octave stays below 2^28 in
    # magnitude (pos * lac^k * freq); waves with a lane beyond that take the
    # guarded loop. Positions just inside and outside the bound, whole waves
    # of each and single lanes over it among small ones, against the oracle.
    freq, lac, octaves = 2.0, 2.5, 4
    spec = S("FractalFBm", source=S(leaf, feature_scale=1.0 / freq), octaves=octaves, lacunarity=lac)
    node = fn.Node.from_encoded(spec.encode())
    limit = 2.0 ** 28 / (freq * lac ** (octaves - 1))
    rng = np.random.default_rng(dims * 7 + len(leaf))
    n = 4096
    pos = [rng.uniform(-1000.0, 1000.0, n) for _ in range(dims)]
    for p in pos:
        p[1024:2048] = rng.uniform(-0.999, 0.999, 1024) * limit            # inside, near the bound
        p[3072:4096] = rng.choice([-1.0, 1.0], 1024) * rng.uniform(1.001, 7.0, 1024) * limit
    for k, v in enumerate([1.01 * limit, -1.5 * limit, 3.0e9, -2.0 ** 31, 0.9999 * limit, 65.0 * limit]):
        pos[k % dims][2048 + 64 * k + (11 * k) % 64] = v                   # one lane per wave
    pos = [p.astype(np.float32) for p in pos]
    old = fn.set_jit_mode(fn.JIT_SYNC)
    try:
        out, mm = node.gen_position_array(pos, seed=-77)
        assert node.jit_state(dims) == 1, fn.last_error()
    finally:
        fn.set_jit_mode(old)
    ref, ref_mm = oracle.Tree(spec.encode()).positions(pos, seed=-77)
    assert_bitexact(out, ref, "%s %dD fractal near the coordinate bound" % (leaf, dims))
    assert mm == ref_mm


@pytest.mark.parametrize("kind", ["FractalFBm", "FractalRidged"])
@pytest.mark.parametrize("ws", [0.0, -0.0, 0.6])
@pytest.mark.parametrize("jit", [fn.JIT_OFF, fn.JIT_SYNC], ids=["interpreter", "specialised"])
def test_fused_fractal_special_positions(gpu, kind, ws, jit):
    # the fused octave loop's weighted-strength step: with weighted strength
    # zero it only has to turn the amplitude into NaN when an octave's value
    # is not finite; infinite, NaN and huge positions reach that case
    spec = S(kind, source=S("Simplex"), octaves=4, gain=0.55, weighted_strength=ws, lacunarity=2.1)
    node = fn.Node.from_encoded(spec.encode())
    xs = np.array([0.0, -0.0, 1e-30, 1e6, -3.0e7, 3.0e9, np.inf, -np.inf, np.nan, 2.5, 1e38, -1e38], dtype=np.float32)
    ys = np.array([0.0, 1.0, -1.0, 5.5, 0.25, 1.0, 0.0, 1.0, 2.0, np.nan, -1e38, 7.0], dtype=np.float32)
    old = fn.set_jit_mode(jit)
    try:
        out, mm = node.gen_position_array([xs, ys], seed=11)
        if jit == fn.JIT_SYNC:
            assert node.jit_state(2) == 1, fn.last_error()
    finally:
        fn.set_jit_mode(old)
    ref, ref_mm = oracle.Tree(spec.encode()).positions([xs, ys], seed=11)
    assert_bitexact(out, ref, "%s ws %g special positions" % (kind, ws))
    assert mm == ref_mm or (np.isnan(mm[0]) and np.isnan(ref_mm[0]))


@pytest.mark.parametrize("kind", ["FractalFBm", "FractalRidged"])
def test_fused_simplex_bricks_on_ragged_grids(gpu, kind):
    # a specialised fused 2D Simplex fractal walks grids in 8 x 8 bricks and
    # hashes the lattice cell once per wave when all lanes share it; ragged
    # sizes leave partial bricks, and offsets from small to 1e7 move waves
    # between the shared-cell and per-lane paths and past the coordinate
    # bound of the conversion select
    spec = S(kind, source=S("Simplex", feature_scale=37.0), octaves=5, gain=0.6, lacunarity=2.0)
    node = fn.Node.from_encoded(spec.encode())
    tree = oracle.Tree(spec.encode())
    old = fn.set_jit_mode(fn.JIT_SYNC)
    try:
        for cnt in ([97, 41], [8, 8], [1, 300], [300, 1], [13, 5], [64, 9]):
            for off in ([0.0, 0.0], [-1.0e4, 3.0e4], [1.0e7, -1.0e7], [-3.5, 2.25]):
                for step in (1.0, 0.37):
                    out, mm = node.gen_uniform_grid_2d(off[0], off[1], cnt[0], cnt[1], step, step, seed=-4321)
                    assert node.jit_state(2) == 1, fn.last_error()
                    ref, ref_mm = tree.grid(off, cnt, [step, step], seed=-4321)
                    assert_bitexact(out, ref, "%s %s at %s step %g" % (kind, cnt, off, step))
                    assert mm == ref_mm
    finally:
        fn.set_jit_mode(old)
