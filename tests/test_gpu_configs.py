"""GPU parity on the benchmark workloads, through the C ABI.

Oracle: the CPU restatement in oracle/ (fused FMA sites, generic branches)
must match bit for bit; the unfused and AVX2-order variants within the
stated tolerances (1e-5 single node, 1e-4 graphs). Full-size grids are
checked on a seeded sample of points evaluated by the oracle through the
position path (grid coordinates are exact integers here, so both paths see
the same positions) plus the size-independent min/max invariant.
"""
import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
import oracle
from conftest import assert_bitexact, assert_close

pytestmark = pytest.mark.gpu

THREADS = 16


def _minmax_of(out):
    v = out[~np.isnan(out)]
    return (float(v.min()), float(v.max())) if v.size else (float("inf"), float("-inf"))


def test_perlin_3d_64(gpu):
    spec = W.perlin_3d()
    enc = spec.encode()
    t = oracle.Tree(enc)
    ref, ref_mm = t.grid([0, 0, 0], [64, 64, 64], seed=1337)
    for node in (fn.Node.from_encoded(enc), spec.create()):
        out, mm = node.gen_uniform_grid_3d(0, 0, 0, 64, 64, 64, 1, 1, 1, 1337)
        assert_bitexact(out, ref, "Perlin 3D 64^3 vs oracle (fused)")
        assert mm == ref_mm
    for mode in (oracle.UNFUSED, oracle.AVX2):
        alt, _ = t.grid([0, 0, 0], [64, 64, 64], seed=1337, mode=mode)
        assert_close(out, alt, 1e-5, "Perlin 3D vs oracle mode %d" % mode)


def _grid_sample_check(node, tree, out, offset, count, seed, n_sample, tol_modes):
    rng = np.random.default_rng(1234)
    total = int(np.prod(count))
    idx = np.sort(rng.choice(total, size=min(n_sample, total), replace=False))
    pos = []
    rem = idx.copy()
    for a, c in enumerate(count):
        pos.append((rem % c).astype(np.float32) + np.float32(offset[a]))
        rem //= c
    ref, _ = tree.positions(pos, seed=seed, lanes=16, threads=THREADS)
    assert_bitexact(out[idx], ref, "grid sample vs oracle (fused)")
    for mode, tol in tol_modes:
        alt, _ = tree.positions(pos, seed=seed, mode=mode, lanes=16, threads=THREADS)
        assert_close(out[idx], alt, tol, "grid sample vs oracle mode %d" % mode)


def test_fbm_simplex_2d_4096(gpu):
    spec = W.fbm_simplex_2d()
    t = oracle.Tree(spec.encode())
    node = fn.Node.from_encoded(spec.encode())
    out, mm = node.gen_uniform_grid_2d(-2048, -2048, 4096, 4096, 1, 1, 42)
    assert np.isfinite(out).all()
    assert mm == _minmax_of(out)
    _grid_sample_check(node, t, out, [-2048, -2048], [4096, 4096], 42, 1 << 16,
                       [(oracle.UNFUSED, 1e-4), (oracle.AVX2, 1e-4)])
    # a full-resolution window bit for bit through the grid path of the oracle
    win, _ = node.gen_uniform_grid_2d(-2048, -2048, 512, 512, 1, 1, 42)
    ref, ref_mm = t.grid([-2048, -2048], [512, 512], seed=42, lanes=16, threads=THREADS)
    assert_bitexact(win, ref, "FBm simplex 512^2 window")


def test_fbm_simplex_tileable_1024(gpu):
    spec = W.fbm_simplex_2d()
    t = oracle.Tree(spec.encode())
    node = fn.Node.from_encoded(spec.encode())
    out, mm = node.gen_tileable_2d(1024, 1024, 1, 1, 42)
    ref, ref_mm = t.tileable(1024, 1024, 1, 1, seed=42, lanes=16, threads=THREADS)
    # the torus coordinates are tabulated with the host's cosf/sinf, as the oracle computes them
    assert_bitexact(out, ref, "FBm simplex tileable 1024^2")
    assert mm == ref_mm == _minmax_of(out)
    for mode in (oracle.UNFUSED, oracle.AVX2):
        alt, _ = t.tileable(1024, 1024, 1, 1, seed=42, mode=mode, lanes=16, threads=THREADS)
        assert_close(out, alt, 1e-4, "tileable vs oracle mode %d" % mode)
    # tileable: the right column continues into the left one
    img = out.reshape(1024, 1024)
    assert np.abs(img[:, 0] - img[:, -1]).max() < 0.1


def test_warped_cellular_3d_256(gpu):
    spec = W.warped_cellular_3d()
    t = oracle.Tree(spec.encode())
    node = fn.Node.from_encoded(spec.encode())
    out, mm = node.gen_uniform_grid_3d(0, 0, 0, 256, 256, 256, 1, 1, 1, 1337)
    assert mm == _minmax_of(out)
    _grid_sample_check(node, t, out, [0, 0, 0], [256, 256, 256], 1337, 1 << 15,
                       [(oracle.UNFUSED, 1e-4), (oracle.AVX2, 1e-4)])
    win, _ = node.gen_uniform_grid_3d(0, 0, 0, 32, 32, 32, 1, 1, 1, 1337)
    ref, _ = t.grid([0, 0, 0], [32, 32, 32], seed=1337, lanes=16, threads=THREADS)
    assert_bitexact(win, ref, "warped cellular 32^3 window")


def test_warped_cellular_positions_16m(gpu):
    spec = W.warped_cellular_3d()
    t = oracle.Tree(spec.encode())
    node = fn.Node.from_encoded(spec.encode())
    n = 1 << 24
    x, y, z = W.splitmix64_positions(n, seed=7)
    out, mm = node.gen_position_array([x, y, z], seed=1337)
    assert mm == _minmax_of(out)
    rng = np.random.default_rng(99)
    idx = np.sort(rng.choice(n, size=1 << 15, replace=False))
    ref, _ = t.positions([x[idx], y[idx], z[idx]], seed=1337, lanes=16, threads=THREADS)
    assert_bitexact(out[idx], ref, "warped cellular positions sample")
    # device-resident inputs and output give the same values
    dpos = [fn.DeviceArray.from_host(a) for a in (x, y, z)]
    dout = fn.DeviceArray(n)
    dmm = node.generate_device(fn.positions_spec(dpos, seed=1337), dout)
    assert_bitexact(dout.to_host(), out, "device-pointer path vs host path")
    assert dmm == mm


def test_gen_single_matches_grid(gpu):
    spec = W.fbm_simplex_2d()
    node = fn.Node.from_encoded(spec.encode())
    t = oracle.Tree(spec.encode())
    for (x, y) in [(0.0, 0.0), (-0.0, 3.5), (123.25, -77.5), (1e5, -1e5)]:
        v = node.gen_single(x, y, seed=42)
        r = t.single(x, y, seed=42)
        assert np.float32(v) == np.float32(r) or (np.isnan(v) and np.isnan(r))


@pytest.mark.parametrize("dims", [2, 3, 4])
def test_mixed_serving_trees(gpu, dims):
    # the 8 seeded random trees of the mixed serving workload (depth 3..12)
    for enc in W.mixed_trees():
        node = fn.Node.from_encoded(enc)
        t = oracle.Tree(enc)
        off = [-1234.5, 777.25, 31.0, -8.5][:dims]
        cnt = [24, 20, 12, 6][:dims]
        gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[dims]
        out, mm = gen(*off, *cnt, seed=99)
        ref, ref_mm = t.grid(off, cnt, seed=99)
        assert_bitexact(out, ref, "mixed tree %s %dD" % (enc[:12], dims))
        assert mm == ref_mm
