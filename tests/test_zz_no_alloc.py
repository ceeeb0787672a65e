"""fn2GenerateDeviceNoAlloc (include/fn2_gpu.h) against fn2GenerateDevice,
the entry point it mirrors: the same output bits and range through the C ABI,
on two calls in a row where the second is smaller (a different tree, fewer
points), so that anything the first call left in the shared device array
would show. Also the size limits, which fail before anything is launched."""
import ctypes

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
from fastnoise2_amd.tree import NodeSpec as S
from conftest import assert_bitexact

pytestmark = pytest.mark.gpu


def _call(entry, node, spec, n, minmax=True):
    out = fn.DeviceArray(max(n, 1))
    mm = np.zeros(2, dtype=np.float32)
    rc = getattr(fn.lib(), entry)(node.handle, ctypes.byref(spec), out.ptr,
                                   mm.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if minmax else None, None)
    assert rc == 0, "%s failed: %s" % (entry, fn.last_error())
    host = out.to_host()[:n]
    out.free()
    return host, (float(mm[0]), float(mm[1]))


def _positions(n, d, seed):
    rng = np.random.default_rng(seed)
    return [fn.DeviceArray.from_host(rng.uniform(-500, 500, n).astype(np.float32)) for _ in range(d)]


def _cases():
    keep = []
    big = W.serving_tree()                      # the deepest tree of the workloads
    cell = W.warped_cellular_3d()
    fbm = W.fbm_simplex_2d()
    small = S("Perlin", feature_scale=20.0)
    pos_big = _positions(300_000, 3, 1)
    pos_small = _positions(1_000, 3, 2)
    keep += pos_big + pos_small
    return keep, [
        ("grid 3D", (big, fn.grid_spec(3, [-40.0, 12.0, 5.0], [96, 80, 40], seed=11)),
         (small, fn.grid_spec(3, [0.0, 0.0, 0.0], [17, 9, 5], seed=3))),
        ("positions 3D", (cell, fn.positions_spec(pos_big, seed=1337)),
         (small, fn.positions_spec(pos_small, offset=[1.5, -2.0, 0.25], seed=8))),
        ("tileable", (fbm, fn.tileable_spec(512, 384, seed=42)),
         (small, fn.tileable_spec(33, 21, seed=5))),
        ("grid 2D then 3D", (fbm, fn.grid_spec(2, [-2048.0, -2048.0], [1024, 512], seed=42)),
         (cell, fn.grid_spec(3, [0.0, 0.0, 0.0], [16, 16, 16], seed=1337))),
    ]


def _points(spec):
    if spec.kind == fn.SOURCE_POSITIONS:
        return spec.count[0]
    if spec.kind == fn.SOURCE_TILEABLE:
        return spec.count[0] * spec.count[1]
    return int(np.prod([spec.count[a] for a in range(spec.dims)]))


def test_no_alloc_matches_generate_device(gpu):
    assert fn.lib().fn2NoAllocScratchBytes() <= 512 * 1024
    keep, cases = _cases()
    for name, first, second in cases:
        print("no-alloc case:", name, flush=True)
        for tree, spec in (first, second):
            node = fn.Node.from_encoded(tree.encode())
            n = _points(spec)
            ref, ref_mm = _call("fn2GenerateDevice", node, spec, n)
            got, got_mm = _call("fn2GenerateDeviceNoAlloc", node, spec, n)
            assert_bitexact(got, ref, "%s (%d points)" % (name, n))
            assert got_mm == ref_mm, (name, got_mm, ref_mm)
            got2, _ = _call("fn2GenerateDeviceNoAlloc", node, spec, n, minmax=False)
            assert_bitexact(got2, ref, "%s without a range" % name)
    del keep


def test_no_alloc_before_and_after_the_specialised_kernel_loads(gpu):
    # before the tree's specialised kernel is loaded the no-allocation call
    # runs the interpreter, and queues no compilation; once an ordinary call
    # has loaded the kernel, it runs that kernel: the same bits every time.
    # Specialised kernels are cached per program for the process, so the tree
    # is one no other test builds (warp amplitude 23.5)
    cell = S("CellularDistance", feature_scale=50, distance_function="Euclidean",
             return_type="Index0Add1", distance_index_0=0, distance_index_1=1, grid_jitter=1.0)
    tree = S("DomainWarpGradient", source=cell, feature_scale=100, warp_amplitude=23.5)
    node = fn.Node.from_encoded(tree.encode())
    spec = fn.grid_spec(3, [3.0, -7.0, 1.0], [64, 48, 32], seed=1337)
    n = 64 * 48 * 32
    old = fn.set_jit_mode(fn.JIT_OFF)
    try:
        assert node.jit_state(3) == -1, "another test loaded this tree's kernel"
        ref, ref_mm = _call("fn2GenerateDevice", node, spec, n)
        fn.set_jit_mode(fn.JIT_SYNC)
        interp, interp_mm = _call("fn2GenerateDeviceNoAlloc", node, spec, n)
        assert node.jit_state(3) == -1, "the no-allocation call requested a compilation"
        _call("fn2GenerateDevice", node, spec, n)
        assert node.jit_state(3) == 1
        spec_out, spec_mm = _call("fn2GenerateDeviceNoAlloc", node, spec, n)
    finally:
        fn.set_jit_mode(old)
    assert_bitexact(interp, ref, "no-allocation call on the interpreter")
    assert_bitexact(spec_out, ref, "no-allocation call on the specialised kernel")
    assert interp_mm == ref_mm and spec_mm == ref_mm


def test_no_alloc_limits_and_empty_calls(gpu):
    node = fn.Node.from_encoded(S("Perlin").encode())
    out = fn.DeviceArray(16)
    mm = np.zeros(2, dtype=np.float32)
    mmp = mm.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    spec = fn.tileable_spec(20000, 20000)
    assert fn.lib().fn2GenerateDeviceNoAlloc(node.handle, ctypes.byref(spec), out.ptr, mmp, None) != 0
    assert "no-allocation limit" in fn.last_error()
    empty = fn.grid_spec(2, [0.0, 0.0], [0, 7])
    assert fn.lib().fn2GenerateDeviceNoAlloc(node.handle, ctypes.byref(empty), out.ptr, mmp, None) == 0
    assert mm[0] == np.inf and mm[1] == -np.inf
    out.free()
