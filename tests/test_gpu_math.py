"""Exhaustive check of the fast correctly rounded sqrt / reciprocal that the
cellular nodes use for the jitter vector length: every float of every binade
the argument can take, against the IEEE results (numpy float32 sqrt and
division are correctly rounded)."""
import numpy as np
import pytest

import fastnoise2_amd as fn

pytestmark = pytest.mark.gpu


def binade(e):
    m = np.arange(1 << 23, dtype=np.uint32)
    return (((e + 127) << 23) | m).view(np.float32)


def run(op, x, din, dout):
    lib = fn.lib()
    assert lib.fn2Memcpy(din.ptr, x.ctypes.data, x.nbytes, 0) == 0
    assert lib.fn2SelfTestMath(op, din.ptr, dout.ptr, x.size) == 0, fn.last_error()
    out = np.empty_like(x)
    assert lib.fn2Memcpy(out.ctypes.data, dout.ptr, x.nbytes, 1) == 0
    return out


@pytest.mark.parametrize("op,lo,hi", [(0, -1, 22), (1, -1, 11), (2, -1, 22), (3, -1, 2)],
                         ids=["sqrt", "reciprocal", "inv_sqrt", "ieee_inv_sqrt"])
def test_fast_math_is_correctly_rounded(gpu, op, lo, hi):
    din = fn.DeviceArray(1 << 23)
    dout = fn.DeviceArray(1 << 23)
    bad = 0
    first = None
    for e in range(lo, hi):
        x = binade(e)
        got = run(op, x, din, dout)
        if op == 0:
            ref = np.sqrt(x)
        elif op == 1:
            ref = np.float32(1.0) / x
        else:
            ref = np.float32(1.0) / np.sqrt(x)
        diff = got.view(np.uint32) != ref.view(np.uint32)
        n = int(diff.sum())
        if n and first is None:
            i = int(np.flatnonzero(diff)[0])
            first = (float(x[i]), float(got[i]), float(ref[i]))
        bad += n
    assert bad == 0, "%d mismatches, first (x, got, ieee) = %r" % (bad, first)



def cell_lengths(dims):
    """Every squared length a D-dimensional cellular jitter vector can have
    (and more): its components are odd / 2 with |odd| <= 2^b - 1 (b = 10 in
    3D, 8 in 4D), so the length is N / 4 for a sum N of D odd squares, and
    N = D (mod 8); all such N up to the largest possible sum."""
    top = {3: 1023, 4: 255}[dims]
    return (np.arange(dims, dims * top * top + 1, 8, dtype=np.float64) / 4).astype(np.float32)


@pytest.mark.parametrize("dims", [3, 4])
def test_upward_corrected_sqrt_exact_on_cell_lengths(gpu, dims):
    # the 3D / 4D cellular nodes correct the hardware square root estimate
    # upward only; on every argument they can pass that must be the
    # correctly rounded root
    x = cell_lengths(dims)
    assert np.array_equal(x.astype(np.float64) * 4, np.arange(dims, dims * {3: 1023, 4: 255}[dims] ** 2 + 1, 8))
    din = fn.DeviceArray(x.size)
    dout = fn.DeviceArray(x.size)
    got = run(5, x, din, dout)
    ref = np.sqrt(x)
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, "%d mismatches, first at x = %r" % (bad.size, float(x[bad[0]]))
