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

