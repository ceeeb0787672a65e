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


def cell_magnitudes(dims):
    """Every squared length a cellular jitter vector can have: its components
    are (odd integer) / 2 with |odd| <= 2^bits - 1 (11 bits in 2D, 10 in 3D,
    8 in 4D), so the length is N / 4 for a sum N of `dims` odd squares, which
    is N = dims (mod 8); all such N up to the maximum (a superset)."""
    top = {2: 2047, 3: 1023, 4: 255}[dims]
    n = np.arange(dims, dims * top * top + 1, 8, dtype=np.float64)
    return (n / 4).astype(np.float32)


@pytest.mark.parametrize("dims", [2, 3, 4])
def test_bare_sqrt_and_reciprocal_exact_on_cell_magnitudes(gpu, dims):
    # the cellular nodes' 1 / sqrt(|v|^2) on the exact values it can take:
    # whether the bare hardware square root and reciprocal round correctly there
    x = cell_magnitudes(dims)
    assert (x.astype(np.float64) * 4 == np.arange(dims, dims * {2: 2047, 3: 1023, 4: 255}[dims] ** 2 + 1, 8)).all()
    din = fn.DeviceArray(x.size)
    dout = fn.DeviceArray(x.size)
    s = run(4, x, din, dout)
    s_ref = np.sqrt(x)
    bad_s = int((s.view(np.uint32) != s_ref.view(np.uint32)).sum())
    r = run(5, s_ref, din, dout)
    r_ref = np.float32(1.0) / s_ref
    bad_r = int((r.view(np.uint32) != r_ref.view(np.uint32)).sum())
    print("dims %d: %d values, sqrt mismatches %d, reciprocal mismatches %d" % (dims, x.size, bad_s, bad_r))
    assert bad_s == 0 and bad_r == 0, (bad_s, bad_r)
