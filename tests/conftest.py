import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests require a device: no device is a failure, not a skip."""
    import fastnoise2_amd as fn
    n = fn.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    # the interpreter by default; tests/test_gpu_jit.py switches specialisation on
    fn.set_jit_mode(fn.JIT_OFF)
    return n


def same_bits(a, b):
    """Element-wise identity of float32 arrays (NaN equals NaN, +0 equals -0)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(a, b, what=""):
    ok = same_bits(a, b)
    if not ok.all():
        bad = np.flatnonzero(~ok)
        i = bad[0]
        diff = np.nanmax(np.abs(np.asarray(a, np.float64)[bad] - np.asarray(b, np.float64)[bad]))
        raise AssertionError("%s: %d of %d values differ (first at %d: %r vs %r; max abs diff %g)"
                             % (what, bad.size, ok.size, i, a[i], b[i], diff))


def assert_close(a, b, tol, what="", rel=False):
    """max |a - b| <= tol; with rel=True the bound per element is tol * max(1, |b|)
    (for nodes whose outputs are not confined to [-1, 1])."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert (nan_a == nan_b).all(), "%s: NaN positions differ" % what
    fin = ~nan_a & np.isfinite(b)
    assert (a[~fin & ~nan_a] == b[~fin & ~nan_a]).all(), "%s: infinities differ" % what
    d = np.abs(a - b)[fin]
    bound = tol * np.maximum(1.0, np.abs(b[fin])) if rel else tol
    worst = (d / (bound if rel else 1.0)).max() if d.size else 0.0
    assert worst <= (1.0 if rel else tol), "%s: max abs diff %g exceeds the bound (tol %g, relative %s)" % (
        what, d.max() if d.size else 0.0, tol, rel)
