"""GPU parity per node kind and dimension (tests/cases.py) through the C ABI.

Each case runs on a grid with non-integer offsets and steps and on the same
points through the position-array path, and is compared with the oracle's
fused variant bit for bit (tolerance 0) or within the case's tolerance when
the node calls sin/cos/pow. The unfused and AVX2-order variants are checked
by index replay (cases.replay): max-abs 1e-5 for single nodes and 1e-4 for
graphs on every point whose decisions agree, flips counted.
"""
import numpy as np
import pytest

import fastnoise2_amd as fn
import oracle
from cases import CASES, REGION, LATTICE, node_count, replay
from conftest import assert_bitexact, assert_close

pytestmark = pytest.mark.gpu


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
    # Index replay against the unfused and AVX2-order variants of the
    # reference: where a point's decision fingerprint agrees with the fused
    # variant's, it took the same lattice cells, vertices, cells and
    # selections, and the result must be within the stated tolerance (1e-5
    # for a single node, 1e-4 for a graph, relative to the output's magnitude
    # for nodes not confined to [-1, 1]); points whose fingerprints differ are
    # flips, and are counted.
    for mode in (oracle.UNFUSED, oracle.AVX2):
        flips, ill, err, bound = replay(spec, dims, seed, tol, out, mode)
        print("%s mode %d: %d flips, %d ill-conditioned of %d points, max error %.3g (bound %.3g)"
              % (name, mode, flips, ill, out.size, err, bound))
        assert err <= bound, "%s: max error %g over agreeing points exceeds %g (oracle mode %d)" % (name, err, bound, mode)
        # the AVX2-order variant evaluates the same positions and changes only
        # the rounding of the 2D gradient products: no lattice index of a
        # single lattice node can move (the unfused variant rounds the grid
        # positions themselves differently, so its flips are counted)
        if mode == oracle.AVX2 and name.split("-")[0] in LATTICE and node_count(spec) == 1:
            assert flips == 0, "%s: %d lattice flips against the AVX2-order variant" % (name, flips)
    # the same points through GenPositionArray
    pos = _grid_points(dims)
    pout, _ = node.gen_position_array(pos, seed=seed)
    pref, _ = t.positions(pos, seed=seed)
    if tol == 0:
        assert_bitexact(pout, pref, name + " positions")
    else:
        assert_close(pout, pref, tol, name + " positions", rel=True)
