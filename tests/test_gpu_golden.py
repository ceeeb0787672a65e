"""GPU results against the committed golden vectors (tests/golden), through
the C ABI: bit-exact against the oracle's fused variant for every entry
except those whose nodes call sin/cos/pow (tolerance 1e-5, relative above 1),
which no reference fixture pins."""
import json
import os

import numpy as np
import pytest

import fastnoise2_amd as fn
from conftest import assert_bitexact, assert_close
from cases import CASES

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "oracle_vectors.json")))
TOL = {name: tol for name, _, _, tol, _ in CASES}


def test_gpu_matches_golden_vectors(gpu):
    with np.load(os.path.join(GOLDEN, "oracle_vectors.npz")) as z:
        vec = {k: z[k] for k in z.files}
    for m in META:
        node = fn.Node.from_encoded(m["encoded"])
        gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d, 4: node.gen_uniform_grid_4d}[m["dims"]]
        out, _ = gen(*m["offset"], *m["count"], *m["step"], seed=m["seed"])
        ref = vec["%d_fused" % m["id"]]
        tol = TOL.get(m["name"], 0.0)
        if tol:
            assert_close(out, ref, tol, m["name"], rel=True)
        else:
            assert_bitexact(out, ref, m["name"])
