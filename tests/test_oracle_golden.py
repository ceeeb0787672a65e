"""Golden fixtures on the CPU: the reference's own encoded trees decode (in
the oracle and in the product library) to the roots the reference gives them
and re-encode to the identical strings; the oracle reproduces the committed
vectors of tests/golden bit for bit."""
import json
import os

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd.tree import decode
import oracle
from conftest import assert_bitexact

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_TREES = json.load(open(os.path.join(GOLDEN, "reference_trees.json")))
META = json.load(open(os.path.join(GOLDEN, "oracle_vectors.json")))


@pytest.fixture(scope="module")
def vectors():
    with np.load(os.path.join(GOLDEN, "oracle_vectors.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("tree", REF_TREES, ids=[t["name"] for t in REF_TREES])
def test_reference_trees_decode_and_reencode(tree):
    enc = tree["encoded"]
    assert oracle.Tree(enc).root_id == fn.meta(tree["root"]).id
    assert fn.Node.from_encoded(enc).name == tree["root"]
    spec = decode(enc)
    assert spec.meta.name == tree["root"]
    # the reference's encoder wrote these strings: writing only non-default
    # members in metadata order must give back the same bytes
    assert spec.encode() == enc


def test_every_case_encodes_and_round_trips():
    for m in META:
        spec = decode(m["encoded"])
        assert spec.encode() == m["encoded"], m["name"]


def test_oracle_reproduces_golden_vectors(vectors):
    trees = {}
    for m in META:
        t = trees.setdefault(m["encoded"], oracle.Tree(m["encoded"]))
        for mode, tag in ((oracle.FUSED, "fused"), (oracle.UNFUSED, "unfused"), (oracle.AVX2, "avx2")):
            out, _ = t.grid(m["offset"], m["count"], m["step"], seed=m["seed"], mode=mode)
            assert_bitexact(out, vectors["%d_%s" % (m["id"], tag)], "%s %s" % (m["name"], tag))
