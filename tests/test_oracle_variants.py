"""Index replay between the oracle's variants on every GPU parity case
(tests/cases.py), on the CPU: the fused variant (which the GPU reproduces
bit for bit) against the unfused and AVX2-order variants. Where a point's
decision fingerprint agrees, the results agree within 1e-5 (single node) or
1e-4 (graph); the AVX2-order variant flips no decision anywhere, the
unfused one (whose grid positions round differently) flips a counted few."""
import numpy as np
import pytest

import oracle
from cases import CASES, REGION, replay

MAX_UNFUSED_FLIP_FRACTION = 0.01   # over all cases' points, White excluded (it hashes position bits)


def test_trace_build_matches_the_checker():
    name, spec, d, tol, seed = CASES[0]
    r = REGION[d]
    vals, _ = oracle.index_trace(spec.encode(), "grid", oracle.FUSED, seed, r["offset"], r["count"], r["step"])
    ref, _ = oracle.Tree(spec.encode()).grid(r["offset"], r["count"], r["step"], seed=seed)
    assert np.array_equal(vals, ref)


def test_fingerprints_see_lattice_moves():
    # shifting every point by one lattice cell changes every fingerprint
    from fastnoise2_amd.tree import NodeSpec as S
    pos = [np.linspace(-20.3, 31.7, 500).astype(np.float32), np.linspace(5.1, -9.9, 500).astype(np.float32)]
    enc = S("Perlin", feature_scale=1.0).encode()
    _, a = oracle.index_trace(enc, "positions", pos=pos)
    _, b = oracle.index_trace(enc, "positions", pos=[pos[0] + np.float32(1.0), pos[1]])
    assert (a != b).all()
    _, c = oracle.index_trace(enc, "positions", pos=pos)
    assert (a == c).all()


@pytest.mark.parametrize("mode", [oracle.UNFUSED, oracle.AVX2])
def test_index_replay_all_cases(mode):
    flips = points = 0
    for name, spec, d, tol, seed in CASES:
        r = REGION[d]
        fused, _ = oracle.index_trace(spec.encode(), "grid", oracle.FUSED, seed, r["offset"], r["count"], r["step"])
        f, ill, err, bound = replay(spec, d, seed, tol, fused, mode)
        assert err <= bound, (name, mode, err, bound)
        if mode == oracle.AVX2:
            assert f == 0, (name, f)
        elif not name.startswith("White"):
            flips += f
            points += fused.size
    if mode == oracle.UNFUSED:
        assert flips <= MAX_UNFUSED_FLIP_FRACTION * points, (flips, points)
