"""Chunk scheduling of the chunked-world client (host logic, no GPU): the
chunks within the load range of a camera, nearest first, by the rule of the
reference's mesh preview (MeshNoisePreview.cpp:327-395 of the reference):
camera moved back by half a chunk and truncated to integers, centre chunk =
that position divided by the chunk size toward zero, in range when the
squared integer distance is at most the truncated squared range."""
import math

import numpy as np
import pytest

from fastnoise2_amd.world import chunks_in_range


def brute(camera, size, load_range):
    p = [int(c - size * 0.5) for c in camera]          # float -> int truncates toward zero
    centre = [int(math.trunc(v / size)) * size for v in p]
    r = math.ceil(load_range / size)
    rsq = int(load_range * load_range)
    cand = []
    for x in range(-r, r + 1):
        for y in range(-r, r + 1):
            for z in range(-r, r + 1):
                c = (x * size + centre[0], y * size + centre[1], z * size + centre[2])
                d2 = sum((p[a] - c[a]) ** 2 for a in range(3))
                if d2 <= rsq:
                    cand.append((d2, c))
    cand.sort(key=lambda t: t[0])   # stable: loop order among equal distances
    return np.array([c for _, c in cand], dtype=np.int32).reshape(-1, 3)


@pytest.mark.parametrize("camera,size,load_range", [
    ((0.0, 0.0, 0.0), 32, 70.0),
    ((100.5, -37.25, 4000.0), 128, 512.0),
    ((-5000.0, 12.0, -77.7), 64, 300.0),
    ((1.0, 2.0, 3.0), 16, 0.0),
    ((-31.9, -32.1, 15.99), 32, 95.5),
])
def test_chunks_in_range_match_the_rule(camera, size, load_range):
    got = chunks_in_range(camera, size, load_range)
    ref = brute(camera, size, load_range)
    assert got.shape == ref.shape
    assert (got == ref).all()
    # every chunk origin is a multiple of the chunk size, and the list has no repeats
    assert (got % size == 0).all()
    assert len({tuple(c) for c in got}) == len(got)


def test_bad_arguments_are_rejected():
    with pytest.raises(ValueError):
        chunks_in_range((0.0, 0.0, 0.0), 0, 10.0)
    with pytest.raises(ValueError):
        chunks_in_range((float("nan"), 0.0, 0.0), 32, 10.0)
    with pytest.raises(ValueError):
        chunks_in_range((0.0, 0.0, 0.0), 32, -1.0)
