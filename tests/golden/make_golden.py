"""Writes the golden fixtures of the parity tests (run from the repository root:
``python tests/golden/make_golden.py``).

* reference_trees.json: the encoded node trees that FastNoise2 itself ships
  (include/FastNoise/FastNoise.h:53 and
  tools/NodeEditor/util/DemoNodeTrees.inl:5-6 of the reference), with the
  root node each decodes to; data, copied verbatim as test inputs.
* oracle_vectors.json / oracle_vectors.npz: outputs of the CPU oracle (fused,
  unfused and AVX2-order variants) for every node case of tests/cases.py on
  a small grid, for the workload trees and for the reference trees. They pin
  the oracle against unintended change and let the GPU tests compare
  against stored vectors as well as against a live oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from fastnoise2_amd import workloads as W  # noqa: E402
from cases import CASES, REGION  # noqa: E402

REFERENCE_TREES = [
    {"name": "FBm over Simplex", "source": "include/FastNoise/FastNoise.h:53",
     "encoded": "DQkGDA==", "root": "FractalFBm"},
    {"name": "Mountain Terrain", "source": "tools/NodeEditor/util/DemoNodeTrees.inl:5",
     "encoded": "E@BBZEG@BD8JFgIECArXIzwECiQIw/UoPwkuAAE@BJDQAH@BC@AIEAJBw@ABZEED0KV78YZmZmPwQDmpkZPwsAAIA/HAMAAHBCBA==",
     "root": "DomainWarpSimplex"},
    {"name": "Cellular Caves", "source": "tools/NodeEditor/util/DemoNodeTrees.inl:6",
     "encoded": "FgIcCS4AAQ@BklCQs@BlRBDNzMw9G@AIMAgAw@ADgC@BCiQIzczMPgkJ@BPkIQH4XrPhjNzEw/DBIkCM3MzD4JCQ@ADBCCAE@BQzczMvhg@B/JAL/BAAL7FE4PgQKFwkNCQg@CQQQDuB4FPwt7FC4/BAOPwnU8DA==",
     "root": "Subtract"},
]

SMALL = {2: [8, 8], 3: [5, 4, 4], 4: [3, 3, 3, 3]}


def entries():
    for name, spec, d, tol, seed in CASES:
        yield name, spec.encode(), d, REGION[d]["offset"], SMALL[d], REGION[d]["step"], seed
    for key, fn_, d, off in (("workload-perlin_3d", W.perlin_3d, 3, [0, 0, 0]),
                             ("workload-fbm_simplex_2d", W.fbm_simplex_2d, 2, [-2048, -2048]),
                             ("workload-warped_cellular_3d", W.warped_cellular_3d, 3, [0, 0, 0]),
                             ("workload-serving_tree-3d", W.serving_tree, 3, [-500, 250, 1000]),
                             ("workload-serving_tree-2d", W.serving_tree, 2, [-500, 250])):
        yield key, fn_().encode(), d, off, [8] * d, [1.0] * d, 1337
    for t in REFERENCE_TREES:
        for d in (2, 3):
            yield "reference-%s-%dd" % (t["name"].replace(" ", ""), d), t["encoded"], d, [-20.5] * d, [6] * d, [3.0] * d, 1337


def main():
    with open(os.path.join(HERE, "reference_trees.json"), "w") as f:
        json.dump(REFERENCE_TREES, f, indent=1)
    meta = []
    arrays = {}
    for i, (name, enc, d, off, cnt, step, seed) in enumerate(entries()):
        t = oracle.Tree(enc)
        for mode, tag in ((oracle.FUSED, "fused"), (oracle.UNFUSED, "unfused"), (oracle.AVX2, "avx2")):
            out, _ = t.grid(off, cnt, step, seed=seed, mode=mode)
            arrays["%d_%s" % (i, tag)] = out
        meta.append({"id": i, "name": name, "encoded": enc, "dims": d, "offset": off, "count": cnt,
                     "step": step, "seed": seed})
    with open(os.path.join(HERE, "oracle_vectors.json"), "w") as f:
        json.dump(meta, f, indent=0)
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **arrays)
    print("wrote %d fixture entries" % len(meta))


if __name__ == "__main__":
    main()
