"""The work-per-point figures the throughput bench prices its roofline with
are the op-counting oracle's counts for each workload's tree."""
import pytest

from fastnoise2_amd import workloads as W
import oracle

WL = W.bench_workloads()


@pytest.mark.parametrize("name", sorted(WL))
def test_committed_op_counts_match_the_counting_oracle(name):
    wl = WL[name]
    s = dict(wl["ops_sample"])
    flop, valu = oracle.op_counts_per_point(wl["tree"]().encode(), s.pop("dims"), s.pop("offset"), s.pop("count"),
                                            seed=wl["seed"], **s)
    assert (flop, valu) == pytest.approx(wl["ops"], rel=1e-6), (flop, valu)


@pytest.mark.parametrize("tree", range(len(W.MIXED_TREE_OPS)))
def test_mixed_tree_op_counts_match_the_counting_oracle(tree):
    enc = W.mixed_trees()[tree]
    for d, count in W.MIXED_OPS_SAMPLE.items():
        got = oracle.op_counts_per_point(enc, d, [0.0] * d, count)
        assert got == pytest.approx(W.MIXED_TREE_OPS[tree][d], rel=1e-6), (tree, d, got)
        # the figures hold away from the origin too
        assert oracle.op_counts_per_point(enc, d, [-777.25] * d, count) == got
