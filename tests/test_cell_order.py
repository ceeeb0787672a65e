"""The pruned cellular walk's cell order (fn2_device_cellwarp.h:
MirroredCellOrder), compiled as host code: for 2D, 3D and 4D every one of
the 3^D cells comes exactly once, centre first, and the classes of (near
axes n, far axes f) come in order of 9 n + 25 f, fewer far axes first on
ties. The walk's results do not depend on the order (the kept slots are the
k smallest distances); this pins the order the skipping relies on."""
import os
import subprocess
import itertools

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fastnoise2_amd", "csrc")

PROGRAM = r'''
#include "fn2_device_cellwarp.h"
#include <cstdio>
template<int D> void dump()
{
    constexpr fn2d::MirroredCellOrder<D> order;
    for( int k = 0; k < fn2d::MirroredCellOrder<D>::kCells; k++ ) std::printf( "%d %d %d\n", D, order.e[k] & 0xff, order.e[k] >> 8 );
}
int main() { dump<2>(); dump<3>(); dump<4>(); return 0; }
'''


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = tmp_path_factory.mktemp("cellorder")
    src, exe = d / "order.cpp", d / "order"
    src.write_text(PROGRAM)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-D__HIP_PLATFORM_AMD__", "-I", CSRC, "-I", "/opt/rocm/include",
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode:
        pytest.skip("host compiler cannot build the device header here: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    t = {2: [], 3: [], 4: []}
    for line in out:
        if line.strip():
            dim, code, cls = map(int, line.split())
            t[dim].append((code, cls))
    return t


@pytest.mark.parametrize("dims", [2, 3, 4])
def test_mirrored_cell_order(tables, dims):
    entries = tables[dims]
    cols = [tuple((code >> (2 * a)) & 3 for a in range(dims)) for code, _ in entries]
    # every cell exactly once, columns 0 (far), 1 (centre), 2 (near) per axis
    assert sorted(cols) == sorted(itertools.product(range(3), repeat=dims))
    assert cols[0] == (1,) * dims
    score = [(9 * c.count(2) + 25 * c.count(0), c.count(0)) for c in cols]
    assert score == sorted(score)
    # class ids change exactly where (n, f) changes, and grow by one
    classes = [cls for _, cls in entries]
    nf = [(c.count(2), c.count(0)) for c in cols]
    for k in range(1, len(cols)):
        assert (classes[k] != classes[k - 1]) == (nf[k] != nf[k - 1])
        assert classes[k] - classes[k - 1] in (0, 1)
    # the near faces follow the centre
    assert all(c.count(2) == 1 and c.count(0) == 0 for c in cols[1:1 + dims])
