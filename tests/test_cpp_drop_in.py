"""The drop-in C++ API from a compiled C++ caller (tests/cpp/drop_in.cpp):
g++ against include/FastNoise, linked with libfastnoise2_amd.so. The program
checks New<T> for every node class (ids in the reference's registration
order), typed setters, SmartNode counts and casts, encoded trees and the
FeatureSet contract, then generates through GenUniformGrid2D/3D,
GenTileable2D, GenPositionArray3D and GenSingle into host std::vectors. On
the GPU its outputs and ranges equal the oracle's bit for bit; without a GPU
they are NaN (no silent fallback)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import fastnoise2_amd as fn
from fastnoise2_amd import workloads as W
from fastnoise2_amd.tree import NodeSpec as S
from conftest import assert_bitexact

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "fastnoise2_amd", "lib")


@pytest.fixture(scope="module")
def run_drop_in(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ is not installed")
    d = tmp_path_factory.mktemp("drop_in")
    exe = str(d / "drop_in")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "drop_in.cpp"), "-L" + LIBDIR, "-lfastnoise2_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    p = subprocess.run([exe, str(d), W.fbm_simplex_2d().encode()], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr

    def load(name):
        v = np.fromfile(str(d / (name + ".f32")), dtype=np.float32)
        r = np.fromfile(str(d / (name + ".range")), dtype=np.float32)
        return v, (float(r[0]), float(r[1]))
    return load


def test_compiled_caller_without_gpu_gets_nan(run_drop_in):
    if fn.device_count() > 0:
        pytest.skip("a GPU is present: the gpu test covers the outputs")
    for name in ("fbm_gain_value_grid2d_4x4", "warp_cell_grid3d_32x24x16", "fbm_tileable_64x48",
                 "warp_cell_positions3d_1000"):
        v, r = run_drop_in(name)
        assert v.size and np.isnan(v).all(), name
        assert np.isnan(r[0]) and np.isnan(r[1]), name


@pytest.mark.gpu
def test_compiled_caller_matches_oracle(gpu, run_drop_in):
    import oracle
    cases = [
        ("fbm_gain_value_grid2d_4x4", S("FractalFBm", source=S("Simplex"), gain=S("Value")),
         lambda t: t.grid([0.0, 0.0], [4, 4], seed=1337)),
        ("warp_cell_grid3d_32x24x16", W.warped_cellular_3d(),
         lambda t: t.grid([-3.5, 2.0, 10.0], [32, 24, 16], [1.0, 1.5, 0.75], seed=1337)),
        ("fbm_tileable_64x48", W.fbm_simplex_2d(), lambda t: t.tileable(64, 48, seed=42)),
    ]
    for name, spec, call in cases:
        v, r = run_drop_in(name)
        ref, ref_mm = call(oracle.Tree(spec.encode()))
        assert_bitexact(v, ref, name)
        assert r == ref_mm, (name, r, ref_mm)
    pos = [run_drop_in("positions_" + a)[0] for a in "xyz"]
    v, r = run_drop_in("warp_cell_positions3d_1000")
    t = oracle.Tree(W.warped_cellular_3d().encode())
    ref, ref_mm = t.positions(pos, offset=[0.5, -1.0, 2.0], seed=1337)
    assert_bitexact(v, ref, "positions")
    assert r == ref_mm
    singles, _ = run_drop_in("singles")
    assert singles[0] == np.float32(t.single(12.25, -7.5, 3.0, seed=1337))
    fbm = oracle.Tree(S("FractalFBm", source=S("Simplex"), gain=S("Value")).encode())
    assert singles[1] == np.float32(fbm.single(1.5, 2.5, seed=1337))
