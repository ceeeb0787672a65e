"""Pins the CPU oracle against what the mathematics and the reference's own
constants fix, independently of the oracle's code:

* gradient selection: the generic bit logic of the oracle against the
  lookup tables the reference writes in its AVX2 / AVX-512 branches
  (include/FastNoise/Generators/Utils.inl:35-37, 71-73, 94-95, 110-113);
* interpolants at their fixed points;
* Perlin and Value noise against a float64, table-driven re-derivation
  written from Perlin.inl / Value.inl and the primes / multiplier of
  Utils.inl:6-19 (lattice indices, hashes, gradients, interpolation and
  ScaleOutput all enter);
* special cases with closed forms: integer lattice points, jitter-0 cellular
  (distance to the nearest lattice point), constant cellular lookups;
* node identities: one-octave and zero-gain fractals equal their source, a
  zero-amplitude warp equals its source, blends and domain modifiers equal
  the element-wise / re-positioned evaluation of their inputs;
* driver invariants: 1 vs 16 lanes, 1 vs many threads, grid vs position
  arrays, tileable seam and torus coordinates.

A plausible mistake in the oracle (a wrong prime, sign, table entry, axis or
seed offset) fails at least one of these.
"""
import math

import numpy as np
import pytest

import oracle
from fastnoise2_amd.tree import NodeSpec as S
from conftest import assert_bitexact, assert_close

from f64ref import (PX, PY, PZ, PW, HASH_A, G2X, G2Y, G3X, G3Y, G3Z, G4X, G4Y, G4Z, G4W, i32, hash_primes,
                    value_coord, quintic, hermite, scale_output, perlin_ref, value_ref)


def rand_points(d, n, lo=-300.0, hi=300.0, seed=5):
    rng = np.random.default_rng(seed + d)
    return [rng.uniform(lo, hi, n).astype(np.float32) for _ in range(d)]


# ---------------------------------------------------------------- gradients
@pytest.mark.parametrize("mode", [oracle.FUSED, oracle.AVX2])
def test_gradient_2d_table(mode):
    L = oracle.load()
    for h in range(64):
        for axis, tab in ((0, G2X), (1, G2Y)):
            f = np.zeros(2, np.float32)
            f[axis] = 1.0
            g = L.fn2o_grad_dot(2, h, f.ctypes.data_as(oracle._fp), mode)
            assert g == np.float32(tab[h & 7]), (h, axis, g)


def test_gradient_3d_table():
    L = oracle.load()
    for h in range(64):
        for axis, tab in enumerate((G3X, G3Y, G3Z)):
            f = np.zeros(3, np.float32)
            f[axis] = 1.0
            assert L.fn2o_grad_dot(3, h, f.ctypes.data_as(oracle._fp), 0) == tab[h & 15], (h, axis)


def test_gradient_4d_table():
    L = oracle.load()
    for h in range(128):
        for axis, tab in enumerate((G4X, G4Y, G4Z, G4W)):
            f = np.zeros(4, np.float32)
            f[axis] = 1.0
            assert L.fn2o_grad_dot(4, h, f.ctypes.data_as(oracle._fp), 0) == tab[h & 31], (h, axis)


# ---------------------------------------------------------------- interpolants
@pytest.mark.parametrize("mode", [oracle.FUSED, oracle.UNFUSED])
def test_interpolants(mode):
    L = oracle.load()
    for kind in (0, 1):
        assert L.fn2o_interp(kind, 0.0, mode) == 0.0
        assert L.fn2o_interp(kind, 1.0, mode) == 1.0
        assert L.fn2o_interp(kind, 0.5, mode) == 0.5
        ts = np.linspace(0, 1, 101)
        vals = [L.fn2o_interp(kind, float(t), mode) for t in ts]
        ref = hermite(ts) if kind == 0 else quintic(ts)
        assert np.abs(np.array(vals) - ref).max() < 1e-6
        assert all(b >= a for a, b in zip(vals, vals[1:]))


# ---------------------------------------------------------------- hashes
def test_hash_probes_match_reference_constants():
    L = oracle.load()
    rng = np.random.default_rng(3)
    for _ in range(200):
        seed = int(rng.integers(-2**31, 2**31))
        n = int(rng.integers(1, 5))
        primed = [int(v) for v in rng.integers(-2**31, 2**31, n)]
        arr = (oracle.ctypes.c_int * n)(*primed)
        assert L.fn2o_hash_primes(seed, n, arr) == int(hash_primes(seed, *primed))
        hb = np.int64(seed)
        for p in primed:
            hb ^= p
        assert L.fn2o_hash_primes_hb(seed, n, arr) == int(i32(hb * HASH_A))
        assert L.fn2o_value_coord(seed, n, arr) == np.float32(value_coord(seed, *primed))


# ---------------------------------------------------------------- lattice noise
@pytest.mark.parametrize("d", [2, 3, 4])
def test_perlin_matches_float64_derivation(d):
    pos = rand_points(d, 4000)
    t = oracle.Tree(S("Perlin", feature_scale=37.0).encode())
    out, _ = t.positions(pos, seed=1337)
    ref = perlin_ref(pos, 1337, 1 / 37.0)
    assert np.abs(out - ref).max() < 2e-5


@pytest.mark.parametrize("d", [2, 3, 4])
def test_value_matches_float64_derivation(d):
    pos = rand_points(d, 4000)
    t = oracle.Tree(S("Value", feature_scale=23.0, seed_offset=3).encode())
    out, _ = t.positions(pos, seed=-99)
    ref = value_ref(pos, -99 + 3, 1 / 23.0)
    assert np.abs(out - ref).max() < 2e-5


@pytest.mark.parametrize("d", [2, 3, 4])
def test_lattice_points_closed_forms(d):
    # at integer positions (frequency 1) Perlin's offsets are 0 and Value
    # reduces to the corner's value hash; ScaleOutput is an fma of float constants
    pos = [np.arange(-20, 20, dtype=np.float32)[(np.arange(40) * (a + 3)) % 40] for a in range(d)]
    per, _ = oracle.Tree(S("Perlin", feature_scale=1.0).encode()).positions(pos, seed=7)
    bound = np.float32({2: 1.726796627044677734375, 3: 1.0363423824310302734375, 4: 1.33858621120452880859375}[d])
    s = np.float32(np.float32(1.0) / (bound + bound)) * np.float32(2.0)
    assert (per == np.float32(np.float64(s) * np.float64(bound) - 1.0)).all()
    val, _ = oracle.Tree(S("Value", feature_scale=1.0).encode()).positions(pos, seed=7)
    prim = [i32(np.asarray(p, np.int64) * pr) for p, pr in zip(pos, (PX, PY, PZ, PW))]
    v = value_coord(7, *prim).astype(np.float32)
    tt = (v + np.float32(2147483648.0)).astype(np.float32)
    expect = (tt.astype(np.float64) * 2.0 ** -31 - 1.0).astype(np.float32)
    assert_bitexact(val, expect, "Value at lattice points")


def test_simplex_zero_at_origin_cells():
    # Simplex / SuperSimplex contributions vanish at lattice vertices of the
    # skewed grid only in special cases; the origin is one for every seed
    for name in ("Simplex", "SuperSimplex", "Perlin"):
        for d in (2, 3, 4):
            t = oracle.Tree(S(name, feature_scale=1.0).encode())
            v = t.single(*([0.0] * d), seed=11)
            # output = ScaleOutput( 0 ) = fma( 2 / (2 bound), bound, -1 ), zero up to rounding
            assert abs(v) < 1e-6, (name, d, v)


# ---------------------------------------------------------------- cellular
@pytest.mark.parametrize("d", [2, 3])
def test_cellular_jitter0_is_lattice_distance(d):
    pos = rand_points(d, 3000, -50, 50)
    t = oracle.Tree(S("CellularDistance", feature_scale=1.0, grid_jitter=0.0, distance_function="Euclidean",
                      return_type="Index0").encode())
    out, _ = t.positions(pos, seed=1)
    p = np.stack([np.asarray(c, np.float64) for c in pos])
    dist = np.sqrt(((p - np.round(p)) ** 2).sum(axis=0))
    jc = {2: 0.437016, 3: 0.396144}[d]
    max_dist = (1 + np.float32(jc)) ** 2
    assert np.abs(out - scale_output(dist, 0.0, max_dist)).max() < 1e-5


def test_cellular_lookup_constant():
    for d in (2, 3, 4):
        pos = rand_points(d, 500)
        t = oracle.Tree(S("CellularLookup", feature_scale=7.0, lookup=S("Constant", value=0.625)).encode())
        out, _ = t.positions(pos, seed=3)
        assert (out == np.float32(0.625)).all()


# ---------------------------------------------------------------- node identities
SRC = S("SuperSimplex", feature_scale=31.0, seed_offset=2)


def _pos_eval(spec, pos, seed=77):
    return oracle.Tree(spec.encode()).positions(pos, seed=seed)[0]


@pytest.mark.parametrize("d", [2, 3, 4])
def test_fractal_identities(d):
    pos = rand_points(d, 2000)
    src = _pos_eval(SRC, pos)
    for fr in ("FractalFBm", "FractalRidged"):
        one = _pos_eval(S(fr, source=SRC, octaves=1), pos)
        if fr == "FractalFBm":
            assert_bitexact(one, src, "FBm 1 octave")
        else:
            assert_close(one, 1 - 2 * np.abs(src), 1e-6, "Ridged 1 octave")
    assert_bitexact(_pos_eval(S("FractalFBm", source=SRC, octaves=4, gain=0.0), pos), src, "FBm gain 0")


@pytest.mark.parametrize("d", [2, 3, 4])
def test_zero_amplitude_warp_is_identity(d):
    pos = rand_points(d, 2000)
    src = _pos_eval(SRC, pos)
    for w in ("DomainWarpGradient", "DomainWarpSimplex", "DomainWarpSuperSimplex"):
        assert_bitexact(_pos_eval(S(w, source=SRC, warp_amplitude=0.0), pos), src, w)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_blend_closed_forms(d):
    pos = rand_points(d, 2000)
    A = S("Perlin", feature_scale=17.0)
    B = S("Value", feature_scale=29.0, seed_offset=5)
    a = _pos_eval(A, pos)
    b = _pos_eval(B, pos)
    f32 = np.float32
    assert_bitexact(_pos_eval(S("Add", lhs=A, rhs=B), pos), a + b, "Add")
    assert_bitexact(_pos_eval(S("Subtract", lhs=A, rhs=B), pos), a - b, "Subtract")
    assert_bitexact(_pos_eval(S("Multiply", lhs=A, rhs=B), pos), a * b, "Multiply")
    assert_bitexact(_pos_eval(S("Divide", lhs=A, rhs=B), pos), a / b, "Divide")
    assert_bitexact(_pos_eval(S("Min", lhs=A, rhs=B), pos), np.minimum(a, b), "Min")
    assert_bitexact(_pos_eval(S("Max", lhs=A, rhs=B), pos), np.maximum(a, b), "Max")
    assert_bitexact(_pos_eval(S("Abs", source=A), pos), np.abs(a), "Abs")
    assert_bitexact(_pos_eval(S("SignedSquareRoot", source=A), pos), np.sign(a) * np.sqrt(np.abs(a)), "SSQRT")
    assert_bitexact(_pos_eval(S("PowInt", value=A, pow=3), pos), a * a * a, "PowInt 3")
    assert_bitexact(_pos_eval(S("Add", lhs=A, rhs=0.25), pos), a + f32(0.25), "Add const")
    assert_close(_pos_eval(S("Remap", source=A, from_min=-0.5, from_max=0.5, to_min=2.0, to_max=-4.0), pos),
                 (a.astype(np.float64) + 0.5) / 1.0 * -6.0 + 2.0, 1e-5, "Remap")
    assert_close(_pos_eval(S("Fade", a=A, b=B, fade=0.25), pos),
                 a + (b.astype(np.float64) - a) * 0.625, 1e-6, "Fade linear")
    terr = _pos_eval(S("Terrace", source=A, step_count=4.0), pos)
    assert_close(terr, np.round(a.astype(np.float64) * 4) / 4, 1e-6, "Terrace")
    assert_close(_pos_eval(S("Modulus", lhs=A, rhs=0.3), pos), np.fmod(a, f32(0.3)), 1e-7, "Modulus")


@pytest.mark.parametrize("d", [2, 3, 4])
def test_domain_modifiers_reposition(d):
    pos = rand_points(d, 2000)
    f32 = np.float32
    scaled = [p * f32(1.75) for p in pos]
    assert_bitexact(_pos_eval(S("DomainScale", source=SRC, scaling=1.75), pos), _pos_eval(SRC, scaled), "DomainScale")
    offs = [3.5, -1.25, 0.5, 7.0][:d]
    moved = [p + f32(o) for p, o in zip(pos, offs)]
    spec = S("DomainOffset", source=SRC, **{"offset_" + "xyzw"[a]: offs[a] for a in range(d)})
    assert_bitexact(_pos_eval(spec, pos), _pos_eval(SRC, moved), "DomainOffset")
    sc = [0.5, 2.0, 1.5, 3.0][:d]
    spec = S("DomainAxisScale", source=SRC, **{"scaling_" + "xyzw"[a]: sc[a] for a in range(d)})
    assert_bitexact(_pos_eval(spec, pos), _pos_eval(SRC, [p * f32(s) for p, s in zip(pos, sc)]), "DomainAxisScale")
    assert_bitexact(_pos_eval(S("SeedOffset", source=SRC, seed_offset=9), pos), _pos_eval(SRC, pos, seed=77 + 9), "SeedOffset")
    assert_bitexact(_pos_eval(S("GeneratorCache", source=SRC), pos), _pos_eval(SRC, pos), "GeneratorCache")
    if d < 4:
        extra = np.full_like(pos[0], f32(2.5))
        assert_bitexact(_pos_eval(S("AddDimension", source=SRC, new_dimension_position=2.5), pos),
                        _pos_eval(SRC, pos + [extra]), "AddDimension")
    if d > 2:
        for ax in range(d):
            rest = [p for a, p in enumerate(pos) if a != ax]
            assert_bitexact(_pos_eval(S("RemoveDimension", source=SRC, remove_dimension="XYZW"[ax]), pos),
                            _pos_eval(SRC, rest), "RemoveDimension %d" % ax)


def test_domain_rotate_yaw_2d():
    # the reference's SetYaw takes radians (Modifiers.h:84): x' = x cos - y sin, y' = x sin + y cos
    pos = rand_points(2, 2000)
    yaw = 0.6
    c, s = math.cos(yaw), math.sin(yaw)
    x, y = (p.astype(np.float64) for p in pos)
    rot = [(x * c - y * s).astype(np.float32), (x * s + y * c).astype(np.float32)]
    a = _pos_eval(S("DomainRotate", source=S("Value", feature_scale=40.0), yaw=yaw), pos)
    b = _pos_eval(S("Value", feature_scale=40.0), rot)
    # rotation matrix entries are float: positions agree to ~1e-5 relative, Value noise is smooth
    assert np.abs(a - b).max() < 1e-3


# ---------------------------------------------------------------- drivers
def test_lanes_and_threads_do_not_change_results():
    for spec, d in ((S("FractalFBm", source=S("Simplex", feature_scale=40.0), octaves=4), 2),
                    (S("DomainWarpGradient", source=S("CellularDistance", feature_scale=12.0), feature_scale=30.0), 3),
                    (S("CellularValue", feature_scale=9.0), 4)):
        t = oracle.Tree(spec.encode())
        off = [-13.5, 7.25, 3.0, -2.0][:d]
        cnt = [37, 29, 11, 5][:d]
        a, mma = t.grid(off, cnt, seed=5, lanes=1, threads=1)
        b, mmb = t.grid(off, cnt, seed=5, lanes=16, threads=1)
        c, mmc = t.grid(off, cnt, seed=5, lanes=16, threads=7)
        assert_bitexact(a, b, "lanes")
        assert_bitexact(a, c, "threads")
        assert mma == mmb == mmc == (float(a.min()), float(a.max()))


def test_grid_equals_position_arrays():
    t = oracle.Tree(S("Simplex", feature_scale=20.0).encode())
    off, cnt, step = [-10.5, 3.25, 1.0], [13, 11, 7], [0.75, 1.5, 2.0]
    g, _ = t.grid(off, cnt, step, seed=9)
    idx = np.arange(int(np.prod(cnt)))
    pos = []
    for a in range(3):
        i = idx % cnt[a]
        idx = idx // cnt[a]
        pos.append((i * np.float64(np.float32(step[a])) + np.float32(off[a])).astype(np.float32))
    p, _ = t.positions(pos, seed=9)
    assert_bitexact(g, p, "grid vs positions")


def test_tileable_torus_coordinates_and_seam():
    spec = S("Simplex", feature_scale=30.0)
    t = oracle.Tree(spec.encode())
    nx, ny = 48, 40
    out, _ = t.tileable(nx, ny, 1.0, 1.0, seed=4)
    f32 = np.float32
    xsp = f32(nx) * f32(0.15915493667)
    ysp = f32(ny) * f32(0.15915493667)
    xi = np.tile(np.arange(nx), ny)
    yi = np.repeat(np.arange(ny), nx)
    xf = (xi.astype(np.float32) * f32(f32(1) / xsp)).astype(np.float32)
    yf = (yi.astype(np.float32) * f32(f32(1) / ysp)).astype(np.float32)
    pos = [np.cos(xf.astype(np.float64)) * xsp, np.cos(yf.astype(np.float64)) * ysp,
           np.sin(xf.astype(np.float64)) * xsp, np.sin(yf.astype(np.float64)) * ysp]
    ref, _ = t.positions([p.astype(np.float32) for p in pos], seed=4)
    assert np.abs(out - ref).max() < 1e-5
    img = out.reshape(ny, nx)
    # wrapping: column nx of the same torus is column 0 again (up to the rounding
    # of the reference's 1/(2 pi) constant 0.15915493667)
    xf_n = np.float32(nx) * f32(f32(1) / xsp)
    col_n = [np.full(ny, np.cos(np.float64(xf_n)) * xsp, np.float32), pos[1][::nx].astype(np.float32),
             np.full(ny, np.sin(np.float64(xf_n)) * xsp, np.float32), pos[3][::nx].astype(np.float32)]
    wrapped, _ = t.positions(col_n, seed=4)
    assert np.abs(wrapped - img[:, 0]).max() < 1e-4


def test_empty_and_single_point_calls():
    t = oracle.Tree(S("Perlin").encode())
    out, mm = t.grid([0, 0], [0, 5], seed=1)
    assert out.size == 0 and mm == (math.inf, -math.inf)
    out, mm = t.grid([1.5, 2.5], [1, 1], seed=1)
    assert mm == (float(out[0]), float(out[0]))
    assert t.single(1.5, 2.5, seed=1) == out[0]
