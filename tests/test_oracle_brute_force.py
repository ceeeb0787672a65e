"""Pins the oracle's lattice, cellular, fractal and warp nodes against float64
restatements written from the nodes' definitions (tests/f64ref.py), not from
their traversal:

* Simplex and SuperSimplex 2D/3D/4D as kernel sums over lattice points found
  by enumeration: SuperSimplex over every lattice point within the kernel
  radius, Simplex over the vertices of the Kuhn simplex that holds the point
  (Simplex.inl:7-936). The enumeration also shows which vertex set each node
  covers: for Simplex 3D/4D (radius^2 0.6) lattice points outside the simplex
  fall inside the radius, and the node does not sum them;
* CellularValue / CellularDistance / CellularLookup as a search over the 3^D
  candidates around the nearest lattice point, for every distance function,
  return type and index pair, jitter != 1 and size jitter, in 2D/3D/4D
  (Cellular.inl:17-807), selection exact and values to 1e-6;
* FractalFBm / FractalRidged as a bit-exact sum of separately evaluated
  octaves (seed + i, position * 2^i, amplitude 2^-i) and, with weighted
  strength and a per-point gain, a float64 recurrence (Fractal.inl:9-67);
* DomainWarpGradient's displacement (DomainWarp.inl:33-206) and the
  Progressive / Independent warp fractals (DomainWarpFractal.inl:4-74), seen
  through a Gradient node that returns one coordinate of the warped position.

Each restatement is also shown to reject plausible slips (a wrong skew or
seed, a swapped bit field, the other vertex set), so that a mistake of that
kind in the oracle fails here.
"""
import numpy as np
import pytest

import oracle
import f64ref as R
from fastnoise2_amd.tree import NodeSpec as S
from conftest import assert_bitexact, assert_close

MODES = [oracle.FUSED, oracle.UNFUSED, oracle.AVX2]


def points(d, n, lo, hi, seed):
    rng = np.random.default_rng(seed * 10 + d)
    return [rng.uniform(lo, hi, n).astype(np.float32) for _ in range(d)]


def oracle_positions(spec, pos, seed, mode=oracle.FUSED):
    return oracle.Tree(spec.encode()).positions(pos, seed=seed, mode=mode)[0]


# ---------------------------------------------------------------- simplex lattices
@pytest.mark.parametrize("name", ["Simplex", "SuperSimplex"])
@pytest.mark.parametrize("d", [2, 3, 4])
def test_simplex_kernel_sum(name, d):
    # scaled positions within +-8 lattice units: float32 rounding of the
    # skewed coordinates stays below 1e-6 there
    pos = points(d, 4000, -100.0, 100.0, 1)
    freq = 1 / 13.0
    ref = R.simplex_ref(name, pos, 1337 + 5, freq)
    for mode in MODES:
        out = oracle_positions(S(name, feature_scale=13.0, seed_offset=5), pos, 1337, mode)
        err = np.abs(out - ref).max()
        assert err < 1e-5, (name, d, mode, err)


@pytest.mark.parametrize("name,d", [("Simplex", 3), ("Simplex", 4), ("SuperSimplex", 4)])
def test_vertex_sets_differ_from_the_full_radius(name, d):
    # Simplex 3D/4D (radius^2 0.6) and SuperSimplex 4D leave out lattice
    # points inside the kernel radius: the node's own vertex set is what
    # matches, and the check tells the two apart
    pos = points(d, 20000, -100.0, 100.0, 2)
    out = oracle_positions(S(name, feature_scale=13.0), pos, 7)
    assert np.abs(out - R.simplex_ref(name, pos, 7, 1 / 13.0)).max() < 1e-5
    assert np.abs(out - R.simplex_ref(name, pos, 7, 1 / 13.0, vertex_set="radius")).max() > 2e-5


@pytest.mark.parametrize("d", [2, 3, 4])
def test_simplex_check_rejects_plausible_slips(d):
    pos = points(d, 2000, -100.0, 100.0, 3)
    out = oracle_positions(S("SuperSimplex", feature_scale=13.0), pos, 9)
    good = R.simplex_ref("SuperSimplex", pos, 9, 1 / 13.0)
    assert np.abs(out - good).max() < 1e-5
    # seed off by one
    assert np.abs(out - R.simplex_ref("SuperSimplex", pos, 8, 1 / 13.0)).max() > 1e-2
    # skew constant of the wrong dimensionality
    saved = dict(R.LATTICE[d])
    try:
        R.LATTICE[d]["F"] = R.LATTICE[{2: 3, 3: 4, 4: 2}[d]]["F"]
        assert np.abs(out - R.simplex_ref("SuperSimplex", pos, 9, 1 / 13.0)).max() > 1e-2
    finally:
        R.LATTICE[d].update(saved)
    # the Simplex vertex set on the SuperSimplex lattice
    assert np.abs(out - R.simplex_ref("SuperSimplex", pos, 9, 1 / 13.0, vertex_set="kuhn")).max() > 1e-3
    # the 2D gradient table on the 4D lattice's gradients (or the reverse)
    saved = R.GRAD[d]
    try:
        R.GRAD[d] = tuple(t[::-1].copy() for t in saved)
        assert np.abs(out - R.simplex_ref("SuperSimplex", pos, 9, 1 / 13.0)).max() > 1e-2
    finally:
        R.GRAD[d] = saved


# ---------------------------------------------------------------- cellular
def _cell_common(d):
    return points(d, 3000, -60.0, 60.0, 4)


@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("func", R.DIST_FUNCS)
def test_cellular_value_selection(d, func):
    pos = _cell_common(d)
    for value_index, jitter, size_jitter in ((0, 1.0, 0.0), (1, 0.8, 0.7), (3, 1.3, 0.0)):
        spec = S("CellularValue", feature_scale=7.0, distance_function=func, value_index=value_index,
                 grid_jitter=jitter, size_jitter=size_jitter, seed_offset=3)
        out = oracle_positions(spec, pos, 99)
        ref, margin = R.cellular_value_ref(pos, 99 + 3, 1 / 7.0, value_index=value_index, jitter=jitter,
                                           func=func, size_jitter=size_jitter)
        clear = margin > 1e-4
        assert clear.mean() > 0.99, (d, func, value_index, clear.mean())
        bad = np.abs(out - ref)[clear] > 1e-6
        assert not bad.any(), (d, func, value_index, int(bad.sum()))


@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("func", R.DIST_FUNCS)
def test_cellular_distance_all_return_types(d, func):
    pos = _cell_common(d)
    for rt in R.RETURN_TYPES:
        for i0, i1, jitter, size_jitter in ((0, 1, 1.0, 0.0), (1, 3, 1.2, 0.5), (2, 0, 0.6, 0.0)):
            spec = S("CellularDistance", feature_scale=7.0, distance_function=func, distance_index_0=i0,
                     distance_index_1=i1, return_type=rt, grid_jitter=jitter, size_jitter=size_jitter)
            out = oracle_positions(spec, pos, 5)
            ref = R.cellular_distance_ref(pos, 5, 1 / 7.0, index0=i0, index1=i1, return_type=rt, func=func,
                                          jitter=jitter, size_jitter=size_jitter)
            # Index0Div1 / Index0Mul1 are not bounded: relative to the value there
            assert_close(out, ref, 1e-5, "%dD %s %s %d/%d" % (d, func, rt, i0, i1), rel=rt in ("Index0Div1", "Index0Mul1"))


@pytest.mark.parametrize("d", [2, 3, 4])
def test_cellular_lookup_nearest_cell(d):
    pos = _cell_common(d)
    for func, jitter in (("Euclidean", 0.9), ("Manhattan", 1.0), ("MaxAxis", 0.5)):
        spec = S("CellularLookup", feature_scale=7.0, seed_offset=2, grid_jitter=jitter, distance_function=func,
                 lookup=S("Value", feature_scale=3.0))
        out = oracle_positions(spec, pos, 5)
        cells, margin = R.cellular_lookup_cells(pos, 5 + 2, 1 / 7.0, 7.0, jitter=jitter, func=func)
        # the lookup runs at seed + 1 of the call (without the node's offset)
        ref = R.value_ref([c.astype(np.float32) for c in cells], 5 + 1, 1 / 3.0)
        clear = margin > 1e-4
        assert clear.mean() > 0.99
        assert np.abs(out - ref)[clear].max() < 1e-5, (d, func)


def test_cellular_check_rejects_plausible_slips():
    pos = _cell_common(3)
    out = oracle_positions(S("CellularValue", feature_scale=7.0, grid_jitter=0.8, value_index=1), pos, 5)
    ref, margin = R.cellular_value_ref(pos, 5, 1 / 7.0, value_index=1, jitter=0.8)
    assert np.abs(out - ref)[margin > 1e-4].max() < 1e-6
    # jitter bit fields of the 3D vector in another order
    saved = R.jitter_vector
    try:
        R.jitter_vector = lambda d, h: saved(d, h)[::-1]
        swapped, _ = R.cellular_value_ref(pos, 5, 1 / 7.0, value_index=1, jitter=0.8)
    finally:
        R.jitter_vector = saved
    assert (np.abs(out - swapped) > 1e-6).mean() > 0.05
    # nearest cell instead of the second nearest
    first, _ = R.cellular_value_ref(pos, 5, 1 / 7.0, value_index=0, jitter=0.8)
    assert (np.abs(out - first) > 1e-6).mean() > 0.5


# ---------------------------------------------------------------- fractals
SOURCES = {
    "Perlin": S("Perlin", feature_scale=23.0),
    "Simplex": S("Simplex", feature_scale=19.0, seed_offset=2),
    "CellularDistance": S("CellularDistance", feature_scale=11.0, return_type="Index0Add1"),
}


@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("src", sorted(SOURCES))
def test_fractal_octave_decomposition_bitexact(d, src):
    # gain 0.5, lacunarity 2, weighted strength 0: octave i is the source at
    # seed + i and position * 2^i with amplitude 2^-i (all exact), summed in
    # float in octave order
    pos = points(d, 2000, -200.0, 200.0, 5)
    seed = 31
    for mode in (oracle.FUSED, oracle.UNFUSED):
        n = [oracle_positions(SOURCES[src], [p * np.float32(2.0 ** i) for p in pos], seed + i, mode) for i in range(5)]
        fbm = n[0].copy()
        ridged = (np.float32(1) - np.abs(n[0]) * np.float32(2)).astype(np.float32)
        for i in range(1, 5):
            amp = np.float32(2.0 ** -i)
            fbm = (fbm + n[i] * amp).astype(np.float32)
            ridged = (ridged + (np.float32(1) - np.abs(n[i]) * np.float32(2)) * amp).astype(np.float32)
        got = oracle_positions(S("FractalFBm", source=SOURCES[src], octaves=5, gain=0.5, lacunarity=2.0), pos, seed, mode)
        assert_bitexact(got, fbm, "FBm %s %dD" % (src, d))
        got = oracle_positions(S("FractalRidged", source=SOURCES[src], octaves=5, gain=0.5, lacunarity=2.0), pos, seed, mode)
        assert_bitexact(got, ridged, "Ridged %s %dD" % (src, d))


@pytest.mark.parametrize("d", [2, 3, 4])
def test_fractal_weighted_strength_and_node_gain(d):
    # amp_{i+1} = amp_i * lerp(1, (n_i + 1) / 2, ws) * gain for FBm and
    # amp_i * lerp(1, 1 - |n_i|, ws) * gain for Ridged, with the gain a node
    # evaluated once at the call's seed and position
    pos = points(d, 2000, -200.0, 200.0, 6)
    seed, oct_, lac, ws = 4, 4, 1.75, 0.65
    src = SOURCES["Simplex"]
    gain_node = S("Value", feature_scale=40.0)
    gain = oracle_positions(gain_node, pos, seed).astype(np.float64)
    n, q = [], [p for p in pos]
    for i in range(oct_):
        n.append(oracle_positions(src, q, seed + i).astype(np.float64))
        q = [(c * np.float32(lac)).astype(np.float32) for c in q]
    amp, fbm = 1.0, n[0].copy()
    for i in range(1, oct_):
        amp = amp * R.lerp(1.0, (n[i - 1] + 1) * 0.5, ws) * gain
        fbm += n[i] * amp
    amp, ridged = 1.0, 1 - 2 * np.abs(n[0])
    for i in range(1, oct_):
        amp = amp * R.lerp(1.0, 1 - np.abs(n[i - 1]), ws) * gain
        ridged += (1 - 2 * np.abs(n[i])) * amp
    got = oracle_positions(S("FractalFBm", source=src, octaves=oct_, gain=gain_node, lacunarity=lac, weighted_strength=ws), pos, seed)
    assert np.abs(got - fbm).max() < 1e-5
    got = oracle_positions(S("FractalRidged", source=src, octaves=oct_, gain=gain_node, lacunarity=lac, weighted_strength=ws), pos, seed)
    assert np.abs(got - ridged).max() < 1e-5


def test_fractal_check_rejects_seed_slip():
    pos = points(2, 1000, -200.0, 200.0, 7)
    src = SOURCES["Perlin"]
    n = [oracle_positions(src, [p * np.float32(2.0 ** i) for p in pos], 10 - i) for i in range(2)]
    wrong = (n[0] + n[1] * np.float32(0.5)).astype(np.float32)
    got = oracle_positions(S("FractalFBm", source=src, octaves=2, gain=0.5, lacunarity=2.0), pos, 10)
    assert (got != wrong).mean() > 0.5


# ---------------------------------------------------------------- warps
def _probe(axis):
    """Gradient node that returns coordinate `axis` of its position exactly."""
    return S("Gradient", **{"multiplier_" + "xyzw"[a]: (1.0 if a == axis else 0.0) for a in range(4)})


def test_gradient_probe_returns_the_coordinate():
    for d in (2, 3, 4):
        pos = points(d, 500, -50.0, 50.0, 8)
        for a in range(d):
            assert_bitexact(oracle_positions(_probe(a), pos, 1), pos[a], "probe axis %d" % a)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_domain_warp_gradient_displacement(d):
    # x' = x + amplitude * axis scale * w(seed + 2 * seed offset, x * frequency):
    # a plain warp node adds its seed offset once in GenT and again in Warp
    pos = points(d, 3000, -40.0, 40.0, 9)
    sc = [1.0, 0.5, 2.0, 1.5][:d]
    vec, _ = R.warp_gradient_vector(pos, 11 + 8, 1 / 6.0)
    for a in range(d):
        spec = S("DomainWarpGradient", feature_scale=6.0, warp_amplitude=3.0, seed_offset=4, source=_probe(a),
                 **{"amplitude_scaling_" + "xyzw"[k]: sc[k] for k in range(d)})
        for mode in (oracle.FUSED, oracle.UNFUSED):
            got = oracle_positions(spec, pos, 11, mode).astype(np.float64) - pos[a]
            assert np.abs(got - 3.0 * sc[a] * vec[a]).max() < 1e-5, (d, a, mode)
    # a single seed offset would not match
    wrong, _ = R.warp_gradient_vector(pos, 11 + 4, 1 / 6.0)
    got = oracle_positions(S("DomainWarpGradient", feature_scale=6.0, warp_amplitude=3.0, seed_offset=4, source=_probe(0)),
                           pos, 11).astype(np.float64) - pos[0]
    assert np.abs(got - 3.0 * wrong[0]).max() > 0.1


@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("kind", ["DomainWarpFractalProgressive", "DomainWarpFractalIndependent"])
def test_warp_fractal_displacement(d, kind):
    pos = points(d, 2000, -10.0, 10.0, 10)
    sc = [1.0, 0.5, 2.0, 1.5][:d]
    indep = kind.endswith("Independent")
    for ws in (0.0, 0.6):
        disp = R.warp_fractal_displacement(pos, 21, 4, 0.6, 2.0, ws, 2.0, 9.0, sc, 3, indep)
        for a in range(d):
            warp = S("DomainWarpGradient", feature_scale=9.0, warp_amplitude=2.0, seed_offset=3, source=_probe(a),
                     **{"amplitude_scaling_" + "xyzw"[k]: sc[k] for k in range(d)})
            spec = S(kind, domain_warp_source=warp, octaves=4, gain=0.6, lacunarity=2.0, weighted_strength=ws)
            got = oracle_positions(spec, pos, 21).astype(np.float64) - pos[a]
            assert np.abs(got - disp[a]).max() < 2e-5, (d, kind, ws, a)
    # the other kind's sampling positions do not match
    other = R.warp_fractal_displacement(pos, 21, 4, 0.6, 2.0, 0.0, 2.0, 9.0, sc, 3, not indep)
    warp = S("DomainWarpGradient", feature_scale=9.0, warp_amplitude=2.0, seed_offset=3, source=_probe(0),
             **{"amplitude_scaling_" + "xyzw"[k]: sc[k] for k in range(d)})
    got = oracle_positions(S(kind, domain_warp_source=warp, octaves=4, gain=0.6, lacunarity=2.0), pos, 21).astype(np.float64) - pos[0]
    assert np.abs(got - other[0]).max() > 1e-3


# ---------------------------------------------------------------- simplex-type warps
def _probe_vectors(d, scheme):
    """The oracle's one-vertex warp vector helper as a vectorised callable
    (hashes, offsets per axis, multipliers) -> vectors per axis."""
    L = oracle.load()

    def helper(hashes, f, m):
        hashes = np.asarray(hashes)
        F = np.ascontiguousarray(np.stack(f, 1).astype(np.float32))
        M = np.broadcast_to(np.asarray(m, np.float32), hashes.shape)
        out = np.zeros((hashes.shape[0], d))
        v = np.zeros(d, np.float32)
        for k in range(hashes.shape[0]):
            if M[k] == 0:
                continue
            v[:] = 0
            fk = np.ascontiguousarray(F[k])
            L.fn2o_warp_vec(d, scheme, int(hashes[k]), fk.ctypes.data_as(oracle._fp), float(M[k]),
                            v.ctypes.data_as(oracle._fp), 0)
            out[k] = v
        return [out[:, a] for a in range(d)]
    return helper


@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("scheme", [0, 1])
def test_warp_vector_helpers_match_table_branches(d, scheme):
    # the oracle follows the generic branches of the helpers; the table
    # (AVX-512) branches are an independent formulation of the same vectors
    # wherever the table indices stay inside their laid-out ranges
    rng = np.random.default_rng(40 + d * 2 + scheme)
    n = 6000
    hashes = rng.integers(-2 ** 31, 2 ** 31, n)
    f = rng.uniform(-1, 1, (n, d))
    m = rng.uniform(0.1, 2.0, n)
    got = np.stack(_probe_vectors(d, scheme)(hashes, [f[:, a] for a in range(d)], m), 1)
    ref = np.stack(R.WARP_VECTOR[(d, scheme)](hashes, [f[:, a].astype(np.float32).astype(np.float64) for a in range(d)],
                                              m.astype(np.float32).astype(np.float64)), 1)
    dom = R.warp_table_domain(d, scheme, hashes)
    assert dom.mean() > 0.35
    assert np.abs(got - ref)[dom].max() < 1e-5, (d, scheme)
    # outside the domain the reference's branches disagree with each other;
    # the difference is a different vector, not rounding
    assert (np.abs(got - ref).max(1)[~dom] > 1e-3).mean() > 0.5


@pytest.mark.parametrize("name", ["DomainWarpSimplex", "DomainWarpSuperSimplex"])
@pytest.mark.parametrize("d", [2, 3, 4])
@pytest.mark.parametrize("scheme", [0, 1])
def test_simplex_warp_lattice_sum(name, d, scheme):
    # displacement = amplitude * axis scale * bounding * sum over the node's
    # lattice vertices of kernel * vector(hash, offset), the per-vertex vector
    # from the oracle's (separately pinned) helper
    pos = points(d, 1200, -30.0, 30.0, 11)
    sc = [1.0, 0.5, 2.0, 1.5][:d]
    # DomainWarpSimplex adds its seed offset twice, DomainWarpSuperSimplex once
    warp_seed = 11 + (6 if name == "DomainWarpSimplex" else 3)
    vec, _ = R.warp_simplex_vector(name[len("DomainWarp"):], pos, warp_seed, 1 / 7.0, scheme, _probe_vectors(d, scheme))
    for a in range(d):
        spec = S(name, feature_scale=7.0, warp_amplitude=1.5, seed_offset=3, vectorization_scheme=scheme,
                 source=_probe(a), **{"amplitude_scaling_" + "xyzw"[k]: sc[k] for k in range(d)})
        got = oracle_positions(spec, pos, 11).astype(np.float64) - pos[a]
        assert np.abs(got - 1.5 * sc[a] * vec[a]).max() < 1e-5, (name, d, scheme, a)
