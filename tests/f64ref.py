"""Float64 restatements of FastNoise2's node math, written from the
mathematical definitions rather than from the oracle's step-by-step
traversal, for the tests that pin the oracle (tests/test_oracle_*.py).

Each function evaluates a node the way its definition reads:

* lattice noise as a kernel sum: every lattice vertex near the point that the
  node's vertex set covers contributes falloff(|d|^2) * gradient(hash) . d,
  found by enumerating a neighbourhood of lattice points, not by the node's
  simplex traversal and its incremental offsets;
* cellular nodes as a search over the 3^D candidate cells around the
  point's nearest lattice point, sorted by distance (ties to the earlier
  candidate in x-major order);
* warps as an explicit displacement of the position.

The shared constants (primes, hash multiplier, gradient tables) are the
reference's (include/FastNoise/Generators/Utils.inl:6-19); the gradient
tables are the ones its AVX2 / AVX-512 branches hold, which
tests/test_oracle_closed_forms.py pins the oracle's generic branch against.
Positions are scaled by the node frequency in float32, as ScalePositions does
(BasicGenerators.inl:16-20), and everything after that in float64.
"""
import itertools
import math

import numpy as np

PX, PY, PZ, PW = 0xF797C5C7, 0x6C060C89, 0x465FD04F, 0xF7A62279
PRIMES = (PX, PY, PZ, PW)
HASH_A = 0xB7E0A5F5
f32 = np.float32

R2F = float(f32(math.sqrt(2.0)))
P2 = float(f32(f32(1.0) + f32(R2F)))   # 1 + kRoot2f in float

# GetGradientDotPerlin 2D, AVX2 table (Utils.inl:174-180)
G2X = np.array([P2, -P2, P2, -P2, 1, -1, 1, -1], np.float64)
G2Y = np.array([1, 1, -1, -1, P2, P2, -P2, -P2], np.float64)
# GetGradientDotCommon 3D, AVX-512 table (Utils.inl:35-37)
G3X = np.array([1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, 0, -1, 0], np.float64)
G3Y = np.array([1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1], np.float64)
G3Z = np.array([0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 1, 0, -1], np.float64)
# GetGradientDotPerlin 4D, AVX-512 two-table permute (Utils.inl:191-196)
G4X = np.array([0] * 8 + [1, -1] * 4 + [1, -1] * 8, np.float64)
G4Y = np.array([1, -1] * 4 + [0] * 8 + [1, 1, -1, -1] * 4, np.float64)
G4Z = np.array([1, 1, -1, -1] * 4 + [0] * 8 + [1, 1, 1, 1, -1, -1, -1, -1], np.float64)
G4W = np.array([1, 1, 1, 1, -1, -1, -1, -1] * 2 + [1, 1, 1, 1, -1, -1, -1, -1] + [0] * 8, np.float64)
GRAD = {2: (G2X, G2Y), 3: (G3X, G3Y, G3Z), 4: (G4X, G4Y, G4Z, G4W)}
GRAD_MASK = {2: 7, 3: 15, 4: 31}


def i32(v):
    v = np.asarray(v, dtype=np.int64) & 0xFFFFFFFF
    return np.where(v >= 1 << 31, v - (1 << 32), v)


def hash_primes(seed, *primed):
    """HashPrimes (Utils.inl:221-230): xor, multiply, arithmetic >> 15 folded in."""
    h = np.int64(seed)
    for p in primed:
        h = h ^ p
    h = i32(h * HASH_A)
    return i32((h >> 15) ^ h)


def hash_primes_hb(seed, *primed):
    """HashPrimesHB (Utils.inl:232-240): the same without the final fold."""
    h = np.int64(seed)
    for p in primed:
        h = h ^ p
    return i32(h * HASH_A)


def value_coord(seed, *primed):
    """GetValueCoord (Utils.inl:242-250) as a float64 of the int32 value."""
    h = np.int64(seed)
    for p in primed:
        h = h ^ p
    h = i32(h * i32(h * HASH_A))
    return h.astype(np.float64)


def primed(lattice):
    """int32 lattice coordinates times the per-axis primes (wrapping)."""
    return [i32(np.asarray(c, np.int64) * pr) for c, pr in zip(lattice, PRIMES)]


def grad_dot(d, hash_, offs):
    tabs = GRAD[d]
    h = np.asarray(hash_) & GRAD_MASK[d]
    return sum(tabs[a][h] * offs[a] for a in range(d))


def scale_pos(pos, freq):
    """ScalePositions: one float32 multiply per axis."""
    return [(np.asarray(c, np.float32) * f32(freq)).astype(np.float64) for c in pos]


def scale_output(v, nmin, nmax, rmin=-1.0, rmax=1.0):
    return (v - nmin) / (nmax - nmin) * (rmax - rmin) + rmin


def quintic(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


def hermite(t):
    return t * t * (3 - 2 * t)


def lerp(a, b, t):
    return a + t * (b - a)


# --------------------------------------------------------------- lattice noise
def perlin_ref(pos, seed, freq):
    """Perlin 2D/3D/4D (Perlin.inl) with table gradients."""
    d = len(pos)
    p = scale_pos(pos, freq)
    fl = [np.floor(c) for c in p]
    base = [i32(f.astype(np.int64) * pr) for f, pr in zip(fl, PRIMES)]
    fr = [c - f for c, f in zip(p, fl)]
    t = [quintic(f) for f in fr]
    vals = []
    for c in range(1 << d):
        prim = [i32(base[a] + (((c >> a) & 1) * PRIMES[a])) for a in range(d)]
        h = hash_primes(seed, *prim)
        vals.append(grad_dot(d, h, [fr[a] - ((c >> a) & 1) for a in range(d)]))
    for a in range(d):
        vals = [lerp(vals[2 * k], vals[2 * k + 1], t[a]) for k in range(len(vals) // 2)]
    bound = {2: 1.726796627044677734375, 3: 1.0363423824310302734375, 4: 1.33858621120452880859375}[d]
    return scale_output(vals[0], -bound, bound)


def value_ref(pos, seed, freq):
    """Value noise 2D/3D/4D (Value.inl)."""
    d = len(pos)
    p = scale_pos(pos, freq)
    fl = [np.floor(c) for c in p]
    base = [i32(f.astype(np.int64) * pr) for f, pr in zip(fl, PRIMES)]
    t = [hermite(c - f) for c, f in zip(p, fl)]
    vals = []
    for c in range(1 << d):
        prim = [i32(base[a] + (((c >> a) & 1) * PRIMES[a])) for a in range(d)]
        vals.append(value_coord(seed, *prim))
    for a in range(d):
        vals = [lerp(vals[2 * k], vals[2 * k + 1], t[a]) for k in range(len(vals) // 2)]
    return scale_output(vals[0], -2147483648.0, 2147483648.0)


# The simplex-type lattices: skew s(v) = v + F (v.1) 1 maps unskewed space to
# the integer lattice; a lattice point L lies at s^-1(L) = L - G (L.1) 1. The
# offset a vertex sees is U(s(p) - L) with U(v) = v - G' (v.1) 1: G' = G gives
# p - s^-1(L) (2D and 4D), G' = 1/2 in 3D gives that vector reflected through
# the plane normal to (1,1,1) (Simplex.inl:65-142, 379-575: kReflectUnskew3),
# which keeps its length.
LATTICE = {
    2: dict(F=1.0 / (math.sqrt(3.0) + 1.0), U=1.0 / (math.sqrt(3.0) + 3.0)),
    3: dict(F=1.0 / 3.0, U=0.5),
    4: dict(F=1.0 / (math.sqrt(5.0) + 1.0), U=1.0 / (math.sqrt(5.0) + 5.0)),
}

# Kernel radius^2, kernel scale and output bounding of each node and
# dimensionality (Simplex.inl: kFalloffRadiusSquared, the 0.5 factor of the
# SuperSimplex falloffs, kBounding).
SIMPLEX_KERNEL = {
    ("Simplex", 2): (0.5, 1.0, 0.0261208079755306243896484375),
    ("Simplex", 3): (0.6, 1.0, 0.030586399137973785400390625),
    ("Simplex", 4): (0.6, 1.0, 0.036730043590068817138671875),
    ("SuperSimplex", 2): (2.0 / 3.0, 1.0, 0.14084912836551666259765625),
    ("SuperSimplex", 3): (0.75, 0.5, 0.0069091119803488254547119140625),
    ("SuperSimplex", 4): (0.8, 0.5, 0.011310013942420482635498046875),
}


def kuhn_vertices(ds):
    """The vertices of the Kuhn simplex that holds skewed fractional offsets
    ds (per axis, in [0,1)): 0, then unit steps along the axes in order of
    decreasing offset, ties to the lower axis. One (N, D) int array per vertex."""
    d = len(ds)
    n = ds[0].shape[0]
    order = np.argsort(-np.stack(ds, axis=1), axis=1, kind="stable")
    verts = [np.zeros((n, d), np.int64)]
    for k in range(d):
        v = verts[-1].copy()
        v[np.arange(n), order[:, k]] += 1
        verts.append(v)
    return verts


def lattice_points(p, d, vertex_set):
    """(vertex lattice coordinates, offsets) of the lattice points a simplex
    node sums over at scaled positions p. vertex_set:

    * "kuhn": the vertices of the Kuhn simplex that holds the point (Simplex);
    * "radius": every lattice point of the neighbourhood (the kernel decides
      which contribute: SuperSimplex 2D, whose four vertices cover its radius);
    * "pairs": the lattice point nearest to p, and for each nonzero offset o
      in {0,1}^D the nearer of nearest + o and nearest - o (SuperSimplex 3D
      and 4D: 8 and 16 vertices; in 4D this leaves out, near the kernel's
      edge, a few lattice points inside its radius)."""
    F, U = LATTICE[d]["F"], LATTICE[d]["U"]
    n = p[0].shape[0]
    ssum = sum(p)
    s = [c + F * ssum for c in p]
    base = [np.floor(c).astype(np.int64) for c in s]

    def offset(L):
        v = [s[a] - L[a] for a in range(d)]
        vs = sum(v)
        return [c - U * vs for c in v]

    def length_sq(L):
        return sum(c * c for c in offset(L))

    if vertex_set == "kuhn":
        verts = [[base[a] + off[:, a] for a in range(d)] for off in kuhn_vertices([c - b for c, b in zip(s, base)])]
    else:
        verts = [[base[a] + o[a] for a in range(d)] for o in itertools.product(range(-1, 3), repeat=d)]
        if vertex_set == "pairs":
            near, near_q = verts[0], length_sq(verts[0])
            for L in verts[1:]:
                q = length_sq(L)
                m = q < near_q
                near = [np.where(m, L[a], near[a]) for a in range(d)]
                near_q = np.where(m, q, near_q)
            verts = [near]
            for o in itertools.product((0, 1), repeat=d):
                if not any(o):
                    continue
                up = [near[a] + o[a] for a in range(d)]
                down = [near[a] - o[a] for a in range(d)]
                m = length_sq(up) <= length_sq(down)
                verts.append([np.where(m, up[a], down[a]) for a in range(d)])
    return [(L, offset(L)) for L in verts]


VERTEX_SET = {("Simplex", 2): "kuhn", ("Simplex", 3): "kuhn", ("Simplex", 4): "kuhn",
              ("SuperSimplex", 2): "radius", ("SuperSimplex", 3): "pairs", ("SuperSimplex", 4): "pairs"}


def simplex_ref(name, pos, seed, freq, vertex_set=None):
    """Simplex / SuperSimplex 2D/3D/4D as a kernel sum over lattice points:
    sum of max(0, k (r^2 - |d|^2))^4 * gradient(HashPrimes(L)) . d."""
    d = len(pos)
    r2, ks, bound = SIMPLEX_KERNEL[(name, d)]
    p = scale_pos(pos, freq)
    total = np.zeros_like(p[0])
    for L, dv in lattice_points(p, d, vertex_set or VERTEX_SET[(name, d)]):
        q = sum(c * c for c in dv)
        fall = np.maximum(0.0, ks * (r2 - q)) ** 4
        g = grad_dot(d, hash_primes(seed, *primed(L)), dv)
        total += np.where(q < r2, fall * g, 0.0)
    return scale_output(total, -bound, bound)


# --------------------------------------------------------------- cellular
JITTER = {2: 0.437016, 3: 0.396144, 4: 0.366025}
DIST_FUNCS = ["Euclidean", "EuclideanSquared", "Manhattan", "Hybrid", "MaxAxis", "Minkowski"]


def jitter_vector(d, h):
    """The cell point's direction from a HashPrimesHB hash (Cellular.inl:57-59,
    141-143, 229-232): bit fields centred on zero."""
    u = np.asarray(h, np.int64) & 0xFFFFFFFF
    if d == 2:
        return [(u & 0x7FF) - 1023.5, (u >> 21) - 1023.5]
    if d == 3:
        return [(u & 0x3FF) - 511.5, ((u >> 11) & 0x3FF) - 511.5, (u >> 22) - 511.5]
    return [(u & 0xFF) - 127.5, ((u >> 8) & 0xFF) - 127.5, ((u >> 16) & 0xFF) - 127.5, (u >> 24) - 127.5]


def size_jitter_factor(d, h, size_jitter):
    """1 + size jitter * (hash bits / full scale) (Cellular.inl:65-69, 152-156, 244-248)."""
    u = np.asarray(h, np.int64) & 0xFFFFFFFF
    bits, full = {2: ((u >> 11) & 0x3FF, 0x3FF), 3: (u & 0xFFFF, 0xFFFF), 4: (u & 0xFFFFF, 0xFFFFF)}[d]
    return 1.0 + size_jitter * bits / full


def distance(func, delta, minkowski_p=1.5, root=False):
    """CalcDistance (Utils.inl:281-335); Euclidean is the squared length unless
    `root` (CellularValue 2D asks for the root)."""
    if func in ("Euclidean", "EuclideanSquared"):
        q = sum(c * c for c in delta)
        return np.sqrt(q) if (func == "Euclidean" and root) else q
    if func == "Manhattan":
        return sum(np.abs(c) for c in delta)
    if func == "Hybrid":
        return sum(c * c + np.abs(c) for c in delta)
    if func == "MaxAxis":
        return np.max(np.abs(np.stack(delta)), axis=0)
    return sum(np.abs(c) ** minkowski_p for c in delta) ** (1.0 / minkowski_p)


def cell_candidates(pos, seed, freq, jitter=1.0, func="Euclidean", size_jitter=0.0, root=False, minkowski_p=1.5):
    """The 3^D candidate cells around each point's nearest lattice point, in
    the reference's visiting order (x outermost): (distances (N, 3^D),
    hashes (N, 3^D), cell points [(N, 3^D)] * D in lattice units)."""
    d = len(pos)
    p = scale_pos(pos, freq)
    centre = [np.rint(c).astype(np.int64) for c in p]
    dists, hashes, points = [], [], [[] for _ in range(d)]
    for off in itertools.product((-1, 0, 1), repeat=d):
        L = [centre[a] + off[a] for a in range(d)]
        h = hash_primes_hb(seed, *primed(L))
        v = jitter_vector(d, h)
        inv = JITTER[d] * jitter / np.sqrt(sum(c * c for c in v))
        cp = [L[a] + v[a] * inv for a in range(d)]
        dist = distance(func, [cp[a] - p[a] for a in range(d)], minkowski_p, root)
        if size_jitter != 0.0:
            dist = dist * size_jitter_factor(d, h, size_jitter)
        dists.append(dist)
        hashes.append(h)
        for a in range(d):
            points[a].append(cp[a])
    return (np.stack(dists, axis=1), np.stack(hashes, axis=1), [np.stack(c, axis=1) for c in points])


def sorted_candidates(dists):
    """Candidate order by distance, ties to the earlier candidate."""
    return np.argsort(dists, axis=1, kind="stable")


def selection_margin(dists, order, k):
    """Relative gap around rank k of the sorted distances: points whose rank-k
    candidate is within float32 rounding of a neighbour in rank are ambiguous
    for a float64 check of a float32 selection."""
    s = np.take_along_axis(dists, order, axis=1)
    gaps = [np.abs(s[:, k] - s[:, k - 1])] if k > 0 else []
    gaps.append(np.abs(s[:, k + 1] - s[:, k]))
    return np.minimum.reduce(gaps) / np.maximum(np.abs(s[:, k]), 1e-30)


def cellular_value_ref(pos, seed, freq, value_index=0, **kw):
    """CellularValue (Cellular.inl:20-292): the hash of the value_index-th
    nearest cell, as float scaled from +-2^31; also the selection margin."""
    d = len(pos)
    dists, hashes, _ = cell_candidates(pos, seed, freq, root=(d == 2), **kw)
    order = sorted_candidates(dists)
    h = np.take_along_axis(hashes, order[:, value_index:value_index + 1], axis=1)[:, 0]
    return scale_output(h.astype(np.float64), -2147483648.0, 2147483648.0), selection_margin(dists, order, value_index)


RETURN_TYPES = ["Index0", "Index0Add1", "Index0Sub1", "Index0Mul1", "Index0Div1"]


def cellular_distance_ref(pos, seed, freq, index0=0, index1=1, return_type="Index0", func="Euclidean", **kw):
    """CellularDistance (Cellular.inl:297-566): the index0-th / index1-th
    smallest distances, rooted for Euclidean, combined and scaled by the
    largest distance (1 + jitter constant)^2."""
    d = len(pos)
    dists, _, _ = cell_candidates(pos, seed, freq, func=func, **kw)
    s = np.sort(dists, axis=1)
    a, b = s[:, index0], s[:, index1]
    if func == "Euclidean":
        a, b = np.sqrt(a), np.sqrt(b)
    m = (1.0 + f32(JITTER[d])) ** 2
    v, hi = {"Index0": (a, m), "Index0Add1": (a + b, 2 * m), "Index0Sub1": (np.abs(a - b), m),
             "Index0Mul1": (a * b, m * m), "Index0Div1": (a / b, m)}[return_type]
    return scale_output(v, 0.0, hi)


def cellular_lookup_cells(pos, seed, freq, scale, **kw):
    """CellularLookup (Cellular.inl:571-807): the nearest cell's point (first
    of equals) times the node's scale, where the lookup is evaluated at
    seed + 1; also the selection margin."""
    dists, _, points = cell_candidates(pos, seed, freq, **kw)
    order = sorted_candidates(dists)
    first = order[:, :1]
    cells = [np.take_along_axis(c, first, axis=1)[:, 0] * scale for c in points]
    return cells, selection_margin(dists, order, 0)


# --------------------------------------------------------------- warps
def warp_gradient_vector(pos, seed, freq):
    """DomainWarpGradient's unit-range warp vector at a position (DomainWarp.inl:
    33-206): Hermite-weighted sum over the 2^D corners of per-axis hash bit
    fields, mapped from [0, full] to [-1, 1]. Returns (vector, |vector|^2)."""
    d = len(pos)
    p = scale_pos(pos, freq)
    fl = [np.floor(c) for c in p]
    w = [hermite(c - f) for c, f in zip(p, fl)]
    L0 = [f.astype(np.int64) for f in fl]
    acc = [np.zeros_like(p[0]) for _ in range(d)]
    for c in range(1 << d):
        bits = [(c >> a) & 1 for a in range(d)]
        h = np.asarray(hash_primes_hb(seed, *primed([L0[a] + bits[a] for a in range(d)])), np.int64) & 0xFFFFFFFF
        wt = np.ones_like(p[0])
        for a in range(d):
            wt = wt * (w[a] if bits[a] else 1.0 - w[a])
        if d == 2:
            comps = [h & 0xFFFF, h >> 16]
        elif d == 3:
            comps = [h & 0x3FF, (h >> 11) & 0x3FF, h >> 22]
        else:
            comps = [h & 0xFF, (h >> 8) & 0xFF, (h >> 16) & 0xFF, h >> 24]
        for a in range(d):
            acc[a] += wt * comps[a]
    full = {2: 0xFFFF, 3: 0x3FF, 4: 0xFF}[d]
    vec = [a / (full / 2.0) - 1.0 for a in acc]
    return vec, sum(v * v for v in vec)


def warp_fractal_displacement(pos, seed, octaves, gain, lacunarity, weighted_strength, amplitude, scale,
                              axis_scale, seed_offset, independent, vector=warp_gradient_vector):
    """DomainWarpFractalProgressive / Independent over a DomainWarpGradient
    (DomainWarpFractal.inl:4-74): octave i warps at seed + i + the warp's
    seed offset, frequency (1/scale) * lacunarity^i, amplitude amp_i with
    amp_{i+1} = amp_i * (1 - weighted strength * |w_i|) * gain; Progressive
    samples the warp at the already warped position, Independent at the
    original one. Returns the total displacement per axis (float64)."""
    d = len(pos)
    orig = [np.asarray(c, np.float64) for c in pos]
    cur = [c.copy() for c in orig]
    freq = float(f32(1.0) / f32(scale))
    amp = amplitude
    for i in range(octaves):
        if i:
            freq = float(f32(freq) * f32(lacunarity))
        at = orig if independent else cur
        vec, strength_sq = vector([c.astype(np.float32) for c in at], seed + i + seed_offset, freq)
        if i:
            amp = amp * (1.0 - weighted_strength * prev_len) * gain
        cur = [cur[a] + vec[a] * amp * axis_scale[a] for a in range(d)]
        prev_len = np.sqrt(strength_sq)
    return [cur[a] - orig[a] for a in range(d)]


# --------------------------------------------------------------- simplex-warp vectors
# One vertex's contribution to the simplex-type warps' vector: multiplier *
# (vector chosen by the hash, applied to the offset f). Restated from the
# table (AVX-512) branches of DomainWarpSimplex.inl:6-724; the oracle follows
# the generic branches, so these pin its bit logic. Table constants are
# float32 values formed in float32 arithmetic as the reference's constexpr
# floats are.
def _u32(v):
    return np.asarray(v, np.int64) & 0xFFFFFFFF


def _s32(v):
    v = _u32(v)
    return np.where(v >= 1 << 31, v - (1 << 32), v)


def _ff(*xs):
    """Sum of float32 values in float32 arithmetic, left to right."""
    acc = f32(xs[0])
    for x in xs[1:]:
        acc = f32(acc + f32(x))
    return float(acc)


_SKEW2F = float(f32(1.0 / (math.sqrt(3.0) + 1.0)))
_SKEW4F = float(f32(1.0 / (math.sqrt(5.0) + 1.0)))
_ROOT3F = float(f32(math.sqrt(3.0)))
_ROOT2F = float(f32(math.sqrt(2.0)))


def _perm2(t1, t2, idx):
    """_mm512_permutex2var_ps: index bit 4 chooses the second table."""
    idx = np.asarray(idx) & 31
    tab = np.array(list(t1) + list(t2), np.float64)
    return tab[idx]


def _ortho2(h, f, m):
    s = _SKEW2F
    s1 = _ff(s, 1)
    tx = [s, -s, s, -s, s1, -s1, s1, -s1, 1, -1, 1, -1, 0, 0, 0, 0]
    ty = [s1, s1, -s1, -s1, s, s, -s, -s, 1, 1, -1, -1, 0, 0, 0, 0]
    idx = _u32((_u32(h) >> 1) * 3) >> 28
    gx, gy = np.array(tx)[idx], np.array(ty)[idx]
    sign = np.where(_u32(h) & 1, -1.0, 1.0)
    return [m * (f[1] * gy + f[0] * gx), m * sign * (f[1] * gx - f[0] * gy)]


def _outer_index(hs):
    iov = _u32(hs * 0xAAAAAAAB) & 0xC0000003
    return (iov & 3) | ((iov >> 30) << 2)


def _outer2(h, f, m):
    r = _ROOT3F
    tx = [r, -r, 1, -1, r, -r, -1, 1, 2, -2, 0, 0, 0, 0, 0, 0]
    ty = [1, -1, r, -r, -1, 1, r, -r, 0, 0, 2, -2, 0, 0, 0, 0]
    hs = _u32(h) >> 1
    ig = _u32(hs * 3) >> 28
    io = _outer_index(hs)
    mm = m * (f[1] * np.array(ty)[ig] + f[0] * np.array(tx)[ig])
    return [mm * np.array(tx)[io], mm * np.array(ty)[io]]


def _outer3(h, f, m):
    tx = [1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0]
    ty = [1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0]
    tz = [0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0]
    hs = _u32(h) >> 1
    ig = _u32(hs * 3) >> 28
    io = _outer_index(hs)
    t = [np.array(tx), np.array(ty), np.array(tz)]
    mm = m * sum(t[a][ig] * f[a] for a in range(3))
    return [mm * t[a][io] for a in range(3)]


def _outer4(h, f, m):
    s, s1 = _SKEW4F, _ff(_SKEW4F, 1)
    tabs = [[s1, s, s, s, -1, 1, 0, 0, -1, 0, 1, 0, -1, 0, 0, 1],
            [s, s1, s, s, 1, -1, 0, 0, 0, -1, 0, 1, 0, -1, 1, 0],
            [s, s, s1, s, 0, 0, -1, 1, 1, 0, -1, 0, 0, 1, -1, 0],
            [s, s, s, s1, 0, 0, 1, -1, 0, 1, 0, -1, 1, 0, 0, -1]]
    hs = _u32(h) >> 2
    ig = _u32(hs * 5) >> 27
    iov = _u32(hs * 0xCCCCCCCD)
    io = _u32((iov & 0xE0000003) * (3 | (1 << 27))) >> 27
    g = [_perm2(t, [-x for x in t], ig) for t in tabs]
    o = [_perm2(t, [-x for x in t], io) for t in tabs]
    mm = m * sum(g[a] * f[a] for a in range(4))
    return [mm * o[a] for a in range(4)]


def _ortho3(h, f, m):
    A, B, C, DE, F = 2.224744871391589, -0.224744871391589, -1.0, 1.0, 2.0
    A, B = float(f32(A)), float(f32(B))
    tA = ([A, A, C, C, -A, -A, C, C, A, A, C, C, -A, -A, C, C],
          [C, B, A, A, C, B, -A, -A, C, -B, A, A, C, -B, -A, -A],
          [B, C, B, B, B, C, B, B, -B, C, -B, -B, -B, C, -B, -B])
    tB = ([B, B, C, C, -B, -B, C, C, B, B, C, C, -B, -B, C, C],
          [C, A, B, B, C, A, -B, -B, C, -A, B, B, C, -A, -B, -B],
          [A, C, A, A, A, C, A, A, -A, C, -A, -A, -A, C, -A, -A])
    tC = ([DE, DE, F, F, C, C, F, F, DE, DE, F, F, C, C, F, F],
          [F, DE, DE, DE, F, DE, C, C, F, C, DE, DE, F, C, C, C],
          [DE, F, DE, DE, DE, F, DE, DE, C, F, C, C, C, F, C, C])
    hs = _u32(h) >> 1
    ifb = _u32(hs * 0xAAAAAAAB)
    rot = _u32((ifb << 2) | (ifb >> 30))
    ip3 = _u32(hs * 3) >> 30
    vab = sum(_perm2(tA[a], tB[a], rot) * f[a] for a in range(3))
    vba = sum(_perm2(tB[a], tA[a], rot) * f[a] for a in range(3))
    vc = sum(np.array(tC[a])[rot & 15] * f[a] for a in range(3))
    mm = m * np.where(_u32(h) & 1, -1.0, 1.0)
    return [mm * np.where(ip3 == 0, vc, vab),
            mm * np.where(ip3 == 1, vc, np.where(ip3 == 2, vba, vab)),
            mm * np.where(ip3 == 2, vc, vba)]


def _ortho4_tables():
    kNegAB, kFillAB = float(f32(-0.375999676691291)), float(f32(0.222726847849776))
    kPosD, kUnD = -_SKEW4F, _SKEW4F
    dPSAB, dPSD = float(f32(-0.124000323308709)), float(f32(0.190983005625053))
    dSE, dABC, dFD = _SKEW4F, float(f32(0.437016024448821)), -_ROOT2F
    dPSEAB, dPSED = _ff(dPSAB, dSE), _ff(dPSD, dSE)
    sIdxABC, sDiagABC, sFillABC = _ff(kNegAB, dPSAB), _ff(kFillAB, dPSAB, dFD), _ff(kFillAB, dPSAB)
    sIdxD, sFillD = _ff(kPosD, dPSD, -1), _ff(kUnD, dPSD)
    pIdxPosAB, pIdxNegAB, pFillAB, pDiagAB = _ff(kNegAB, 1), kNegAB, kFillAB, _ff(kFillAB, dFD)
    pIdxPosC, pIdxNegC, pFillC = _ff(kNegAB, dABC, 1), _ff(kNegAB, dABC), _ff(kFillAB, dABC)
    pIdxPosD, pIdxNegD, pFillD = kPosD, _ff(kPosD, -1), kUnD
    eIdxABC, eDiagABC, eFillABC = _ff(kNegAB, dPSEAB, 1), _ff(kFillAB, dPSEAB, dFD), _ff(kFillAB, dPSEAB)
    eIdxD, eFillD = _ff(kPosD, dPSED), _ff(kUnD, dPSED)
    z = [0] * 12
    return {
        "A_gX": ([sIdxABC, sDiagABC, sDiagABC, sDiagABC, pIdxPosAB, pIdxNegAB, pDiagAB, pDiagAB, pIdxPosAB, pDiagAB, pIdxNegAB, pDiagAB, pIdxPosAB, pDiagAB, pDiagAB, pIdxNegAB], [eIdxABC, eDiagABC, eDiagABC, eDiagABC] + z),
        "B_gX": ([sIdxABC, sFillABC, sFillABC, sFillABC, pIdxPosAB, pIdxNegAB, pFillAB, pFillAB, pIdxPosAB, pFillAB, pIdxNegAB, pFillAB, pIdxPosAB, pFillAB, pFillAB, pIdxNegAB], [eIdxABC, eFillABC, eFillABC, eFillABC] + z),
        "C_gX": ([sIdxABC, sFillABC, sFillABC, sFillABC, pIdxPosC, pIdxNegC, pFillC, pFillC, pIdxPosC, pFillC, pIdxNegC, pFillC, pIdxPosC, pFillC, pFillC, pIdxNegC], [eIdxABC, eFillABC, eFillABC, eFillABC] + z),
        "D_gX": ([sIdxD, sFillD, sFillD, sFillD, pIdxPosD, pIdxNegD, pFillD, pFillD, pIdxPosD, pFillD, pIdxNegD, pFillD, pIdxPosD, pFillD, pFillD, pIdxNegD], [eIdxD, eFillD, eFillD, eFillD] + z),
        "A_gY": ([sDiagABC, sIdxABC, sFillABC, sFillABC, pIdxNegAB, pIdxPosAB, pFillAB, pFillAB, pDiagAB, pIdxPosAB, pDiagAB, pIdxNegAB, pDiagAB, pIdxPosAB, pIdxNegAB, pDiagAB], [eDiagABC, eIdxABC, eFillABC, eFillABC] + z),
        "B_gY": ([sFillABC, sIdxABC, sDiagABC, sDiagABC, pIdxNegAB, pIdxPosAB, pDiagAB, pDiagAB, pFillAB, pIdxPosAB, pFillAB, pIdxNegAB, pFillAB, pIdxPosAB, pIdxNegAB, pFillAB], [eFillABC, eIdxABC, eDiagABC, eDiagABC] + z),
        "C_gY": ([sFillABC, sIdxABC, sFillABC, sFillABC, pIdxNegC, pIdxPosC, pFillC, pFillC, pFillC, pIdxPosC, pFillC, pIdxNegC, pFillC, pIdxPosC, pIdxNegC, pFillC], [eFillABC, eIdxABC, eFillABC, eFillABC] + z),
        "D_gY": ([sFillD, sIdxD, sFillD, sFillD, pIdxNegD, pIdxPosD, pFillD, pFillD, pFillD, pIdxPosD, pFillD, pIdxNegD, pFillD, pIdxPosD, pIdxNegD, pFillD], [eFillD, eIdxD, eFillD, eFillD] + z),
        "A_gZ": ([sFillABC, sFillABC, sIdxABC, sFillABC, pDiagAB, pDiagAB, pIdxPosAB, pIdxNegAB, pIdxNegAB, pFillAB, pIdxPosAB, pFillAB, pFillAB, pIdxNegAB, pIdxPosAB, pFillAB], [eFillABC, eFillABC, eIdxABC, eFillABC] + z),
        "B_gZ": ([sDiagABC, sDiagABC, sIdxABC, sFillABC, pFillAB, pFillAB, pIdxPosAB, pIdxNegAB, pIdxNegAB, pDiagAB, pIdxPosAB, pDiagAB, pDiagAB, pIdxNegAB, pIdxPosAB, pDiagAB], [eDiagABC, eDiagABC, eIdxABC, eFillABC] + z),
        "C_gZ": ([sFillABC, sFillABC, sIdxABC, sDiagABC, pFillC, pFillC, pIdxPosC, pIdxNegC, pIdxNegC, pFillC, pIdxPosC, pFillC, pFillC, pIdxNegC, pIdxPosC, pFillC], [eFillABC, eFillABC, eIdxABC, eDiagABC] + z),
        "D_gZ": ([sFillD, sFillD, sIdxD, sFillD, pFillD, pFillD, pIdxPosD, pIdxNegD, pIdxNegD, pFillD, pIdxPosD, pFillD, pFillD, pIdxNegD, pIdxPosD, pFillD], [eFillD, eFillD, eIdxD, eFillD] + z),
        "A_gW": ([sFillABC, sFillABC, sFillABC, sIdxABC, pFillAB, pFillAB, pIdxNegAB, pIdxPosAB, pFillAB, pIdxNegAB, pFillAB, pIdxPosAB, pIdxNegAB, pFillAB, pFillAB, pIdxPosAB], [eFillABC, eFillABC, eFillABC, eIdxABC] + z),
        "B_gW": ([sFillABC, sFillABC, sFillABC, sIdxABC, pDiagAB, pDiagAB, pIdxNegAB, pIdxPosAB, pDiagAB, pIdxNegAB, pDiagAB, pIdxPosAB, pIdxNegAB, pDiagAB, pDiagAB, pIdxPosAB], [eFillABC, eFillABC, eFillABC, eIdxABC] + z),
        "C_gW": ([sDiagABC, sDiagABC, sDiagABC, sIdxABC, pFillC, pFillC, pIdxNegC, pIdxPosC, pFillC, pIdxNegC, pFillC, pIdxPosC, pIdxNegC, pFillC, pFillC, pIdxPosC], [eDiagABC, eDiagABC, eDiagABC, eIdxABC] + z),
        "D_gW": ([sFillD, sFillD, sFillD, sIdxD, pFillD, pFillD, pIdxNegD, pIdxPosD, pFillD, pIdxNegD, pFillD, pIdxPosD, pIdxNegD, pFillD, pFillD, pIdxPosD], [eFillD, eFillD, eFillD, eIdxD] + z),
    }


_ORTHO4 = None


def _ortho4(h, f, m):
    global _ORTHO4
    if _ORTHO4 is None:
        _ORTHO4 = _ortho4_tables()
    hs = _u32(h) >> 2
    basis = _u32(hs * 5)
    idx = basis >> 27
    ip3 = _s32(hs * 0xD5555556) >> 29          # arithmetic shift of the int32
    ip8 = _s32(basis) >> 24
    vals = {k: sum(_perm2(*_ORTHO4["%s_g%s" % (k, ax)], idx) * f[a] for a, ax in enumerate("XYZW")) for k in "ABCD"}
    ix = ip8 >> 1
    iy = ix + 1 + ip3
    iz0 = ip8 & 1
    iz = ix + 1 + np.where(iz0 >= ip3, iz0 + 1, iz0)
    isum = ix + iy + iz
    pick = lambda i, order: np.choose(np.asarray(i) & 3, [vals[k] for k in order])
    mm = m * np.where(_u32(h) & 1, -1.0, 1.0)
    return [mm * pick(ix, "ABCD"), mm * pick(iy, "ABCD"), mm * pick(iz, "ABCD"), mm * pick(isum, "CBAD")]


WARP_VECTOR = {(2, 0): _ortho2, (3, 0): _ortho3, (4, 0): _ortho4, (2, 1): _outer2, (3, 1): _outer3, (4, 1): _outer4}

# Output bounding of the simplex-type warps (DomainWarpSimplex.inl:823-825,
# 919-921, 1073-1075, 1208-1210, 1434-1436, 1821-1823): (scheme 0, scheme 1)
WARP_BOUNDING = {
    ("Simplex", 2): (70.1480577066486, 49.918426513671875 / 2.0),
    ("Simplex", 3): (16.281631889139874, 32.69428253173828125 / 1.4142135623730951),
    ("Simplex", 4): (30.88161777516092, 33.653125584827855 / 1.4142135623730951),
    ("SuperSimplex", 2): (12.814453124999995, 9.28993664146183 / 2.0),
    ("SuperSimplex", 3): (37.63698669623629, 144.736422163332608 / 1.4142135623730951),
    ("SuperSimplex", 4): (48.80058117543753, 115.21625311930542 / 1.4142135623730951),
}


def warp_simplex_vector(name, pos, seed, freq, scheme, helper=None):
    """DomainWarpSimplex / SuperSimplex's warp vector before amplitude and axis
    scale: bounding * sum over the node's lattice vertices of
    kernel(|d|^2) * vector(HashPrimes(L), d), over the same vertex sets and
    kernels as the noise nodes. `seed` is the warp's hash seed: the call's
    seed plus twice the node's seed offset for DomainWarpSimplex (added in
    DomainWarp::GenT and again in its Warp), once for DomainWarpSuperSimplex,
    whose Warp adds none (DomainWarpSimplex.inl:731-766, 1094-1128).
    Returns (vector per axis, |vector|^2 at unit axis scale)."""
    d = len(pos)
    r2, ks, _ = SIMPLEX_KERNEL[(name, d)]
    bound = float(f32(WARP_BOUNDING[(name, d)][scheme]))
    p = scale_pos(pos, freq)
    acc = [np.zeros_like(p[0]) for _ in range(d)]
    for L, dv in lattice_points(p, d, VERTEX_SET[(name, d)]):
        q = sum(c * c for c in dv)
        fall = np.where(q < r2, np.maximum(0.0, ks * (r2 - q)) ** 4, 0.0)
        v = (helper or WARP_VECTOR[(d, scheme)])(hash_primes(seed, *primed(L)), dv, fall)
        for a in range(d):
            acc[a] += v[a]
    if d == 3 and scheme == 1:
        # the outer-product vectors are turned to the 3D lattice's reflected
        # offsets: v += k (v.1) 1 with k = -1 (Simplex) or -2/3 (SuperSimplex)
        # (DomainWarpSimplex.inl:909-917, 1424-1432)
        k = float(f32(-1.0)) if name == "Simplex" else float(f32(-2.0 / 3.0))
        t = k * sum(acc)
        acc = [a + t for a in acc]
    vec = [a * bound for a in acc]
    return vec, sum(v * v for v in vec)


def warp_table_domain(d, scheme, h):
    """Hashes for which the reference's table branches and generic branch of
    a simplex-warp vector helper compute the same vector: every table index
    inside the range the tables were laid out for ([0,12) gradient and outer
    indices, [0,3) permutation selectors in 2D/3D; [0,20) basis indices in
    4D) and, for the 4D orthogonal matrix, the pairwise basis group [4,16)
    only (its generic branch compares the basis index against 0x10 << 27 and
    0x04 << 27 as signed integers, which folds the single and extra groups
    together). Outside this domain the reference's SIMD levels disagree with
    each other; the oracle follows the generic branch."""
    u = _u32(h)
    hs1, hs2 = u >> 1, u >> 2
    if d in (2, 3):
        grad_ok = (_u32(hs1 * 3) >> 28) < 12
        outer_ok = (_u32(hs1 * 0xAAAAAAAB) >> 30) < 3
        if (d, scheme) == (2, 0):
            return grad_ok
        if scheme == 1:
            return grad_ok & outer_ok
        return outer_ok & ((_u32(hs1 * 3) >> 30) < 3)
    basis = _u32(hs2 * 5) >> 27
    if scheme == 0:
        return (basis >= 4) & (basis < 16)
    outer = _u32((_u32(hs2 * 0xCCCCCCCD) & 0xE0000003) * (3 | (1 << 27))) >> 27
    return (basis < 20) & (outer < 20)
