"""The benchmark workloads of the generator path, as node trees and inputs.

Frequencies are given the way FastNoise2's node editor shows them: a node
with frequency f has Feature Scale 1/f (BasicGenerators.h:20-24 of the
reference: mFrequency = 1 / scale), so 0.02 -> 50, 0.005 -> 200, 0.01 -> 100.
"""
import numpy as np

from .tree import NodeSpec, decode

S = NodeSpec


def perlin_3d():
    """Perlin, frequency 0.02: GenUniformGrid3D 64^3 at (0,0,0), seed 1337."""
    return S("Perlin", feature_scale=50)


def fbm_simplex_2d():
    """FractalFBm(Simplex), 5 octaves, gain 0.5, lacunarity 2, weighted strength 0,
    frequency 0.005: GenUniformGrid2D 4096^2 at (-2048,-2048) seed 42, and
    GenTileable2D 1024^2."""
    return S("FractalFBm", source=S("Simplex", feature_scale=200), octaves=5, gain=0.5,
             lacunarity=2.0, weighted_strength=0.0)


def warped_cellular_3d():
    """DomainWarpGradient (amplitude 30, frequency 0.01) over CellularDistance
    (Euclidean, Index0Add1, jitter 1, frequency 0.02): GenUniformGrid3D 256^3
    at (0,0,0) seed 1337, and GenPositionArray3D on 16,777,216 positions."""
    cell = S("CellularDistance", feature_scale=50, distance_function="Euclidean",
             return_type="Index0Add1", distance_index_0=0, distance_index_1=1, grid_jitter=1.0)
    return S("DomainWarpGradient", source=cell, feature_scale=100, warp_amplitude=30.0)


FBM_SUPERSIMPLEX_SUM = 1.96875  # amplitude sum of 6 octaves at gain 0.5 (FBm output is not normalised)


def serving_tree():
    """Add( Remap( DomainWarpFractalProgressive( 3 octaves, DomainWarpSimplex(
    amplitude 40, FractalFBm( SuperSimplex, 6 octaves ) ) ), from +-1.96875 to
    [0,1] ), Multiply( CellularValue, 0.3 ) )."""
    fbm = S("FractalFBm", source=S("SuperSimplex"), octaves=6)
    warp = S("DomainWarpSimplex", source=fbm, warp_amplitude=40.0)
    prog = S("DomainWarpFractalProgressive", domain_warp_source=warp, octaves=3)
    remap = S("Remap", source=prog, from_min=-FBM_SUPERSIMPLEX_SUM, from_max=FBM_SUPERSIMPLEX_SUM, to_min=0.0, to_max=1.0)
    cell = S("Multiply", lhs=S("CellularValue"), rhs=0.3)
    return S("Add", lhs=remap, rhs=cell)


def world_tree():
    """Terrain density for the chunked-world client: a height gradient
    (value 0.01 per unit of y) plus FractalFBm(Perlin, 4 octaves, frequency
    0.01); the iso surface 0 lies near y = 0."""
    return S("Add", lhs=S("Gradient", multiplier_y=0.01),
             rhs=S("FractalFBm", source=S("Perlin", feature_scale=100.0), octaves=4))


def splitmix64_positions(n, seed=7, lo=-1000.0, hi=1000.0):
    """SplitMix64 stream seeded `seed`; draws 3i, 3i+1, 3i+2 give point i's
    x, y, z = lo + (hi - lo) * ((r >> 40) * 2^-24), computed in float64 and
    rounded once to float32."""
    out = np.empty((3, n), dtype=np.float32)
    golden = np.uint64(0x9E3779B97F4A7C15)
    chunk = 1 << 21
    with np.errstate(over="ignore"):
        for base in range(0, 3 * n, chunk):
            k = np.arange(base, min(3 * n, base + chunk), dtype=np.uint64)
            z = np.uint64(seed) + (k + np.uint64(1)) * golden
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            u = (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
            v = (lo + (hi - lo) * u).astype(np.float32)
            idx = k.astype(np.int64)
            out[idx % 3, idx // 3] = v
    return out[0], out[1], out[2]


SERVING_SIZES = [("3d", 16), ("3d", 32), ("3d", 64), ("2d", 128), ("2d", 512)]


def serving_trace(n=20000, rate=2000.0, seed=2024):
    """Poisson request stream against the serving tree: mean `rate` requests/s;
    each request has a uniform int32 seed, a size drawn uniformly from
    16^3, 32^3, 64^3 (3D grids) and 128^2, 512^2 (2D grids), and an origin on a
    chunk-aligned multiple in [-1e5, 1e5]. Returns (specs, arrival seconds)."""
    from . import grid_spec
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    specs = []
    for _ in range(n):
        kind, c = SERVING_SIZES[int(rng.integers(len(SERVING_SIZES)))]
        s = int(rng.integers(-2**31, 2**31))
        d = 3 if kind == "3d" else 2
        k = int(1e5 // c)
        origin = [float(v * c) for v in rng.integers(-k, k + 1, d)]
        specs.append(grid_spec(d, origin, [c] * d, seed=s))
    return specs, arrivals


# ---------------------------------------------------------------- mixed trees
_LEAVES = ["Perlin", "Simplex", "SuperSimplex", "Value", "CellularValue", "CellularDistance"]


def random_tree(rng, depth):
    """A seeded random node tree of about `depth` levels built from lattice and
    cellular leaves, FBm/Ridged, Add/Multiply/MinSmooth/Fade, Remap/Terrace,
    DomainScale/DomainRotate and domain warps."""
    def leaf():
        name = _LEAVES[int(rng.integers(len(_LEAVES)))]
        kw = {"feature_scale": float(rng.choice([20.0, 40.0, 80.0, 150.0])), "seed_offset": int(rng.integers(0, 5))}
        if name.startswith("Cellular"):
            kw["distance_function"] = str(rng.choice(["Euclidean", "Euclidean Squared", "Manhattan", "Hybrid"]))
        return S(name, **kw)

    def build(d):
        if d <= 1:
            return leaf()
        kind = int(rng.integers(8))
        if kind == 0:
            return S(str(rng.choice(["FractalFBm", "FractalRidged"])), source=leaf(), octaves=int(rng.integers(2, 6)))
        if kind == 1:
            return S(str(rng.choice(["Add", "Multiply"])), lhs=build(d - 1), rhs=build(d - 2))
        if kind == 2:
            return S("MinSmooth", lhs=build(d - 1), rhs=build(d - 2), smoothness=float(rng.uniform(0.05, 0.5)))
        if kind == 3:
            return S("Fade", a=build(d - 1), b=build(d - 2), fade=build(d - 2))
        if kind == 4:
            return S("Remap", source=build(d - 1), from_min=-1.0, from_max=1.0, to_min=0.0, to_max=1.0)
        if kind == 5:
            return S("Terrace", source=build(d - 1), step_count=float(rng.integers(2, 9)), smoothness=float(rng.uniform(0, 1)))
        if kind == 6:
            if rng.integers(2):
                return S("DomainScale", source=build(d - 1), scaling=float(rng.uniform(0.5, 2.0)))
            return S("DomainRotate", source=build(d - 1), yaw=float(rng.uniform(0, 3)), pitch=float(rng.uniform(0, 1)))
        warp = str(rng.choice(["DomainWarpGradient", "DomainWarpSimplex", "DomainWarpSuperSimplex"]))
        w = S(warp, source=build(d - 1), warp_amplitude=float(rng.uniform(5, 40)), feature_scale=float(rng.choice([50.0, 100.0])))
        if rng.integers(2):
            return S(str(rng.choice(["DomainWarpFractalProgressive", "DomainWarpFractalIndependent"])),
                     domain_warp_source=w, octaves=int(rng.integers(2, 4)))
        return w

    return build(depth)


def mixed_trees(seed=5, count=8):
    """The 8 mixed-workload trees (depths 3..12), as encoded strings."""
    rng = np.random.default_rng(seed)
    depths = np.linspace(3, 12, count).round().astype(int)
    return [random_tree(rng, int(d)).encode() for d in depths]


# (FP32 FLOPs, vector ALU operations) per point of each mixed tree in 2D, 3D
# and 4D, counted by the op-counting oracle on a 32^2 / 16^3 / 8^4 grid at the
# origin (the trees branch on nothing, so the counts do not depend on position)
MIXED_TREE_OPS = [
    {2: (590.0, 936.0), 3: (592.0, 937.0), 4: (2174.0, 3446.0)},
    {2: (288.0, 607.0), 3: (381.0, 1059.0), 4: (653.0, 2239.0)},
    {2: (785.0, 1326.0), 3: (2398.0, 4171.0), 4: (8417.0, 14258.0)},
    {2: (751.0, 1333.0), 3: (2114.0, 3643.0), 4: (6971.0, 11024.0)},
    {2: (389.0, 1101.0), 3: (391.0, 1102.0), 4: (672.0, 2398.0)},
    {2: (6027.0, 10739.0), 3: (10071.0, 19003.0), 4: (36019.0, 65490.0)},
    {2: (4160.0, 8807.0), 3: (5332.0, 10688.0), 4: (15361.0, 28849.0)},
    {2: (12865.0, 20614.0), 3: (24583.0, 41557.0), 4: (85815.0, 137879.0)},
]
MIXED_OPS_SAMPLE = {2: [32, 32], 3: [16, 16, 16], 4: [8, 8, 8, 8]}


def mixed_trace(n=4000, seed=77, burst_ms=10.0, burst_rate=10000.0, off_mean_ms=40.0, ntrees=8):
    """Bursty requests over the mixed trees: on/off arrivals (bursts of
    `burst_ms` at `burst_rate` requests/s, exponential off periods), uniform
    tree, dimension 2/3/4, size log-uniform in [2^12, 2^22] points (a grid of
    side round(N^(1/d))), uniform seed and origin. Returns (tree indices,
    specs, arrival seconds)."""
    from . import grid_spec
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while len(arrivals) < n:
        end = t + burst_ms / 1000.0
        while t < end and len(arrivals) < n:
            arrivals.append(t)
            t += rng.exponential(1.0 / burst_rate)
        t = end + rng.exponential(off_mean_ms / 1000.0)
    trees, specs = [], []
    for _ in range(n):
        d = int(rng.integers(2, 5))
        pts = 2.0 ** rng.uniform(12, 22)
        side = max(2, int(round(pts ** (1.0 / d))))
        origin = [float(v) for v in rng.integers(-100000, 100000, d)]
        trees.append(int(rng.integers(ntrees)))
        specs.append(grid_spec(d, origin, [side] * d, seed=int(rng.integers(-2**31, 2**31))))
    return trees, specs, np.array(arrivals)


def grid_spec_like(dims, points=1 << 14, seed=3):
    """A small grid request of `dims` dimensions (warm-up of compile and tuning)."""
    from . import grid_spec
    side = max(2, int(round(points ** (1.0 / dims))))
    return grid_spec(dims, [0.0] * dims, [side] * dims, seed=seed)


def bench_workloads():
    """The single-tree throughput workloads: name -> tree, call shape and the
    algorithmic work per point.

    `ops` is (FP32 FLOPs with an fma counted as 2, vector ALU operations) per
    point, counted by the op-counting build of the oracle over `ops_sample`
    (a corner of the workload's own grid, or a coarse grid over the position
    cube); the counts are committed here and recounted by the CPU tests.
    `pruned`: the tree's CellularDistance takes the pruned cell walk, which
    leaves out cells the reference's fixed 27-cell walk (and so the count)
    includes."""
    return {
        "fbm_simplex_2d_grid_4096x4096": dict(
            tree=fbm_simplex_2d, kind="grid", dims=2, offset=[-2048, -2048], count=[4096, 4096], seed=42,
            ops=(299.0, 572.0), ops_sample=dict(dims=2, offset=[-2048, -2048], count=[32, 32])),
        "fbm_simplex_tileable_1024x1024": dict(
            tree=fbm_simplex_2d, kind="tileable", dims=2, count=[1024, 1024], seed=42,
            ops=(653.0, 1703.0), ops_sample=dict(dims=2, offset=None, count=[64, 64], tileable=True)),
        "perlin_3d_grid_64x64x64": dict(
            tree=perlin_3d, kind="grid", dims=3, offset=[0, 0, 0], count=[64, 64, 64], seed=1337,
            ops=(69.0, 209.0), ops_sample=dict(dims=3, offset=[0, 0, 0], count=[32, 32, 32])),
        "warped_cellular_3d_grid_256x256x256": dict(
            tree=warped_cellular_3d, kind="grid", dims=3, offset=[0, 0, 0], count=[256, 256, 256], seed=1337,
            ops=(837.0, 1229.0), ops_sample=dict(dims=3, offset=[0, 0, 0], count=[32, 32, 32]), pruned=True),
        "warped_cellular_positions_16777216": dict(
            tree=warped_cellular_3d, kind="positions", dims=3, count=[1 << 24], seed=1337,
            ops=(837.0, 1229.0), pruned=True,
            ops_sample=dict(dims=3, offset=[-1000, -1000, -1000], count=[16, 16, 16], step=[125, 125, 125])),
        "serving_tree_3d_grid_256x256x256": dict(
            tree=serving_tree, kind="grid", dims=3, offset=[0, 0, 0], count=[256, 256, 256], seed=1337,
            ops=(1950.0, 4255.0), ops_sample=dict(dims=3, offset=[0, 0, 0], count=[32, 32, 32])),
        # cheap single-node grids: bounded by HBM writes, not by the ALU
        "constant_2d_grid_16384x16384": dict(
            tree=lambda: S("Constant"), kind="grid", dims=2, offset=[0, 0], count=[16384, 16384], seed=1337,
            ops=(4.0, 2.0), ops_sample=dict(dims=2, offset=[0, 0], count=[64, 64])),
        "checkerboard_2d_grid_16384x16384": dict(
            tree=lambda: S("Checkerboard"), kind="grid", dims=2, offset=[0, 0], count=[16384, 16384], seed=1337,
            ops=(10.0, 12.0), ops_sample=dict(dims=2, offset=[0, 0], count=[64, 64])),
        "mixed_tree_deepest_3d_grid_128x128x128": dict(
            tree=lambda: decode(mixed_trees()[-1]), kind="grid", dims=3, offset=[0, 0, 0], count=[128, 128, 128],
            seed=1337, ops=(24583.0, 41557.0), ops_sample=dict(dims=3, offset=[0, 0, 0], count=[32, 32, 32])),
    }
