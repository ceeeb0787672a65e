"""Node-by-node parity cases shared by the GPU parity tests.

Every case is a node tree (NodeSpec), the dimensions it is evaluated in and
a tolerance: 0 means bit-exact against the oracle's fused variant. A
tolerance is only given where the result goes through a transcendental
function whose FastSIMD implementation no reference fixture pins (FS::Sin,
FS::Cos, FS::Pow), so device and host math libraries may round differently.
"""
from fastnoise2_amd.tree import NodeSpec as S

LATTICE = ["Perlin", "Value", "Simplex", "SuperSimplex"]
DISTANCES = ["Euclidean", "Euclidean Squared", "Manhattan", "Hybrid", "Max Axis", "Minkowski"]
RETURNS = ["Index0", "Index0Add1", "Index0Sub1", "Index0Mul1", "Index0Div1"]

# grid regions per dimension: non-integer offsets and steps, negative coordinates
REGION = {
    2: dict(offset=[-40.25, 17.5], count=[48, 40], step=[0.9, 1.7]),
    3: dict(offset=[-12.5, 3.25, -30.0], count=[20, 16, 14], step=[1.1, 0.8, 2.3]),
    4: dict(offset=[5.5, -7.25, 2.0, -1.5], count=[9, 8, 7, 6], step=[1.3, 0.7, 1.9, 2.2]),
}


def _cases():
    out = []

    def add(name, spec, dims=(2, 3, 4), tol=0.0, seed=1337):
        for d in dims:
            out.append(("%s-%dd" % (name, d), spec, d, tol, seed))

    # ---- generators
    add("Constant", S("Constant", value=0.7))
    add("White", S("White", seed_offset=3))
    add("White-range", S("White", output_min=-3.0, output_max=5.0))
    add("Checkerboard", S("Checkerboard", feature_scale=7.0))
    add("SineWave", S("SineWave", feature_scale=23.0), tol=1e-5)
    add("Gradient", S("Gradient", multiplier_x=0.3, multiplier_y=-1.2, multiplier_z=0.5, multiplier_w=2.0,
                      offset_x=1.5, offset_y=-2.0, offset_z=0.25, offset_w=3.0))
    add("Gradient-offset-node", S("Gradient", multiplier_x=1.0, offset_y=S("Perlin", feature_scale=30)))
    for df in DISTANCES:
        tol = 1e-5 if df == "Minkowski" else 0.0
        add("DistanceToPoint-" + df.replace(" ", ""),
            S("DistanceToPoint", distance_function=df, point_x=3.0, point_y=-4.0, point_z=1.5, point_w=-2.0), tol=tol)
    for n in LATTICE:
        add(n, S(n))
        add(n + "-scaled", S(n, feature_scale=13.0, seed_offset=-5, output_min=0.0, output_max=10.0), seed=-77)
    # ---- cellular
    for df in DISTANCES:
        tol = 1e-5 if df == "Minkowski" else 0.0
        dn = df.replace(" ", "")
        add("CellularValue-" + dn, S("CellularValue", feature_scale=9.0, distance_function=df), tol=tol)
        add("CellularDistance-" + dn, S("CellularDistance", feature_scale=9.0, distance_function=df), tol=tol)
        add("CellularLookup-" + dn, S("CellularLookup", feature_scale=9.0, distance_function=df,
                                      lookup=S("Simplex", feature_scale=3.0)), tol=tol)
    for vi in range(4):
        add("CellularValue-index%d" % vi, S("CellularValue", feature_scale=11.0, value_index=vi, grid_jitter=0.7))
    for rt in RETURNS:
        add("CellularDistance-" + rt, S("CellularDistance", feature_scale=11.0, distance_function="Euclidean",
                                         return_type=rt, distance_index_0=0, distance_index_1=2))
    add("CellularDistance-idx3", S("CellularDistance", feature_scale=11.0, distance_index_0=3, distance_index_1=3,
                                   distance_function="Euclidean", return_type="Index0Add1"))
    add("CellularDistance-sizejitter", S("CellularDistance", feature_scale=8.0, size_jitter=0.4, grid_jitter=0.8))
    add("CellularValue-sizejitter", S("CellularValue", feature_scale=8.0, size_jitter=0.3))
    add("CellularDistance-jitter-node", S("CellularDistance", feature_scale=8.0, grid_jitter=S("Perlin", feature_scale=40)))
    add("CellularValue-jitter0", S("CellularValue", feature_scale=6.0, grid_jitter=0.0))
    add("CellularLookup-seed", S("CellularLookup", feature_scale=10.0, seed_offset=4, lookup=S("White")))
    # ---- fractals
    for fr in ["FractalFBm", "FractalRidged"]:
        for n in LATTICE:
            add("%s-%s" % (fr, n), S(fr, source=S(n, feature_scale=40.0), octaves=4))
        add(fr + "-weighted", S(fr, source=S("Simplex", feature_scale=40.0), octaves=3, gain=0.7, weighted_strength=0.6,
                                lacunarity=2.5))
        add(fr + "-gain-node", S(fr, source=S("Perlin", feature_scale=40.0), octaves=3,
                                 gain=S("Value", feature_scale=90.0), weighted_strength=S("Simplex", feature_scale=70.0)))
        add(fr + "-generic-source", S(fr, source=S("Abs", source=S("Simplex", feature_scale=30.0)), octaves=3))
        add(fr + "-1oct", S(fr, source=S("SuperSimplex", feature_scale=30.0), octaves=1))
    add("PingPong", S("PingPong", source=S("Simplex", feature_scale=30.0), ping_pong_strength=3.5))
    add("PingPong-node", S("PingPong", source=S("Perlin", feature_scale=30.0), ping_pong_strength=S("Value", feature_scale=50)))
    # ---- domain warps
    for w in ["DomainWarpGradient", "DomainWarpSimplex", "DomainWarpSuperSimplex"]:
        add(w, S(w, source=S("Perlin", feature_scale=20.0), warp_amplitude=12.0, feature_scale=35.0))
        add(w + "-axis", S(w, source=S("Value", feature_scale=20.0), warp_amplitude=8.0, feature_scale=25.0, seed_offset=2,
                          amplitude_scaling_x=0.5, amplitude_scaling_y=2.0, amplitude_scaling_z=1.5, amplitude_scaling_w=0.25))
        add(w + "-amp-node", S(w, source=S("Simplex", feature_scale=20.0), feature_scale=30.0,
                              warp_amplitude=S("Perlin", feature_scale=60.0)))
        if w != "DomainWarpGradient":
            add(w + "-outer", S(w, source=S("Perlin", feature_scale=20.0), warp_amplitude=10.0, feature_scale=30.0,
                                vectorization_scheme="Gradient Outer Product"))
        for fr in ["DomainWarpFractalProgressive", "DomainWarpFractalIndependent"]:
            add("%s-%s" % (fr, w), S(fr, domain_warp_source=S(w, source=S("Simplex", feature_scale=25.0), warp_amplitude=9.0,
                                                                 feature_scale=40.0), octaves=3, weighted_strength=0.4))
    # ---- blends
    a = S("Perlin", feature_scale=25.0)
    b = S("Simplex", feature_scale=35.0)
    for op in ["Add", "Multiply", "Min", "Max"]:
        add(op, S(op, lhs=a, rhs=b))
        add(op + "-const", S(op, lhs=a, rhs=0.35))
    for op in ["Subtract", "Divide"]:
        add(op, S(op, lhs=a, rhs=b))
        add(op + "-const-lhs", S(op, lhs=1.5, rhs=b))
    add("Modulus", S("Modulus", lhs=S("Value", feature_scale=20.0, output_min=-10.0, output_max=10.0), rhs=1.3))
    for op in ["MinSmooth", "MaxSmooth"]:
        add(op, S(op, lhs=a, rhs=b, smoothness=0.3))
        add(op + "-node", S(op, lhs=a, rhs=0.1, smoothness=S("Value", feature_scale=40.0, output_min=0.0, output_max=0.5)))
    add("Abs", S("Abs", source=b))
    add("SignedSquareRoot", S("SignedSquareRoot", source=b))
    add("PowFloat", S("PowFloat", value=S("Perlin", feature_scale=30.0), pow=1.7), tol=1e-5)
    add("PowFloat-node", S("PowFloat", value=2.5, pow=S("Value", feature_scale=30.0)), tol=1e-5)
    for p in (0, 1, 2, 3, 5):
        add("PowInt-%d" % p, S("PowInt", value=a, pow=p))
    add("Fade-linear", S("Fade", a=a, b=b, fade=0.3))
    for interp in ["Linear", "Hermite", "Quintic"]:
        add("Fade-" + interp, S("Fade", a=a, b=b, interpolation=interp, fade=S("Value", feature_scale=50.0), fade_min=-0.5,
                                fade_max=0.8))
    # ---- modifiers
    add("DomainScale", S("DomainScale", source=b, scaling=1.7))
    add("DomainOffset", S("DomainOffset", source=b, offset_x=3.5, offset_y=-1.25, offset_z=S("Value", feature_scale=10.0),
                          offset_w=7.0))
    add("DomainRotate-yaw", S("DomainRotate", source=b, yaw=30.0))
    add("DomainRotate-all", S("DomainRotate", source=b, yaw=15.0, pitch=40.0, roll=-25.0))
    add("DomainAxisScale", S("DomainAxisScale", source=b, scaling_x=0.5, scaling_y=2.0, scaling_z=1.5, scaling_w=3.0))
    add("SeedOffset", S("SeedOffset", source=b, seed_offset=17))
    add("ConvertRGBA8", S("ConvertRGBA8", source=b, min=-0.8, max=0.9))
    add("GeneratorCache", S("GeneratorCache", source=b))
    add("Remap", S("Remap", source=b, from_min=-0.5, from_max=0.5, to_min=10.0, to_max=-3.0))
    add("Remap-clamp", S("Remap", source=b, from_min=-0.5, from_max=0.5, to_min=0.0, to_max=1.0, clamp_output="True"))
    add("Remap-nodes", S("Remap", source=b, from_min=S("Constant", value=-0.7), to_max=S("Value", feature_scale=30.0)))
    add("Terrace", S("Terrace", source=b, step_count=5.0))
    add("Terrace-smooth", S("Terrace", source=b, step_count=4.0, smoothness=0.6))
    add("Terrace-smooth-node", S("Terrace", source=b, step_count=3.0, smoothness=S("Value", feature_scale=40.0, output_min=0.0)))
    add("AddDimension", S("AddDimension", source=b, new_dimension_position=2.5), dims=(2, 3))
    add("AddDimension-node", S("AddDimension", source=b, new_dimension_position=S("Perlin", feature_scale=20.0)), dims=(2, 3))
    for ax in "XYZW":
        add("RemoveDimension-" + ax, S("RemoveDimension", source=b, remove_dimension=ax), dims=(3, 4))
    add("DomainRotatePlane-XY", S("DomainRotatePlane", source=b, rotation_type="Improve XY Planes"))
    add("DomainRotatePlane-XZ", S("DomainRotatePlane", source=b, rotation_type="Improve XZ Planes"))
    # ---- deep graphs
    shared = S("Simplex", feature_scale=45.0)
    add("Shared-node", S("Add", lhs=shared, rhs=S("Multiply", lhs=shared, rhs=S("Perlin", feature_scale=15.0))))
    deep = S("Add",
             lhs=S("Remap", source=S("DomainWarpFractalProgressive",
                                     domain_warp_source=S("DomainWarpSimplex", source=S("FractalFBm", source=S("SuperSimplex"),
                                                                                        octaves=3), warp_amplitude=20.0),
                                     octaves=2), from_min=-1.75, from_max=1.75),
             rhs=S("Multiply", lhs=S("CellularValue"), rhs=0.3))
    add("Serving-like", deep, dims=(2, 3))
    return out


CASES = _cases()


def node_count(spec, seen=None):
    """Distinct nodes of a tree (a node used twice counts once)."""
    from fastnoise2_amd.tree import NodeSpec
    seen = set() if seen is None else seen
    if id(spec) in seen:
        return 0
    seen.add(id(spec))
    kids = list(spec.lookups.values()) + [h for h in spec.hybrids.values() if isinstance(h, NodeSpec)]
    return 1 + sum(node_count(k, seen) for k in kids)


def replay(spec, dims, seed, tol, values, mode):
    """Index replay of one case's grid against an oracle variant (unfused or
    AVX2 order): `values` (the fused result, from the GPU or the oracle) is
    compared with the variant on the points whose decision fingerprints
    (oracle.index_trace) agree with the fused variant's. Returns (flips,
    ill-conditioned points, max error over the rest, its bound).

    The bound is 1e-5 for a single node and 1e-4 for a graph (or the case's
    own tolerance when larger), per point relative to max(1, |value|) for
    outputs not confined to [-1, 1]; the returned error is that ratio times
    the bound's base. A Divide's quotient is ill-conditioned where its
    denominator is below 1e-2 in magnitude (an input's last-bit difference
    then grows past 1e-5 of the quotient); those points are counted, not
    bounded."""
    import numpy as np
    import oracle
    r = REGION[dims]
    enc = spec.encode()
    fused, fp0 = oracle.index_trace(enc, "grid", oracle.FUSED, seed, r["offset"], r["count"], r["step"])
    alt, fp1 = oracle.index_trace(enc, "grid", mode, seed, r["offset"], r["count"], r["step"])
    agree = fp0 == fp1
    ill = np.zeros(agree.shape, bool)
    if spec.meta.name == "Divide":
        den = spec.hybrids.get(1, 1.0)
        if hasattr(den, "encode"):
            dv, _ = oracle.index_trace(den.encode(), "grid", oracle.FUSED, seed, r["offset"], r["count"], r["step"])
            ill = np.abs(dv) < 1e-2
    values = np.asarray(values, np.float64)
    ok = agree & ~ill & np.isfinite(alt) & np.isfinite(values)
    bound = max(tol, 1e-5 if node_count(spec) == 1 else 1e-4)
    rel = np.abs(values[ok] - alt[ok]) / np.maximum(1.0, np.abs(alt[ok].astype(np.float64)))
    err = float(rel.max(initial=0.0))
    return int((~agree).sum()), int((ill & agree).sum()), err, bound
