"""Throughput of FastNoise2's generator path on one MI355X.

One step = one GenUniformGrid2D-equivalent pass of the workload's node tree
over all of its points, written to a device-resident output buffer with the
output range folded in (the batch call without the PCIe copy). The default
workload is FractalFBm(Simplex), 5 octaves, frequency 0.005, on a 4096x4096
grid at (-2048,-2048), seed 42.

Prints one JSON line: Gpoints/s over K timed steps (host wall clock from the
first launch to the completion of the last), the dominant kernel's roofline
figures from HIP events on its launch stream, and the host baseline: the CPU
oracle (a restatement of the reference algorithm, vectorised to 16 lanes and
threaded) timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpoints/s per graph (GPU vs host SIMD); p50/p99 chunk latency batched vs plain"
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak (an FMA counted as 2)
# vector ALU issue peak behind it: 157.3 / 2 = 78.6 T lane operations per second
# (256 CUs x 128 lanes per clock x 2.4 GHz), one operation per lane per instruction
VALU_PEAK_TOPS = FP32_PEAK_TFLOPS / 2
HBM_PEAK_GBS = 8000.0
COPY_CAL_MIB = 256                # device-to-device copy that scales the HBM counters


def make_spec(fn, W, wl, keep):
    if wl["kind"] == "grid":
        return fn.grid_spec(wl["dims"], wl["offset"], wl["count"], seed=wl["seed"])
    if wl["kind"] == "tileable":
        return fn.tileable_spec(wl["count"][0], wl["count"][1], seed=wl["seed"])
    xyz = W.splitmix64_positions(wl["count"][0], seed=7)
    dev = [fn.DeviceArray.from_host(a) for a in xyz]
    keep.extend(dev)
    return fn.positions_spec(dev, seed=wl["seed"])


def cpu_baseline(wl, enc, target_s=12.0):
    """Times the oracle (16 lanes, one thread per core of this process's CPU
    share, at most 16) on a bounded sample of the workload: a slice of it when
    the whole takes longer than target_s, else whole passes repeated for about
    target_s."""
    import oracle
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t = oracle.Tree(enc)

    def run(rows, reps=1):
        t0 = time.perf_counter()
        for _ in range(reps):
            if wl["kind"] == "grid":
                count = list(wl["count"])
                count[-1] = rows
                t.grid(wl["offset"], count, seed=wl["seed"], lanes=16, threads=threads)
                what = "GenUniformGrid%dD %s" % (wl["dims"], "x".join(map(str, count)))
            elif wl["kind"] == "tileable":
                t.tileable(wl["count"][0], rows, seed=wl["seed"], lanes=16, threads=threads)
                what = "GenTileable2D %dx%d" % (wl["count"][0], rows)
            else:
                t.positions(pos[:rows * 4096], seed=wl["seed"], lanes=16, threads=threads)
                what = "GenPositionArray3D first %d positions" % (rows * 4096)
        n = reps * rows * (int(np.prod(wl["count"][:-1])) if wl["kind"] != "positions" else 4096)
        return n, time.perf_counter() - t0, what if reps == 1 else "%s x %d passes" % (what, reps)

    if wl["kind"] == "positions":
        import fastnoise2_amd.workloads as W
        pos = np.stack(W.splitmix64_positions(wl["count"][0], seed=7))
        pos = [pos[0], pos[1], pos[2]]
        pos = _RowView(pos)
    full_rows = wl["count"][-1] if wl["kind"] != "positions" else wl["count"][0] // 4096
    rows = max(1, full_rows // 64)
    while True:
        n, dt, what = run(rows)
        if dt >= 0.5 or rows == full_rows:
            break
        rows = min(full_rows, rows * 4)
    per_row_s = max(dt, 1e-9) / rows
    if per_row_s * full_rows >= target_s:
        n, dt, what = run(max(rows, int(target_s / per_row_s)))
    else:
        n, dt, what = run(full_rows, max(1, int(round(target_s / (per_row_s * full_rows)))))
    return {"value": n / dt / 1e9, "unit": "Gpoints/s", "cores": threads, "kind": "port",
            "sample": "%s (%d points, %.1f s), CPU oracle 16-lane SIMD, %d threads" % (what, n, dt, threads)}


class _RowView:
    """Slices each of the x, y, z position arrays alike."""

    def __init__(self, arrays):
        self.arrays = arrays

    def __getitem__(self, sl):
        return [a[sl] for a in self.arrays]


def pmc_child(fn, W, wl, name):
    """Counter pass body: three 256 MiB device-to-device copies (the scale
    reference), then the workload's kernel three times; prints its name."""
    nbytes = COPY_CAL_MIB << 20
    src, dst = fn.DeviceArray(nbytes // 4), fn.DeviceArray(nbytes // 4)
    for _ in range(3):
        if fn.lib().fn2Memcpy(dst.ptr, src.ptr, nbytes, 2):
            raise SystemExit("copy failed: " + fn.last_error())
    node = fn.Node.from_encoded(wl["tree"]().encode())
    keep = []
    spec = make_spec(fn, W, wl, keep)
    out = fn.DeviceArray(int(np.prod(wl["count"])))
    fn.set_jit_mode(fn.JIT_SYNC)
    node.bench_device(spec, out, warmup=0, iters=1)
    fn.set_jit_mode(fn.JIT_ASYNC)
    node.bench_device(spec, out, warmup=0, iters=3)
    dims = wl["dims"] if wl["kind"] != "tileable" else 4
    print(json.dumps({"kernel": "fn2_jit_single" if node.jit_state(dims) == 1 else "fn2_generate"}), flush=True)


def measure_traffic(workload):
    """HBM bytes per launch of the workload's kernel: one rocprofv3 --pmc pass
    for FETCH_SIZE, one for WRITE_SIZE (KiB per dispatch), each scaled by what
    the same counter reads for the 256 MiB copy (which moves 256 MiB each way).
    None when rocprofv3 is missing or a pass fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return None
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="fn2_pmc_")
        cmd = ["timeout", "-s", "KILL", "120", "rocprofv3", "--pmc", counter, "--kernel-trace", "-f", "csv",
               "-d", d, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--workload", workload,
               "--pmc-child"]
        p = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, FN2_JIT="async"))
        kernel = None
        for line in p.stdout.splitlines():
            if line.startswith("{"):
                kernel = json.loads(line).get("kernel")
        rows = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    with open(os.path.join(root, f)) as fh:
                        rows += list(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        copy = [float(r["Counter_Value"]) for r in rows
                if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer") and int(r["Grid_Size"]) >= 1 << 16]
        gen = [float(r["Counter_Value"]) for r in rows if kernel and kernel in r["Kernel_Name"]]
        if not copy or len(gen) < 2:
            progress("%s pass gave no usable rows (exit %d, kernel %r, %d rows); its output ends with: %s"
                     % (counter, p.returncode, kernel, len(rows), (p.stdout + p.stderr)[-1500:]))
            return None
        scale = (COPY_CAL_MIB << 10) / (sum(copy) / len(copy))
        kib = sum(gen[1:]) / len(gen[1:])   # the first dispatch includes compile-time warm-up
        res[counter] = {"kib_per_dispatch": round(kib, 1), "copy_scale": round(scale, 3)}
    total = sum(v["kib_per_dispatch"] * v["copy_scale"] * 1024 for v in res.values())
    return {"bytes": int(total), "kernel": kernel, **res}


def progress(msg):
    """One progress line on stderr per finished phase (the JSON line stays alone on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def serving_bench(W, trace_requests=20000, rate=2000.0, saturate_requests=20000):
    """Chunk latency and throughput of the serving engine on the serving tree:
    the Poisson trace (open loop) for p50/p99 latency, and a saturating closed
    loop (512 outstanding) of the same request mix for throughput, into pinned
    host buffers and into device buffers, in the plain (one launch + copy per
    chunk), batched, graph-replay and persistent (one work-queue launch per
    batch) modes."""
    import fastnoise2_amd as fn
    from fastnoise2_amd.serve import ServeEngine, SERVE_PLAIN, SERVE_BATCHED, SERVE_GRAPH, SERVE_PERSISTENT
    enc = W.serving_tree().encode()
    specs, arrivals = W.serving_trace(trace_requests, rate)
    pts = np.array([s.count[0] * s.count[1] * (s.count[2] if s.dims == 3 else 1) for s in specs], dtype=np.float64)
    out = {"tree": enc, "trace": "Poisson %g req/s, %d requests, sizes 16^3/32^3/64^3/128^2/512^2" % (rate, trace_requests),
           "budget_bytes": 16 << 30}
    for name, mode in (("plain", SERVE_PLAIN), ("batched", SERVE_BATCHED), ("graph", SERVE_GRAPH), ("persistent", SERVE_PERSISTENT)):
        eng = ServeEngine(16 << 30, 1 << 23, mode)
        tid = eng.add_tree(enc)
        eng.replay([tid] * 64, specs[:64], None, 16)   # compile + warm up
        fn.jit_wait()   # specialised kernels ready before timing
        rec = {}
        if name != "graph":
            lat, wall = eng.replay([tid] * len(specs), specs, arrivals, 512)
            rec.update({"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                        "mean_ms": round(float(lat.mean()), 3), "trace_wall_s": round(wall, 2)})
        n = min(saturate_requests, len(specs))
        lat, wall = eng.replay([tid] * n, specs[:n], None, 512)
        rec.update({"saturated_Gpoints_s": round(float(pts[:n].sum()) / wall / 1e9, 3),
                    "saturated_requests_s": round(n / wall, 1),
                    "saturated_p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "saturated_p99_ms": round(float(np.percentile(lat, 99)), 3),
                    "batches": eng.stats()["batches"]})
        # the same closed loop with outputs left in device memory (consumers on
        # the GPU): no PCIe copy, so launch and batching costs show
        lat, wall = eng.replay([tid] * n, specs[:n], None, 512, device_outputs=True)
        rec.update({"saturated_device_Gpoints_s": round(float(pts[:n].sum()) / wall / 1e9, 3),
                    "saturated_device_p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "saturated_device_p99_ms": round(float(np.percentile(lat, 99)), 3)})
        eng.close()
        out[name] = rec
        progress("serving %s done" % name)
    out["batched_vs_plain_throughput"] = round(out["batched"]["saturated_Gpoints_s"] / out["plain"]["saturated_Gpoints_s"], 2)
    out["batched_vs_plain_device_throughput"] = round(out["batched"]["saturated_device_Gpoints_s"]
                                                      / out["plain"]["saturated_device_Gpoints_s"], 2)
    return out


def serving_mixed_bench(W, trace_requests=2000, saturate_requests=1000):
    """The mixed-tree serving workload: 8 seeded random trees (depth 3..12),
    2D/3D/4D requests of 2^12..2^22 points, bursty on/off arrivals (10 ms at
    10k req/s); per-tree tuned batched launches vs graph replay vs the
    persistent work-queue kernel vs plain."""
    import fastnoise2_amd as fn
    from fastnoise2_amd.serve import ServeEngine, SERVE_PLAIN, SERVE_BATCHED, SERVE_GRAPH, SERVE_PERSISTENT
    encs = W.mixed_trees()
    trees, specs, arrivals = W.mixed_trace(trace_requests)
    pts = np.array([float(np.prod(list(s.count)[:s.dims])) for s in specs])
    out = {"trees": len(encs), "trace": "on/off bursts 10 ms at 10k req/s, %d requests, 2^12..2^22 points, dims 2-4"
           % trace_requests, "mean_points": round(float(pts.mean()))}
    warm = [W.grid_spec_like(d) for d in (2, 3, 4)]
    for name, mode in (("plain", SERVE_PLAIN), ("batched", SERVE_BATCHED), ("graph", SERVE_GRAPH), ("persistent", SERVE_PERSISTENT)):
        eng = ServeEngine(16 << 30, 1 << 23, mode)
        tids = [eng.add_tree(e) for e in encs]
        wt = [t for t in tids for _ in warm]
        eng.replay(wt, warm * len(tids), None, 8)   # compile + tune every tree and dimensionality
        fn.jit_wait()
        rec = {}
        lat, wall = eng.replay([tids[t] for t in trees], specs, arrivals, 512)
        rec.update({"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                    "trace_wall_s": round(wall, 2)})
        n = min(saturate_requests, len(specs))
        lat, wall = eng.replay([tids[t] for t in trees[:n]], specs[:n], None, 64)
        rec.update({"saturated_Gpoints_s": round(float(pts[:n].sum()) / wall / 1e9, 3),
                    "saturated_p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "saturated_p99_ms": round(float(np.percentile(lat, 99)), 3)})
        if mode != SERVE_PLAIN:
            rec["tuned_points_per_lane"] = [[eng.tree_tuning(t, d) for d in (2, 3, 4)] for t in tids]
        eng.close()
        out[name] = rec
        progress("mixed serving %s done" % name)
    out["batched_vs_plain_throughput"] = round(out["batched"]["saturated_Gpoints_s"] / out["plain"]["saturated_Gpoints_s"], 2)
    return out


def world_bench(W, chunk_size=64, load_range=400.0, step=64.0, frames=48):
    """Chunked-world fly-through: a camera stepping along x by one chunk per
    frame through the terrain tree; each frame streams every missing chunk in
    range (nearest first, at most 128 in flight) before the camera moves on,
    so both modes deliver the same chunks (the first frame's sphere, then a
    slab per frame). Range culling and face counts run on the GPU. Chunks per
    second, their points per second and request-to-consumer latency, batched
    vs plain."""
    import fastnoise2_amd as fn
    from fastnoise2_amd.serve import ServeEngine, SERVE_PLAIN, SERVE_BATCHED
    from fastnoise2_amd.world import ChunkWorld, CHUNK_SURFACE
    enc = W.world_tree().encode()
    out = {"tree": enc, "chunk": "%d^3 (+2 margin)" % chunk_size, "load_range": load_range, "frames": frames,
           "camera_step": step}
    for name, mode in (("plain", SERVE_PLAIN), ("batched", SERVE_BATCHED)):
        eng = ServeEngine(16 << 30, 1 << 23, mode)
        tid = eng.add_tree(enc)
        world = ChunkWorld(eng, tid, chunk_size=chunk_size, iso_surface=0.0, load_range=load_range, seed=1337,
                           max_in_flight=128)
        world.update((0.0, 0.0, 0.0))
        world.poll(wait_all=True)        # compile, tune and warm up
        fn.jit_wait()
        world.new_version()
        lat, surface = [], 0
        t0 = time.perf_counter()
        for f in range(frames):
            cam = (f * step, 0.0, 0.0)
            while True:
                n = world.update(cam)
                for r in world.poll(wait_all=True):
                    lat.append(r["latency_ms"])
                    surface += r["state"] == CHUNK_SURFACE
                if n == 0:
                    break
        wall = time.perf_counter() - t0
        g = chunk_size + 2
        out[name] = {"chunks": len(lat), "chunks_s": round(len(lat) / wall, 1),
                     "Gpoints_s": round(len(lat) * g ** 3 / wall / 1e9, 3),
                     "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                     "surface_fraction": round(surface / max(1, len(lat)), 3), "wall_s": round(wall, 3)}
        world.close()
        eng.close()
        progress("world %s done" % name)
    out["batched_vs_plain_chunks"] = round(out["batched"]["chunks_s"] / out["plain"]["chunks_s"], 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="fbm_simplex_2d_grid_4096x4096")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--jit", default="sync", choices=["off", "sync"],
                    help="run-time specialised kernel of the tree (compiled before timing) or the interpreter")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the reference kernel for HBM counter scaling)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes for roofline.traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every point is independent and this build runs on a single MI355X; "
              "splitting the index space over several GPUs is designed but not run (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import fastnoise2_amd.workloads as W
    wls = W.bench_workloads()
    # counter passes first: they run in child processes, started before this
    # process touches the GPU
    traffic = None
    if not args.no_traffic and not args.pmc_child and not args.copy_calibration:
        traffic = measure_traffic(args.workload)
        progress("hbm counter passes done")
    import fastnoise2_amd as fn
    if fn.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.pmc_child:
        pmc_child(fn, W, wls[args.workload], args.workload)
        return
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src, dst = fn.DeviceArray(nbytes // 4), fn.DeviceArray(nbytes // 4)
        for _ in range(3):
            if fn.lib().fn2Memcpy(dst.ptr, src.ptr, nbytes, 2):
                raise SystemExit("copy failed: " + fn.last_error())
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)
        return
    wl = wls[args.workload]
    spec_tree = wl["tree"]()
    enc = spec_tree.encode()
    node = fn.Node.from_encoded(enc)
    keep = []
    spec = make_spec(fn, W, wl, keep)
    n = int(np.prod(wl["count"]))
    out = fn.DeviceArray(n)

    # the interpreter's figure first, then the tree's specialised kernel
    fn.set_jit_mode(fn.JIT_OFF)
    interp_ms, _ = node.bench_device(spec, out, warmup=args.warmup, iters=max(3, args.steps // 5))
    kernel_name = "fn2_generate"
    if args.jit == "sync":
        fn.set_jit_mode(fn.JIT_SYNC)
        t0 = time.perf_counter()
        node.bench_device(spec, out, warmup=0, iters=1)
        jit_compile_s = time.perf_counter() - t0
        if node.jit_state(wl["dims"] if wl["kind"] != "tileable" else 4) == 1:
            kernel_name = "fn2_jit_single"
    ms_event, ms_wall = node.bench_device(spec, out, warmup=args.warmup, iters=args.steps)
    value = n / (ms_wall * 1e-3) / 1e9
    fn.set_jit_mode(fn.JIT_ASYNC)
    progress("%s: %d timed steps done, %.2f Gpoints/s" % (args.workload, args.steps, value))

    # algorithmic work per point (counted by the op-counting oracle, committed
    # with the workload and recounted by the CPU tests)
    flop_pt, valu_pt = wl["ops"]
    kernel_s = ms_event * 1e-3
    tflops = flop_pt * n / kernel_s / 1e12
    bytes_pt = 4 + (4 * wl["dims"] if wl["kind"] == "positions" else 0)
    valu_tops = valu_pt * n / kernel_s / 1e12
    roofline = {
        # vector-ALU bound: integer hashing and per-lane FP32, no contraction
        # for MFMA and 4 bytes of HBM traffic per point
        "bound": "valu",
        "achieved": round(valu_tops, 3),
        "peak": round(VALU_PEAK_TOPS, 2),
        "unit": "T VALU lane-ops/s",
        "frac": round(valu_tops / VALU_PEAK_TOPS, 4),
        "traffic": traffic["bytes"] if traffic else None,
        "traffic_detail": traffic,
        "traffic_algorithmic_bytes": bytes_pt * n,
        "kernel": kernel_name,
        "interpreter_kernel_ms": round(interp_ms, 4),
        "kernel_ms": round(ms_event, 4),
        "valu_ops_per_point": valu_pt,
        "flop_per_point": flop_pt,
        "fp32_TFLOPs": round(tflops, 3),
        "fp32_peak_TFLOPs": FP32_PEAK_TFLOPS,
        "hbm_algorithmic_GBs": round(bytes_pt * n / kernel_s / 1e9, 1),
        "hbm_peak_GBs": HBM_PEAK_GBS,
    }
    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: grid coordinates / SplitMix64(7) positions, output device-resident",
        "config": {"workload": args.workload, "tree": enc, "points": n, "seed": wl["seed"]},
        "jit_compile_s": round(jit_compile_s, 2) if args.jit == "sync" else None,
        "roofline": roofline,
    }
    if wl["kind"] == "grid":
        # the reference's own call shape: host output buffer, PCIe copy included
        d = wl["dims"]
        gen = {2: node.gen_uniform_grid_2d, 3: node.gen_uniform_grid_3d}[d]
        gen(*wl["offset"], *wl["count"], seed=wl["seed"])
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            gen(*wl["offset"], *wl["count"], seed=wl["seed"])
        rec["api_roundtrip_Gpoints_s"] = round(n * reps / (time.perf_counter() - t0) / 1e9, 3)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(wl, enc, args.cpu_seconds)
        rec["gpu_vs_cpu"] = round(value / rec["cpu_baseline"]["value"], 1)
        progress("cpu baseline done")
    if not args.no_serving and args.workload == "fbm_simplex_2d_grid_4096x4096":
        rec["serving"] = serving_bench(W)
        rec["serving_mixed"] = serving_mixed_bench(W)
        rec["world"] = world_bench(W)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
