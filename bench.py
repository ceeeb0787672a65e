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

The default run also measures the other plain workloads the same way under
"configs" (Perlin 64^3, the tileable 1024^2 FBm image, the warped cellular
256^3 grid and 16.7 M positions, and two cheap single-node 16384^2 grids that
HBM writes bound), each with its roofline, hardware counters and host
baseline, then the serving engine and the chunked-world client.

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
SIMDS = 1024                      # 256 CUs x 4 SIMDs
XCDS = 8                          # GRBM_GUI_ACTIVE reads as the sum over the 8 accelerator dies
VALU_ISSUE_CYCLES = 2             # cycles per wave64 VALU instruction at the 78.6 T peak

HEADLINE = "fbm_simplex_2d_grid_4096x4096"
# the other plain workloads, measured under "configs" by the default run
CONFIGS = ["perlin_3d_grid_64x64x64", "fbm_simplex_tileable_1024x1024", "warped_cellular_3d_grid_256x256x256",
           "warped_cellular_positions_16777216", "constant_2d_grid_16384x16384", "checkerboard_2d_grid_16384x16384"]
# workloads whose kernel writes more than it computes: priced against HBM
HBM_BOUND = {"constant_2d_grid_16384x16384", "checkerboard_2d_grid_16384x16384"}
# counter passes: one for FETCH_SIZE (3 TCC counters), one for WRITE_SIZE
# (2 TCC) with the instruction and busy counters beside it
SQ_COUNTERS = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES",
               "SQ_INSTS_VALU_TRANS_F32", "GRBM_GUI_ACTIVE"]
COUNTER_PASSES = [["FETCH_SIZE"], ["WRITE_SIZE"] + SQ_COUNTERS]
# where the wave-cycles go (an optional third pass: the bench goes on without
# it if the counters are not there): alive, waiting on anything, waiting to
# issue, issuing
STALL_COUNTERS = ["SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_SCA",
                  "SQ_ACTIVE_INST_MISC"]


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
    """Times the oracle (16 lanes, one thread per core in this process's CPU
    affinity mask) on a bounded sample of the workload: a slice of it when
    the whole takes longer than target_s, else whole passes repeated for about
    target_s. Workloads of at most 2^22 points are too short to split over
    every core (thread start-up outweighs the work), so their passes run as
    concurrent single-threaded calls, one per core."""
    import threading
    import oracle
    threads = max(1, len(os.sched_getaffinity(0)))
    t = oracle.Tree(enc)
    total_points = int(np.prod(wl["count"]))

    def one(rows, nthreads):
        if wl["kind"] == "grid":
            count = list(wl["count"])
            count[-1] = rows
            t.grid(wl["offset"], count, seed=wl["seed"], lanes=16, threads=nthreads)
        elif wl["kind"] == "tileable":
            t.tileable(wl["count"][0], rows, seed=wl["seed"], lanes=16, threads=nthreads)
        else:
            t.positions(pos[:rows * 4096], seed=wl["seed"], lanes=16, threads=nthreads)

    def run_concurrent():
        # ctypes releases the GIL during each call, so the calls run side by
        # side; one timed single-threaded pass sets the count for target_s
        t0 = time.perf_counter()
        one(wl["count"][-1], 1)
        per = max(1, int(round(target_s / max(time.perf_counter() - t0, 1e-6))))
        workers = [threading.Thread(target=lambda: [one(wl["count"][-1], 1) for _ in range(per)]) for _ in range(threads)]
        t0 = time.perf_counter()
        for w in workers:
            w.start()
        for w in workers:
            w.join()
        dt = time.perf_counter() - t0
        what = "whole workload x %d passes as %d concurrent single-threaded calls" % (per * threads, threads)
        return per * threads * total_points, dt, what

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
    elif total_points <= 1 << 22 and threads > 1:
        n, dt, what = run_concurrent()
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


def pmc_child(fn, W, names):
    """Counter pass body: three 256 MiB device-to-device copies (the scale
    reference), then each workload's kernel four times (one launch that
    compiles the tree's specialised kernel, three more); prints one JSON line
    per workload naming its kernel."""
    nbytes = COPY_CAL_MIB << 20
    src, dst = fn.DeviceArray(nbytes // 4), fn.DeviceArray(nbytes // 4)
    for _ in range(3):
        if fn.lib().fn2Memcpy(dst.ptr, src.ptr, nbytes, 2):
            raise SystemExit("copy failed: " + fn.last_error())
    src.free()
    dst.free()
    wls = W.bench_workloads()
    for name in names:
        wl = wls[name]
        node = fn.Node.from_encoded(wl["tree"]().encode())
        keep = []
        spec = make_spec(fn, W, wl, keep)
        out = fn.DeviceArray(int(np.prod(wl["count"])))
        fn.set_jit_mode(fn.JIT_SYNC)
        node.bench_device(spec, out, warmup=0, iters=1)
        fn.set_jit_mode(fn.JIT_ASYNC)
        node.bench_device(spec, out, warmup=0, iters=3)
        dims = wl["dims"] if wl["kind"] != "tileable" else 4
        fn.lib().fn2DeviceSynchronize()
        out.free()
        for d in keep:
            d.free()
        print(json.dumps({"workload": name, "kernel": "fn2_jit_single" if node.jit_state(dims) == 1 else "fn2_generate",
                          "dispatches": 4}), flush=True)


def _is_generator(kernel_name):
    return kernel_name.startswith("fn2_jit_single") or kernel_name.startswith("void fn2_generate<")


def measure_counters(names):
    """Hardware counters per launch of each workload's kernel, from
    rocprofv3 --pmc passes over one child run of all the workloads each: HBM
    bytes (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch, each scaled by
    what the same counter reads for the 256 MiB copy, which moves 256 MiB each
    way) with the instruction / busy counters, then the wave-cycle counters.
    {} when rocprofv3 is missing or one of the first two passes fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return {}
    res = {n: {} for n in names}
    for counters in COUNTER_PASSES + [STALL_COUNTERS]:
        optional = counters is STALL_COUNTERS
        d = tempfile.mkdtemp(prefix="fn2_pmc_")
        cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc"] + counters + [
            "--kernel-trace", "-f", "csv", "-d", d, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__),
            "--pmc-child", ",".join(names)]
        p = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, FN2_JIT="async"))
        order = [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]
        rows = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    with open(os.path.join(root, f)) as fh:
                        rows += list(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        # counter value per (dispatch, counter), dispatches in launch order
        per = {}
        for r in rows:
            key = int(r["Dispatch_Id"])
            e = per.setdefault(key, {"kernel": r["Kernel_Name"], "grid": int(r["Grid_Size"]), "c": {}})
            e["c"][r["Counter_Name"]] = e["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        disp = [per[k] for k in sorted(per)]
        copies = [e for e in disp if e["kernel"].startswith("__amd_rocclr_copyBuffer") and e["grid"] >= 1 << 16]
        gens = [e for e in disp if _is_generator(e["kernel"])]
        if len(order) != len(names) or not copies or len(gens) != sum(o["dispatches"] for o in order):
            progress("counter pass %s gave no usable rows (exit %d, %d workloads, %d generator dispatches); its output "
                     "ends with: %s" % (counters[0], p.returncode, len(order), len(gens), (p.stdout + p.stderr)[-1500:]))
            if optional:
                continue
            return {}
        at = 0
        for o in order:
            mine = gens[at:at + o["dispatches"]][1:]   # the first dispatch includes compile-time warm-up
            at += o["dispatches"]
            rec = res[o["workload"]]
            rec["kernel"] = o["kernel"]
            for c in counters:
                v = sum(e["c"].get(c, 0.0) for e in mine) / len(mine)
                if c in ("FETCH_SIZE", "WRITE_SIZE"):
                    scale = (COPY_CAL_MIB << 10) / (sum(e["c"][c] for e in copies) / len(copies))
                    rec[c] = {"kib_per_dispatch": round(v, 1), "copy_scale": round(scale, 3)}
                else:
                    rec.setdefault("counters", {})[c] = round(v)
    for rec in res.values():
        if "FETCH_SIZE" in rec and "WRITE_SIZE" in rec:
            rec["bytes"] = int(sum(rec[c]["kib_per_dispatch"] * rec[c]["copy_scale"] * 1024 for c in ("FETCH_SIZE", "WRITE_SIZE")))
    return res


def roofline_for(name, wl, n, kernel_ms, counters, kernel_name, interp_ms=None):
    """Roofline figures of one workload's kernel: algorithmic work per point
    (counted by the op-counting oracle, committed with the workload) over the
    event-timed launch, against the VALU issue peak, or against HBM for the
    workloads bound by their writes; plus what the counters show the
    hardware issued and how busy its vector ALUs were."""
    flop_pt, valu_pt = wl["ops"]
    kernel_s = kernel_ms * 1e-3
    bytes_pt = 4 + (4 * wl["dims"] if wl["kind"] == "positions" else 0)
    valu_tops = valu_pt * n / kernel_s / 1e12
    hbm_gbs = bytes_pt * n / kernel_s / 1e9
    if name in HBM_BOUND:
        r = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
    else:
        # vector-ALU bound: integer hashing and per-lane FP32, no contraction
        # for MFMA and 4 bytes of HBM traffic per point
        r = {"bound": "valu", "achieved": round(valu_tops, 3), "peak": round(VALU_PEAK_TOPS, 2),
             "unit": "T VALU lane-ops/s", "frac": round(valu_tops / VALU_PEAK_TOPS, 4)}
    r.update({
        "traffic": counters.get("bytes"),
        "traffic_algorithmic_bytes": bytes_pt * n,
        "kernel": kernel_name,
        "kernel_ms": round(kernel_ms, 4),
        "valu_ops_per_point": valu_pt,
        "flop_per_point": flop_pt,
        "fp32_TFLOPs": round(flop_pt * n / kernel_s / 1e12, 3),
        "fp32_peak_TFLOPs": FP32_PEAK_TFLOPS,
        "hbm_algorithmic_GBs": round(hbm_gbs, 1),
        "hbm_peak_GBs": HBM_PEAK_GBS,
    })
    if wl.get("pruned"):
        r["note"] = ("ops per point count the reference's fixed 27-cell walk; the kernel's pruned walk leaves out "
                     "cells that cannot change the result, so frac can exceed 1: hw.issue_frac is the share of the "
                     "issue peak the kernel used")
    if interp_ms is not None:
        r["interpreter_kernel_ms"] = round(interp_ms, 4)
    if counters.get("FETCH_SIZE"):
        r["traffic_detail"] = {c: counters[c] for c in ("FETCH_SIZE", "WRITE_SIZE") if c in counters}
        if counters.get("WRITE_SIZE"):
            w = counters["WRITE_SIZE"]
            r["hbm_write_GBs_counted"] = round(w["kib_per_dispatch"] * w["copy_scale"] * 1024 / kernel_s / 1e9, 1)
    c = counters.get("counters")
    if c and c.get("SQ_INSTS_VALU"):
        # SQ_INSTS_VALU counts wave64 instructions: 64 lane operations each
        issued = c["SQ_INSTS_VALU"] * 64 / kernel_s / 1e12
        cycles = c.get("GRBM_GUI_ACTIVE", 0) / XCDS
        if kernel_ms < 0.1:
            # a profiled dispatch's active cycles include its launch and drain
            # (several microseconds): for short kernels they say nothing of
            # the kernel's own busy share or clock
            cycles = 0
        r["hw"] = {
            "valu_instr_per_point": round(c["SQ_INSTS_VALU"] * 64 / n, 1),
            "issued_T_lane_ops_s": round(issued, 3),
            "issue_frac": round(issued / VALU_PEAK_TOPS, 4),
            # share of the kernel's cycles each SIMD spent issuing VALU work
            "valu_busy": round(c["SQ_INSTS_VALU"] * VALU_ISSUE_CYCLES / SIMDS / cycles, 4) if cycles else None,
            "clock_GHz": round(cycles / kernel_s / 1e9, 3) if cycles else None,
            "counters": c,
        }
        if c.get("SQ_WAVE_CYCLES"):
            # shares of the waves' lifetime (wave-cycles): waiting on anything
            # (a dependency, memory, a barrier), waiting for an issue slot,
            # and issuing VALU / scalar instructions
            wc = float(c["SQ_WAVE_CYCLES"])
            r["hw"]["wave_cycles"] = {k: round(c[n] / wc, 4) for k, n in (
                ("wait_any", "SQ_WAIT_ANY"), ("wait_inst_any", "SQ_WAIT_INST_ANY"), ("active_inst_any", "SQ_ACTIVE_INST_ANY"),
                ("active_inst_valu", "SQ_ACTIVE_INST_VALU"), ("active_inst_scalar", "SQ_ACTIVE_INST_SCA"),
                ("active_inst_misc", "SQ_ACTIVE_INST_MISC")) if n in c}
    return r


def progress(msg):
    """One progress line on stderr per finished phase (the JSON line stays alone on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def _clock_marks():
    return {"monotonic": time.clock_gettime_ns(time.CLOCK_MONOTONIC),
            "boottime": time.clock_gettime_ns(time.CLOCK_BOOTTIME)}


def window(name, t0, t1, requests=None):
    """Names a phase's wall-clock window on stderr, in both host clocks a
    kernel trace may be stamped in, so that --account can sum the kernels
    that ran inside it; `requests`: how many requests of the mixed trace
    (from its start) the phase served."""
    js = {k: [t0[k], t1[k]] for k in t0}
    if requests is not None:
        js["requests"] = requests
    print("bench-window %s %s" % (name, json.dumps(js)), file=sys.stderr, flush=True)


def account(trace_csv, log):
    """Device-time accounting of the phases a bench run named (bench-window
    lines in `log`) from a rocprofv3 --kernel-trace CSV of the same run:
    kernel count, the sum of kernel durations, the time at least one kernel
    was running (overlapping kernels on several streams counted once), both
    against the window's wall time, and the idle gaps between busy spans.
    For a window over the mixed trace's requests, also the algorithmic vector
    ALU work of those requests (committed op counts per tree and
    dimensionality) against the VALU issue peak over the wall, the busy time
    and the kernel time; in plain mode (one launch per request, in submission
    order) per tree as well."""
    import csv
    import fastnoise2_amd.workloads as W
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    wins = []
    for line in open(log):
        if line.startswith("bench-window "):
            _, name, js = line.split(" ", 2)
            wins.append((name, json.loads(js)))
    out = []
    mixed = None
    for name, clocks in wins:
        nreq = clocks.pop("requests", None)
        # the clock whose window holds the most kernels is the trace's clock
        best = max(clocks, key=lambda c: sum(1 for s, e, _ in rows if clocks[c][0] <= s and e <= clocks[c][1]))
        t0, t1 = clocks[best]
        ks = [(s, e, k) for s, e, k in rows if t0 <= s and e <= t1]
        busy, gaps, cur_s, cur_e = 0, [], None, None
        for s, e, _ in ks:
            if cur_e is None or s > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                    gaps.append(s - cur_e)
                cur_s, cur_e = s, e
            else:
                cur_e = max(cur_e, e)
        if cur_e is not None:
            busy += cur_e - cur_s
        wall = t1 - t0
        total = sum(e - s for s, e, _ in ks)
        by = {}
        for s, e, k in ks:
            key = k.split("(")[0]
            c = by.setdefault(key, [0, 0])
            c[0] += 1
            c[1] += e - s
        g = np.array(gaps, dtype=np.float64) if gaps else np.zeros(1)
        alg = {}
        if nreq:
            if mixed is None:
                mixed = W.mixed_trace(2000)
            trees, specs = mixed[0][:nreq], mixed[1][:nreq]
            pts = np.array([float(np.prod(list(sp.count)[:sp.dims])) for sp in specs])
            valu = np.array([W.MIXED_TREE_OPS[t][sp.dims][1] for t, sp in zip(trees, specs)]) * pts
            peak = VALU_PEAK_TOPS * 1e12
            gen = [(s, e) for s, e, k in ks if _is_generator(k) or k.startswith("fn2_jit_batch")]
            alg = {"requests": nreq, "Gpoints": round(float(pts.sum()) / 1e9, 4),
                   "valu_frac_over_wall": round(float(valu.sum()) / (wall * 1e-9) / peak, 3),
                   "valu_frac_over_busy": round(float(valu.sum()) / (busy * 1e-9) / peak, 3) if busy else None,
                   "valu_frac_over_kernel_sum": round(float(valu.sum()) / (total * 1e-9) / peak, 3) if total else None}
            if name.endswith("-plain") and len(gen) == nreq:
                dur = np.array([(e - s) * 1e-9 for s, e in gen])
                tr = np.array(trees)
                alg["per_tree"] = [{"tree": t, "requests": int((tr == t).sum()),
                                    "kernel_ms": round(float(dur[tr == t].sum()) * 1e3, 3),
                                    "Gpoints_s": round(float(pts[tr == t].sum()) / float(dur[tr == t].sum()) / 1e9, 3),
                                    "valu_frac": round(float(valu[tr == t].sum()) / float(dur[tr == t].sum()) / peak, 3)}
                                   for t in range(len(W.MIXED_TREE_OPS)) if (tr == t).any()]
        out.append({"window": name, "clock": best, "wall_ms": round(wall / 1e6, 3), "kernels": len(ks),
                    "kernel_sum_ms": round(total / 1e6, 3), "busy_ms": round(busy / 1e6, 3),
                    "kernel_sum_over_wall": round(total / wall, 3), "busy_over_wall": round(busy / wall, 3),
                    "gaps": len(gaps), "gap_ms_total": round(float(g.sum()) / 1e6, 3),
                    "gap_us_p50": round(float(np.percentile(g, 50)) / 1e3, 1),
                    "gap_us_p99": round(float(np.percentile(g, 99)) / 1e3, 1),
                    "by_kernel": {k: {"count": c[0], "ms": round(c[1] / 1e6, 3)} for k, c in
                                  sorted(by.items(), key=lambda kv: -kv[1][1])[:8]}, **alg})
    return out


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
        # compile + warm up, then tune the tile on the specialised kernel
        for _ in range(2):
            eng.replay([tid] * 64, specs[:64], None, 16)
            fn.jit_wait()
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


def serving_mixed_bench(W, trace_requests=2000, saturate_requests=1000, modes=("plain", "batched", "graph", "persistent")):
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
        if name not in modes:
            continue
        eng = ServeEngine(16 << 30, 1 << 23, mode)
        tids = [eng.add_tree(e) for e in encs]
        wt = [t for t in tids for _ in warm]
        # compile and tune every tree and dimensionality: the interpreter's
        # tile, then (specialised kernels loaded) the specialised tile
        for _ in range(2):
            eng.replay(wt, warm * len(tids), None, 8)
            fn.jit_wait()
        rec = {}
        lat, wall = eng.replay([tids[t] for t in trees], specs, arrivals, 512)
        rec.update({"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                    "trace_wall_s": round(wall, 2)})
        n = min(saturate_requests, len(specs))
        w0 = _clock_marks()
        lat, wall = eng.replay([tids[t] for t in trees[:n]], specs[:n], None, 64)
        window("mixed-saturated-" + name, w0, _clock_marks(), requests=n)
        rec.update({"saturated_Gpoints_s": round(float(pts[:n].sum()) / wall / 1e9, 3),
                    "saturated_p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "saturated_p99_ms": round(float(np.percentile(lat, 99)), 3)})
        if mode != SERVE_PLAIN:
            # per tree, per dimensionality 2/3/4: threads per block, points per
            # lane, octave unroll (0 = the compiler's choice; the unroll stage
            # is off by default)
            rec["tuned_tiles"] = [[eng.tree_tile(t, d) for d in (2, 3, 4)] for t in tids]
        eng.close()
        out[name] = rec
        progress("mixed serving %s done" % name)
    if "plain" in out and "batched" in out:
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


def time_workload(fn, W, name, wl, steps, warmup, jit=True):
    """Device-resident throughput of one workload: the interpreter's event time
    per launch, then the tree's specialised kernel (compiled before timing)
    over `steps` launches. Returns (Gpoints/s, event ms, interpreter ms,
    kernel name, node)."""
    node = fn.Node.from_encoded(wl["tree"]().encode())
    keep = []
    spec = make_spec(fn, W, wl, keep)
    n = int(np.prod(wl["count"]))
    out = fn.DeviceArray(n)
    fn.set_jit_mode(fn.JIT_OFF)
    interp_ms, _ = node.bench_device(spec, out, warmup=min(warmup, 2), iters=max(3, steps // 5))
    kernel_name = "fn2_generate"
    if jit:
        fn.set_jit_mode(fn.JIT_SYNC)
        node.bench_device(spec, out, warmup=0, iters=1)
        if node.jit_state(wl["dims"] if wl["kind"] != "tileable" else 4) == 1:
            kernel_name = "fn2_jit_single"
    ms_event, ms_wall = node.bench_device(spec, out, warmup=warmup, iters=steps)
    fn.set_jit_mode(fn.JIT_ASYNC)
    out.free()
    for d in keep:
        d.free()
    return n / (ms_wall * 1e-3) / 1e9, ms_event, interp_ms, kernel_name, node


def api_roundtrip(fn, W, node, wl, reps=5):
    """The reference's own call shape: GenUniformGrid / GenTileable2D /
    GenPositionArray3D into a caller-owned pageable host buffer (allocated and
    touched once, as a caller reuses it), host-to-device position copies and
    the device-to-host output copy included. Gpoints/s over `reps` calls."""
    import ctypes
    n = int(np.prod(wl["count"]))
    out = np.zeros(n, dtype=np.float32)
    mm = np.zeros(2, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    o, m = out.ctypes.data_as(fp), mm.ctypes.data_as(fp)
    L = fn.lib()
    if wl["kind"] == "grid":
        off, cnt, seed = wl["offset"], wl["count"], wl["seed"]
        if wl["dims"] == 2:
            call = lambda: L.fnGenUniformGrid2D(node.handle, o, *map(float, off), *cnt, 1.0, 1.0, seed, m)
        else:
            call = lambda: L.fnGenUniformGrid3D(node.handle, o, *map(float, off), *cnt, 1.0, 1.0, 1.0, seed, m)
    elif wl["kind"] == "tileable":
        call = lambda: L.fnGenTileable2D(node.handle, o, wl["count"][0], wl["count"][1], 1.0, 1.0, wl["seed"], m)
    else:
        xyz = [np.ascontiguousarray(a) for a in W.splitmix64_positions(n, seed=7)]
        px, py, pz = (a.ctypes.data_as(fp) for a in xyz)
        call = lambda: L.fnGenPositionArray3D(node.handle, o, n, px, py, pz, 0.0, 0.0, 0.0, wl["seed"], m)
    call()
    if fn.last_error():
        raise RuntimeError("API call failed: " + fn.last_error())
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    return round(n * reps / (time.perf_counter() - t0) / 1e9, 3)


def config_record(fn, W, name, wl, counters, args):
    """One plain workload: throughput, roofline, counters, API round trip and
    host baseline."""
    n = int(np.prod(wl["count"]))
    # enough launches for about 0.3 s of GPU time, at least 10
    value, ms, interp_ms, kname, node = time_workload(fn, W, name, wl, 3, 1)
    steps = int(min(2000, max(10, 0.3 / max(ms * 1e-3, 1e-6))))
    value, ms, interp_ms, kname, node = time_workload(fn, W, name, wl, steps, 3)
    rec = {"value": round(value, 3), "unit": "Gpoints/s", "points": n, "steps": steps, "kernel_ms": round(ms, 4),
           "tree": wl["tree"]().encode(), "seed": wl["seed"],
           "roofline": roofline_for(name, wl, n, ms, counters.get(name, {}), kname, interp_ms)}
    rec["api_roundtrip_Gpoints_s"] = api_roundtrip(fn, W, node, wl, reps=5 if n >= 1 << 20 else 50)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(wl, rec["tree"], args.config_cpu_seconds)
        rec["gpu_vs_cpu"] = round(value / rec["cpu_baseline"]["value"], 1)
    progress("config %s: %.2f Gpoints/s, kernel %.4f ms" % (name, value, ms))
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=HEADLINE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--config-cpu-seconds", type=float, default=4.0,
                    help="host baseline sample per workload of the configs section")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-mixed-only", default=None, metavar="MODES",
                    help="only the mixed-tree serving replays, in the comma-separated modes "
                         "(plain,batched,graph,persistent); for kernel-trace accounting")
    ap.add_argument("--no-configs", action="store_true", help="skip the other plain workloads")
    ap.add_argument("--jit", default="sync", choices=["off", "sync"],
                    help="run-time specialised kernel of the tree (compiled before timing) or the interpreter")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the reference kernel for HBM counter scaling)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--pmc-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--account", nargs=2, metavar=("KERNEL_TRACE_CSV", "BENCH_STDERR"), default=None,
                    help="no GPU: device-time accounting of the phases a bench run named on stderr, "
                         "from a rocprofv3 --kernel-trace CSV of that run")
    args = ap.parse_args()
    if args.account:
        print(json.dumps(account(*args.account), indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every point is independent and this build runs on a single MI355X; "
              "splitting the index space over several GPUs is designed but not run (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import fastnoise2_amd.workloads as W
    if args.serving_mixed_only:
        print(json.dumps({"serving_mixed": serving_mixed_bench(W, modes=args.serving_mixed_only.split(","))}), flush=True)
        return
    wls = W.bench_workloads()
    with_configs = args.workload == HEADLINE and not args.no_configs
    names = [args.workload] + (CONFIGS if with_configs else [])
    # counter passes first: they run in child processes, started before this
    # process touches the GPU
    counters = {}
    if not args.no_traffic and not args.pmc_child and not args.copy_calibration:
        counters = measure_counters(names)
        progress("counter passes done")
    import fastnoise2_amd as fn
    if fn.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.pmc_child:
        pmc_child(fn, W, args.pmc_child.split(","))
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
    enc = wl["tree"]().encode()
    n = int(np.prod(wl["count"]))
    t0 = time.perf_counter()
    value, ms_event, interp_ms, kernel_name, node = time_workload(fn, W, args.workload, wl, args.steps, args.warmup,
                                                                   jit=args.jit == "sync")
    ms_wall = n / (value * 1e9) * 1e3
    progress("%s: %d timed steps done, %.2f Gpoints/s" % (args.workload, args.steps, value))
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
        "roofline": roofline_for(args.workload, wl, n, ms_event, counters.get(args.workload, {}), kernel_name, interp_ms),
    }
    rec["api_roundtrip_Gpoints_s"] = api_roundtrip(fn, W, node, wl)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(wl, enc, args.cpu_seconds)
        rec["gpu_vs_cpu"] = round(value / rec["cpu_baseline"]["value"], 1)
        progress("cpu baseline done")
    if with_configs:
        rec["configs"] = {name: config_record(fn, W, name, wls[name], counters, args) for name in CONFIGS}
    if not args.no_serving and args.workload == HEADLINE:
        rec["serving"] = serving_bench(W)
        rec["serving_mixed"] = serving_mixed_bench(W)
        rec["world"] = world_bench(W)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
