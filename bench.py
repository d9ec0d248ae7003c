"""Closest-point throughput on one MI355X.

Workload (one "step" = one pass of the closest-point kernel over the batch):
a seeded torus R=1, r=0.3 tessellated to 1,048,576 float32 triangles (1024 x 512
grid, per-vertex jitter 1e-3) and 16,777,216 closest-point queries (half uniform
in the 1.2x AABB, half within 0.05 of the surface), squared max radius inf.
Queries are resident in HBM as GPUBoundingSphere records before the timed
region; results are written as GPUInteraction records (the drop-in layout).

The printed JSON line carries
  value        forward closest-point queries/s (whole job, one GPU)
  fwd_bwd      the differentiable step on the same queries: compact forward
               (d, p, primitive, leaf slot) + the adjoint kernel writing dL/dx and
               dL/dV (dL/dd = 1); queries/s and its ratio to the forward rate
  rays         16,777,216 rays (origins uniform in the AABB, directions uniform
               on S^2, tMax = inf) through the ray-intersection kernel
  sphere_intersection  16,777,216 spheres of radius 0.1 at the query points
               through the stochastic sphere-intersection kernel
  closest_silhouette   every 4th query (4,194,304) through the closest
               silhouette point kernel, on the torus built with silhouettes
  bvh_build    the torus' SAH tree built on the device against the host
               build + transfer (the same tree either way)
  instances    a scene of 256 transformed instances of a 20,480-triangle
               icosphere, 4,194,304 closest-point queries through the two-level
               kernel, against the same triangles baked into one mesh
  end_to_end   GPUScene3D.find_closest_points(points, r2, hits) on the same
               16M queries from host memory: packing, upload, ordering, kernel
               and the copy of the records back into the list (median of 3)
  star2d       the 2D workload: a 262,144-segment star polygon, 8,388,608
               queries, plus rays and ten refit steps after vertex perturbations
  diff         the differentiable loop: a 200,000-triangle geodesic sphere,
               4,194,304 near-surface queries, 100 steps of vertex upload +
               device refit + forward + adjoint + Adam
  batch        256 meshes (1k-500k triangles, 1k-1M queries each, log-uniform)
               in one persistent launch replayed from a HIP graph, against 256
               separate launches; and the batch's fwd+bwd
  roofline     algorithmic bytes per query x queries per launch / the traversal
               kernel's mean launch duration (HIP events on the launch stream),
               against MI355X HBM bandwidth.  Bytes per query
               B = 16 + 40 + 32 (1 + 2 I) + 40 P, with I (records expanded) and P
               (primitives tested) counted by the kernel's own counting variant
               over the whole batch.  traffic = measured HBM bytes per launch from
               the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (null if absent)
  cpu_baseline the reference CPU path timed on this host's cores over a strided
               sample of the same queries (both halves), median of 5 runs:
               "scalar", the restatement of Scene<3>::findClosestPoints with the
               binary SAH BVH, and "vectorized", the MBVH aggregate (4-wide nodes,
               SIMD leaf groups); each with the thread count, a 1-thread figure on
               a smaller sample, and the host record (CPU model, sockets, SMT)

Options: --legs selects the extra legs, --opt KEY=VALUE sets a GPUScene option on
every scene, --workload runs one of the other configs alone, and the
--summarize-* modes turn rocprofv3 databases into the text summaries kept under
profiles/.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "closest-point queries/s, 1M-tri mesh, fwd and fwd+bwd; % HBM roofline vs CPU"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "closest_points_coop__torus_1M_tris_16M_queries__traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=("torus", "batch", "star2d", "diff"), default="torus",
                   help="torus: 1M-tri torus, 16M queries (the headline line); batch: 256 uneven meshes; "
                        "star2d: 262,144-segment polygon with refit; diff: 200k-tri differentiable distance")
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--queries", type=int, default=1 << 24)
    p.add_argument("--grid", type=int, nargs=2, default=[1024, 512], help="torus U V (2UV triangles)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=1 << 19, help="queries timed on the CPU path")
    p.add_argument("--no-extra", action="store_true", help="forward only (skip fwd+bwd, rays and the other configs)")
    p.add_argument("--legs", default="all",
                   help="comma-separated extra legs of the headline run: fwd_bwd, e2e, rays, spheres, silhouettes, "
                        "build, instances, configs (star2d / diff / batch sub-objects); default all")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="after the timed work, copy a buffer of this size 3x (counter calibration)")
    for k in ("block", "stack", "sort", "coop", "hold_min", "xcds"):
        p.add_argument("--" + k.replace("_", "-"), type=int, default=None)
    p.add_argument("--opt", action="append", default=[], metavar="KEY=VALUE",
                   help="a GPUScene option set on every scene the run builds (repeatable)")
    p.add_argument("--summarize-kernels", metavar="DB", help="write a kernel-trace summary and exit")
    p.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                   help="write the HBM traffic summary from two counter runs and exit")
    p.add_argument("--summarize-counters", nargs="+", metavar="DB",
                   help="write a per-kernel summary of rocprofv3 --pmc runs (one counter group each) and exit")
    p.add_argument("--kernel", action="append", default=[], help="kernel name pattern (regex) for --summarize-counters")
    p.add_argument("--workload-note", default="", help="workload line of a summary")
    p.add_argument("--calib-bytes", type=int, default=0, help="bytes of the calibration copy (summaries)")
    p.add_argument("--command", action="append", default=[], help="command lines recorded in a summary")
    p.add_argument("--out", default=None, help="summary file")
    return p.parse_args()


EXTRA_OPTS = []


def apply_opts(g):
    """the --opt settings, on a freshly transferred GPUScene"""
    for kv in EXTRA_OPTS:
        k, v = kv.split("=")
        g.set_option(k, int(v))
    return g


def algorithmic_bytes(n_internal, n_prims, sphere_in=16, record_out=40, node_bytes=32, prim_bytes=40):
    """B = in + out + 32 (1 + 2 I) + 40 P per query: the root box, two child
    boxes per expanded internal node and one primitive record per primitive
    test, at the reference GPU record sizes (bvh_interop_structures.h:55-119)."""
    return sphere_in + record_out + node_bytes * (1.0 + 2.0 * n_internal) + prim_bytes * n_prims


def read_traffic():
    """HBM bytes per launch of the forward kernel from the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                m = re.match(r"\s*traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", line)
                if m:
                    return float(m.group(1))
    except OSError:
        pass
    return None


def host_record():
    """The host the CPU baseline runs on: model, sockets, cores, SMT, usable threads."""
    rec = {"model": None, "sockets": None, "physical_cores": None, "smt": None,
           "online_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
           "simd": None}
    try:
        phys, cores, flags = set(), set(), ""
        with open("/proc/cpuinfo") as fh:
            cur = {}
            for line in fh:
                if ":" not in line:
                    if cur:
                        phys.add(cur.get("physical id"))
                        cores.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name" and rec["model"] is None:
                    rec["model"] = v
                if k == "flags" and not flags:
                    flags = v
        rec["sockets"] = len(phys)
        rec["physical_cores"] = len(cores)
        rec["smt"] = (rec["online_cpus"] or 0) > len(cores) > 0
        f = set(flags.split())
        rec["simd"] = "avx512" if {"avx512f", "avx512vl", "avx512dq"} <= f else ("avx2" if "avx2" in f else "none")
    except OSError:
        pass
    return rec


def _median_rate(fn, n, runs):
    """n / median time of `runs` calls of fn, after one warm-up call"""
    fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return n / float(np.median(times)), [round(n / t, 1) for t in times]


def cpu_baseline(V, F, q, r2, sample, runs=5):
    """The reference CPU path on this host's cores, BVH build excluded: (1) the
    restated non-vectorized Scene<3>::findClosestPoints (binary SAH BVH, static
    thread partition as fcpw.inl:1547-1559) and (2) the restated vectorized
    aggregate (4-wide MBVH with SIMD leaf groups, mbvh.inl / wide_query_operations.h),
    each on all usable threads and on one thread.  The sample is every k-th query
    of the whole batch, so it covers the uniform and the near-surface half in
    their proportions; the one-thread sample is 1/16 of it."""
    import oracle
    threads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30)
    n = min(sample, len(q))
    stride = max(1, len(q) // n)
    qs, rs = np.ascontiguousarray(q[::stride][:n]), np.ascontiguousarray(r2[::stride][:n])
    n1 = max(1, n // 16)
    q1, r1 = np.ascontiguousarray(qs[::16][:n1]), np.ascontiguousarray(rs[::16][:n1])
    t0 = time.time()
    ob = oracle.Bvh(V, F)
    log(f"[cpu] scalar BVH built in {time.time() - t0:.1f}s")
    t0 = time.time()
    vb = oracle.VectorizedBvh(V, F)
    log(f"[cpu] vectorized BVH ({vb.width} lanes, {vb.info()}) built in {time.time() - t0:.1f}s")
    res = {}
    for kind, fn in (("scalar", lambda a, b, t: ob.closest_points_cpu(a, b, nthreads=t)),
                     ("vectorized", lambda a, b, t: vb.closest_points(a, b, nthreads=t))):
        v_all, runs_all = _median_rate(lambda: fn(qs, rs, threads), n, runs)
        v_one, runs_one = _median_rate(lambda: fn(q1, r1, 1), n1, runs)
        res[kind] = {"value": v_all, "threads": threads, "value_1_thread": v_one, "runs": runs_all,
                     "runs_1_thread": runs_one}
        log(f"[cpu] {kind}: {v_all:.3e} queries/s on {threads} threads, {v_one:.3e} on 1 thread")
    # the two paths answer the same (the reference's own vectorized check, |dd| <= 1e-6)
    a, b = ob.closest_points_cpu(q1, r1, nthreads=threads), vb.closest_points(q1, r1, nthreads=threads)
    dd = float(np.abs(a["d"].astype(np.float64) - b["d"]).max())
    best = max(("scalar", "vectorized"), key=lambda k: res[k]["value"])
    return {"value": res[best]["value"], "unit": "queries/s", "cores": threads, "kind": "port",
            "path": best, "scalar": res["scalar"],
            "vectorized": {**res["vectorized"], "simd_width": vb.width, "mbvh": vb.info()},
            "max_abs_d_scalar_vs_vectorized": dd, "host": host_record(),
            "sample": f"every {stride}th of the same queries ({n}; uniform and near-surface halves alike), "
                      f"{n1} of those for the 1-thread figure, median of {runs} runs after a warm-up run, "
                      f"BVH build excluded; value = the faster path ({best}) on {threads} threads"}


# ---------------------------------------------------------------------------
# rocprofv3 summaries (no GPU needed): turn the SQLite output of a kernel-trace
# run or of FETCH_SIZE / WRITE_SIZE counter runs into the text kept in profiles/
# ---------------------------------------------------------------------------
MAIN_KERNEL = r"k_closest_points_coop<3, \d+, 0, false>"  # any LDS stack depth


def _db(path):
    import glob
    import sqlite3
    if os.path.isdir(path):
        found = sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))
        if not found:
            raise SystemExit(f"no rocprofv3 database under {path}")
        path = found[0]
    elif not os.path.isfile(path):
        raise SystemExit(f"no rocprofv3 database at {path}")
    return sqlite3.connect(path), path


def summarize_kernels(db, command, out):
    con, path = _db(db)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration), sum(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[5] for r in rows) or 1
    lines = [f"# rocprofv3 kernel trace summary ({os.path.basename(path)})", f"# command: {command}", "",
             f"{'calls':>6} {'avg_us':>12} {'min_us':>12} {'max_us':>12} {'total_%':>8}  kernel"]
    for name, n, avg, mn, mx, tot in rows:
        lines.append(f"{n:6d} {avg / 1e3:12.3f} {mn / 1e3:12.3f} {mx / 1e3:12.3f} {100 * tot / total:8.2f}  {name}")
    _emit(lines, out)


def _counter(con, counter, pattern, pick_largest_grid=False):
    rows = con.execute("select dispatch_id, kernel_name, grid_size, sum(value) from counters_collection "
                       "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
    rows = [r for r in rows if re.search(pattern, r[1])]
    if pick_largest_grid and rows:
        g = max(r[2] for r in rows)
        rows = [r for r in rows if r[2] == g]
    if not rows:
        raise SystemExit(f"no dispatch of /{pattern}/ with counter {counter}")
    return float(np.mean([r[3] for r in rows])), len(rows), rows[0][1]


def summarize_traffic(fetch_db, write_db, calib_bytes, commands, out):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the forward kernel, each
    scaled by calib_bytes / (the counter's reading for a copy of calib_bytes)."""
    fcon, fpath = _db(fetch_db)
    wcon, wpath = _db(write_db)
    pat = MAIN_KERNEL
    calib = r"copyBuffer|neg_kernel|direct_copy_kernel"  # read N bytes, write N bytes
    kf, nf, name = _counter(fcon, "FETCH_SIZE", pat)
    kw, nw, _ = _counter(wcon, "WRITE_SIZE", pat)
    cf, ncf, cname = _counter(fcon, "FETCH_SIZE", calib, pick_largest_grid=True)
    cw, ncw, _ = _counter(wcon, "WRITE_SIZE", calib, pick_largest_grid=True)
    ff, fw = calib_bytes / (cf * 1024.0), calib_bytes / (cw * 1024.0)
    fetch_b, write_b = kf * 1024.0 * ff, kw * 1024.0 * fw
    lines = ["# HBM traffic per launch of the forward closest-point kernel",
             "# workload: torus 1,048,576 tris, 16,777,216 queries (bench.py defaults)"]
    lines += [f"# command: {c}" for c in commands]
    lines += ["", f"kernel: {name}",
              f"dispatches: {nf} (FETCH_SIZE pass), {nw} (WRITE_SIZE pass)",
              f"FETCH_SIZE_kib_per_dispatch: {kf:.1f}",
              f"WRITE_SIZE_kib_per_dispatch: {kw:.1f}",
              f"calibration: {cname[:90]}, {ncf}+{ncw} dispatches copying {calib_bytes} bytes",
              f"calibration_FETCH_SIZE_kib: {cf:.1f}  -> fetch factor {ff:.4f}",
              f"calibration_WRITE_SIZE_kib: {cw:.1f}  -> write factor {fw:.4f}",
              f"fetch_bytes_per_launch: {fetch_b:.4e}",
              f"write_bytes_per_launch: {write_b:.4e}",
              f"traffic_bytes_per_launch: {fetch_b + write_b:.6e}"]
    _emit(lines, out)


def summarize_counters(dbs, kernels, commands, workload, out):
    """Per-dispatch counter values (mean over the dispatches of each kernel with
    the largest grid) from several rocprofv3 --pmc runs, one counter group per
    run, for each kernel pattern; plus the L2 hit rate from TCC_HIT / TCC_MISS."""
    lines = ["# rocprofv3 counters per dispatch", f"# workload: {workload}"]
    lines += [f"# command: {c}" for c in commands]
    for pat in kernels:
        vals, name = {}, None
        for db in dbs:
            con, _ = _db(db)
            names = [r[0] for r in con.execute("select distinct counter_name from counters_collection")]
            for cn in names:
                try:
                    v, n, kname = _counter(con, cn, pat, pick_largest_grid=True)
                except SystemExit:
                    continue
                vals[cn] = (v, n)
                name = kname
        if not vals:
            continue
        lines += ["", f"kernel: {name}"]
        for cn in sorted(vals):
            v, n = vals[cn]
            lines.append(f"  {cn:28s} {v:16.6g}   ({n} dispatches)")
        if "TCC_HIT_sum" in vals and "TCC_MISS_sum" in vals:
            h, m = vals["TCC_HIT_sum"][0], vals["TCC_MISS_sum"][0]
            lines.append(f"  {'L2_hit_rate':28s} {h / max(h + m, 1):16.4f}   (TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum))")
    _emit(lines, out)


def _emit(lines, out):
    text = "\n".join(lines) + "\n"
    if out:
        with open(out, "w") as fh:
            fh.write(text)
    print(text, end="")


def timed(torch, stream, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    args = parse()
    EXTRA_OPTS.extend(args.opt)
    if args.summarize_kernels:
        summarize_kernels(args.summarize_kernels, " && ".join(args.command), args.out)
        return
    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.kernel, args.command, args.workload_note, args.out)
        return
    if args.summarize_traffic:
        summarize_traffic(*args.summarize_traffic, args.calib_bytes, args.command, args.out)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    runners = {"torus": run_torus, "batch": run_batch, "star2d": run_star2d, "diff": run_diff}
    res = runners[args.workload](args, torch)
    print(json.dumps(res), flush=True)


def run_torus(args, torch):
    """The 1,048,576-triangle torus with 16M closest-point queries (plus the
    fwd+bwd and ray rates and the CPU baseline on the same queries)."""
    import fcpw_amd as fcpw
    from fcpw_amd import synth

    dev = torch.device("cuda:0")
    U, Vg = args.grid
    t0 = time.time()
    v, f = synth.torus_mesh(U, Vg)
    V, F = synth.f32(v), f.astype(np.int32)
    lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
    rq = synth.Rng(synth.seed(2, 2))
    n_u = args.queries // 2
    q = synth.f32(np.concatenate([synth.uniform_in_box(n_u, lo, hi, 1.2, rq),
                                  synth.near_surface(args.queries - n_u, v, f, 0.05, rq)]))
    r2 = np.full(len(q), np.inf, np.float32)
    nq = len(q)
    log(f"[setup] {len(F)} triangles, {nq} queries generated in {time.time() - t0:.1f}s")

    t0 = time.time()
    scene = fcpw.Scene3D()
    scene.set_object_count(1)
    scene.set_object_vertices(V, 0)
    scene.set_object_triangles(F, 0)
    scene.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    gpu = fcpw.GPUScene3D(fcpw.slang_directory_path)
    gpu.transfer_to_gpu(scene)
    apply_opts(gpu)
    for k in ("block", "stack", "sort", "coop", "hold_min", "xcds"):
        val = getattr(args, k)
        if val is not None:
            gpu.set_option(k, val)
    opts = {k: gpu.get_option(k) for k in ("coop", "stack", "hold_min", "sort", "morton_bits", "xcds", "block")}
    log(f"[setup] BVH built + uploaded in {time.time() - t0:.1f}s: {gpu.info('n_records')} records, "
        f"depth {gpu.info('max_depth')}, options {opts}")

    sph = torch.empty((nq, 4), dtype=torch.float32, device=dev)
    sph[:, :3] = torch.from_numpy(q).to(dev)
    sph[:, 3] = float("inf")
    d_out = torch.empty((nq, 10), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def fwd():
        gpu.find_closest_points_device(sph, d_out, stream=stream)

    # ---- forward: the timed region (device-resident inputs, one launch per step)
    for _ in range(args.warmup):
        fwd()
    torch.cuda.synchronize()
    gpu.set_option("timing", 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fwd()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = gpu.info("kernel_ns") / 1e9 / gpu.info("kernel_launches")
    gpu.set_option("timing", 0)
    ms_step = wall * 1e3 / args.steps
    qps = nq / (wall / args.steps)
    log(f"[fwd] {ms_step:.3f} ms/step wall, traversal kernel {kernel_s * 1e3:.3f} ms, {qps:.4e} queries/s")

    # ---- traversal counts for the byte model (counting variant, same order)
    gpu.set_option("stats", 1)
    fwd()
    I = gpu.info("stat_internal_nodes") / gpu.info("stat_queries")
    P = gpu.info("stat_prim_tests") / gpu.info("stat_queries")
    waves = max(gpu.info("stat_waves"), 1)
    log(f"[fwd] per wave: {gpu.info('stat_wave_iterations') / waves:.1f} loop iterations, "
        f"{gpu.info('stat_leaf_passes') / waves:.1f} leaf passes, {gpu.info('stat_leaf_groups') / waves:.1f} "
        f"64-item groups (one query per lane, no refill)")
    gpu.set_option("stats", 0)
    B = algorithmic_bytes(I, P)
    achieved = B * nq / kernel_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(),
                "bytes_per_query": round(B, 1), "internal_nodes_per_query": round(I, 3),
                "prim_tests_per_query": round(P, 3), "kernel_ms": round(kernel_s * 1e3, 4),
                "queries_per_launch": nq}
    log(f"[fwd] I={I:.2f} P={P:.2f} B={B:.0f} B/query -> {achieved:.0f} GB/s algorithmic")

    res = {"metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"torus {len(F)} tris, {nq} closest-point queries (half uniform in 1.2x AABB, "
                                  f"half within 0.05 of the surface), r2=inf, forward",
                      "triangles": len(F), "queries": nq, **opts}}

    legs = set(args.legs.split(",")) if args.legs != "all" else None

    def want(leg):
        return not args.no_extra and (legs is None or leg in legs)

    if not args.no_extra:
        if want("fwd_bwd"):
            # ---- forward + backward (compact forward, adjoint kernel for dL/dx and dL/dV)
            x = sph[:, :3].contiguous()
            d = torch.empty(nq, dtype=torch.float32, device=dev)
            p = torch.empty((nq, 3), dtype=torch.float32, device=dev)
            prim = torch.empty(nq, dtype=torch.int32, device=dev)
            slot = torch.empty(nq, dtype=torch.int32, device=dev)
            gd = torch.ones(nq, dtype=torch.float32, device=dev)
            gx = torch.empty_like(x)
            gV = torch.zeros((len(V), 3), dtype=torch.float32, device=dev)
            h = stream.cuda_stream

            def fwd_bwd():
                gpu._closest_points_compact_device(x.data_ptr(), 0, nq, d.data_ptr(), p.data_ptr(), prim.data_ptr(),
                                                   slot.data_ptr(), h)
                gV.zero_()
                gpu._distance_backward_device(x.data_ptr(), slot.data_ptr(), gd.data_ptr(), nq, gx.data_ptr(),
                                              gV.data_ptr(), h)
            t_fb = timed(torch, stream, fwd_bwd, args.steps, args.warmup)
            res["fwd_bwd"] = {"value": nq / t_fb, "unit": "queries/s", "ms_per_step": t_fb * 1e3,
                              "ratio_to_fwd": t_fb * 1e3 / ms_step,
                              "bwd_order": gpu.get_option("bwd_order"), "bwd_reduce": gpu.get_option("bwd_reduce")}
            log(f"[fwd+bwd] {t_fb * 1e3:.3f} ms/step, {nq / t_fb:.4e} queries/s ({t_fb * 1e3 / ms_step:.2f}x fwd)")
            # the adjoint's other query orders on the same step (same dL/dx, dL/dV up to summation order)
            alt = {}
            for order, reduce in ((0, 0), (1, 0), (1, 1)):
                gpu.set_option("bwd_order", order)
                gpu.set_option("bwd_reduce", reduce)
                alt[f"order{order}_reduce{reduce}_ms_per_step"] = timed(torch, stream, fwd_bwd, args.steps,
                                                                         args.warmup) * 1e3
            gpu.set_option("bwd_order", 1)
            gpu.set_option("bwd_reduce", res["fwd_bwd"]["bwd_reduce"])
            res["fwd_bwd"]["other_adjoint_orders"] = alt
            log(f"[fwd+bwd] other adjoint orders: {alt}")
            # same closest points as the record-layout forward
            rec = d_out.view(-1)[: 4096 * 10].view(4096, 10)
            assert torch.equal(rec[:, 8], d[:4096]), "compact and record forward passes differ"
            del x, d, p, prim, slot, gd, gx, gV

        if want("e2e"):
            # ---- end to end through the drop-in API: host query arrays in, host records
            # out (GPUScene::findClosestPoints, fcpw_gpu.inl:219-274: pack, upload,
            # kernel, readback), against the device-resident forward above
            hits = fcpw.GPUInteractionList()
            gpu.find_closest_points(q, r2, hits)  # warm-up: pinned buffers and the record list
            e2e = []
            for _ in range(3):
                t0 = time.perf_counter()
                gpu.find_closest_points(q, r2, hits)
                e2e.append(time.perf_counter() - t0)
            t_e2e = float(np.median(e2e))
            dev_rec = d_out[:, 8:10].cpu().numpy()
            assert np.array_equal(hits.get_distances().view(np.uint32), dev_rec[:, 0].view(np.uint32)) and \
                np.array_equal(hits.get_indices(), dev_rec[:, 1].view(np.uint32)), "API and device differ"
            del dev_rec
            res["end_to_end"] = {"value": nq / t_e2e, "unit": "queries/s", "ms_per_call": t_e2e * 1e3,
                                 "ratio_to_device_resident": t_e2e * 1e3 / ms_step,
                                 "what": "GPUScene3D.find_closest_points(points, r2, hits) on the same 16M host "
                                         "queries: GPUBoundingSphere packing, host-to-device copy, ordering and "
                                         "kernel, copy of the GPUInteraction records back into the list; median "
                                         "of 3 after a warm-up call"}
            log(f"[e2e] API call {t_e2e * 1e3:.1f} ms ({nq / t_e2e:.4e} queries/s, {t_e2e * 1e3 / ms_step:.2f}x the "
                f"device-resident step)")
            del hits

        if want("rays"):
            # ---- rays
            rr = synth.Rng(synth.seed(2, 3))
            o = synth.f32(synth.uniform_in_box(nq, lo, hi, 1.0, rr))
            dr = synth.f32(rr.sphere(nq, 3))
            rays = torch.empty((nq, 10), dtype=torch.float32, device=dev)
            rays[:, 0:3] = torch.from_numpy(o).to(dev)
            rays[:, 3:6] = torch.from_numpy(dr).to(dev)
            rays[:, 6:9] = 1.0 / rays[:, 3:6]
            rays[:, 9] = float("inf")
            ray_out = torch.empty((nq, 10), dtype=torch.float32, device=dev)
            del sph

            def ray_step():
                gpu.intersect_device(rays, ray_out, stream=stream)
            t_r = timed(torch, stream, ray_step, args.steps, args.warmup)
            res["rays"] = {"value": nq / t_r, "unit": "rays/s", "ms_per_step": t_r * 1e3,
                           "workload": f"{nq} rays, origins uniform in the AABB, directions uniform on S^2, tMax=inf",
                           "ray_coop": gpu.get_option("ray_coop")}
            log(f"[rays] {t_r * 1e3:.3f} ms/step, {nq / t_r:.4e} rays/s (ray_coop={gpu.get_option('ray_coop')})")
            # the other ray kernel on the same rays: same bits, its own time
            first = ray_out.clone()
            other = 1 - gpu.get_option("ray_coop")
            gpu.set_option("ray_coop", other)
            t_o = timed(torch, stream, ray_step, args.steps, args.warmup)
            assert torch.equal(first.view(torch.int32), ray_out.view(torch.int32)), "ray kernels differ"
            gpu.set_option("ray_coop", 1 - other)
            res["rays_other_kernel"] = {"value": nq / t_o, "unit": "rays/s", "ms_per_step": t_o * 1e3,
                                        "ray_coop": other}
            log(f"[rays] ray_coop={other}: {t_o * 1e3:.3f} ms/step, {nq / t_o:.4e} rays/s")
            del rays, first, ray_out

        if want("spheres"):
            # ---- stochastic sphere intersection at the query points, radius 0.1
            balls = _spheres(torch, q, dev)
            balls[:, 3] = 0.01
            rnd = torch.rand((nq, 3), dtype=torch.float32, device=dev)
            ray_out = torch.empty((nq, 10), dtype=torch.float32, device=dev)

            def sphere_step():
                gpu.intersect_spheres_device(balls, rnd, ray_out, stream=stream)
            t_s = timed(torch, stream, sphere_step, args.steps, args.warmup)
            found = float((ray_out[:, 9].view(torch.int32) != -1).float().mean())
            res["sphere_intersection"] = {"value": nq / t_s, "unit": "queries/s", "ms_per_step": t_s * 1e3,
                                          "workload": f"{nq} spheres of radius 0.1 at the closest-point queries, "
                                                      f"uniform random numbers; {found:.3f} find a primitive"}
            log(f"[spheres] {t_s * 1e3:.3f} ms/step, {nq / t_s:.4e} queries/s, found {found:.3f}")
            del balls, rnd, ray_out

        if want("silhouettes"):
            # ---- closest silhouette points on the same torus built with silhouettes,
            # every fourth closest-point query (uniform and near-surface alike)
            t0 = time.time()
            ss = fcpw.Scene3D()
            ss.set_object_count(1)
            ss.set_object_vertices(V, 0)
            ss.set_object_triangles(F, 0)
            ss.compute_silhouettes()
            ss.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
            gs = fcpw.GPUScene3D(fcpw.slang_directory_path)
            gs.transfer_to_gpu(ss)
            apply_opts(gs)
            log(f"[silhouettes] silhouettes + cones built and transferred in {time.time() - t0:.1f}s")
            qs = q[::4]
            ns = len(qs)
            balls = _spheres(torch, qs, dev)
            flip = torch.zeros(ns, dtype=torch.int32, device=dev)
            sil_out = torch.empty((ns, 10), dtype=torch.float32, device=dev)

            def sil_step():
                gs.find_closest_silhouette_points_device(balls, flip, sil_out, stream=stream)
            t_sil = timed(torch, stream, sil_step, args.steps, args.warmup)
            found = float((sil_out[:, 9].view(torch.int32) != -1).float().mean())
            res["closest_silhouette"] = {"value": ns / t_sil, "unit": "queries/s", "ms_per_step": t_sil * 1e3,
                                         "workload": f"{ns} queries (every 4th closest-point query), r2=inf, "
                                                     f"flip=0; {found:.3f} find a silhouette"}
            log(f"[silhouettes] {t_sil * 1e3:.3f} ms/step, {ns / t_sil:.4e} queries/s, found {found:.3f}")
            # the one-query-per-lane kernel on the same queries: same records, its own time
            first = sil_out.clone()
            gs.set_option("coop", 0)
            t_sil1 = timed(torch, stream, sil_step, args.steps, args.warmup)
            assert torch.equal(first.view(torch.int32), sil_out.view(torch.int32)), "silhouette kernels differ"
            gs.set_option("coop", 1)
            res["closest_silhouette"]["per_lane_ms_per_step"] = t_sil1 * 1e3
            log(f"[silhouettes] one query per lane: {t_sil1 * 1e3:.3f} ms/step")
            del balls, flip, sil_out, gs, ss, first

        if want("build"):
            # ---- BVH construction: the reference's SAH tree built on the device
            # (geometry upload included) against the host build + transfer
            sb = fcpw.Scene3D()
            sb.set_object_count(1)
            sb.set_object_vertices(V, 0)
            sb.set_object_triangles(F, 0)
            gb = fcpw.GPUScene3D(fcpw.slang_directory_path)
            gb.build_on_gpu(sb)
            torch.cuda.synchronize()
            t_dev = []
            for _ in range(5):
                t0 = time.perf_counter()
                gb.build_on_gpu(sb)
                torch.cuda.synchronize()
                t_dev.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            sb.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
            t_hb = time.perf_counter() - t0
            t0 = time.perf_counter()
            gb.transfer_to_gpu(sb)
            torch.cuda.synchronize()
            t_tr = time.perf_counter() - t0
            t_db = float(np.median(t_dev))
            res["bvh_build"] = {"device_ms": t_db * 1e3, "host_build_ms": t_hb * 1e3, "host_transfer_ms": t_tr * 1e3,
                                "speedup": (t_hb + t_tr) / t_db,
                                "workload": f"{len(F)} triangles, SAH, the same tree either way; device time includes "
                                            f"the geometry upload (median of 5)"}
            log(f"[build] device {t_db * 1e3:.2f} ms vs host build {t_hb * 1e3:.1f} ms + transfer {t_tr * 1e3:.1f} ms")
            del gb, sb

        if want("instances"):
            res["instances"] = run_instances(args, torch, fcpw, synth, dev, stream)

    if want("configs"):
        # the other configurations, each as a sub-object of the headline line
        keep = ("value", "unit", "ms_per_step", "config", "rays", "refit", "forward_only", "ratio_step_to_forward",
                "loss_every_10_steps", "eager", "separate_launches", "speedup_vs_separate", "fwd_bwd",
                "closest_after_refit_ms", "metric")
        for name, fn in (("star2d", run_star2d), ("diff", run_diff), ("batch", run_batch)):
            t0 = time.time()
            sub = fn(args, torch)
            res[name] = {k: v for k, v in sub.items() if k in keep}
            log(f"[{name}] done in {time.time() - t0:.1f}s")
            torch.cuda.empty_cache()

    if args.calib_copy_mib > 0:
        nbytes = args.calib_copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"[calib] copied {nbytes} bytes 3x")

    res["roofline"] = roofline
    res["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(V, F, q, r2, args.cpu_sample)
        log(f"[cpu] {res['cpu_baseline']['value']:.3e} queries/s on {res['cpu_baseline']['cores']} threads")
    return res


def run_instances(args, torch, fcpw, synth, dev, stream, n_inst=256, n_queries=1 << 22):
    """Scene of transformed instances: 256 copies of the 20,480-triangle noisy
    icosphere (random rotation, uniform scale in [0.3, 1], centres uniform in
    [-8, 8]^3), 4M closest-point queries uniform in the scene box, r2 = inf.
    Timed against the same triangles baked into one 5.2M-triangle mesh."""
    v, f = synth.icosphere(5)
    V = synth.f32(synth.radial_noise(v, 0.01, synth.Rng(synth.seed(6, 1))))
    F = f.astype(np.int32)
    rng = synth.Rng(synth.seed(6, 2))
    mats = []
    for _ in range(n_inst):
        ax = rng.sphere(1, 3)[0]
        a = rng.uniform(1, 0.0, 2 * np.pi)[0]
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
        M = np.eye(4)
        M[:3, :3] = rng.uniform(1, 0.3, 1.0)[0] * R
        M[:3, 3] = rng.uniform(3, -8.0, 8.0)
        mats.append(M)
    mats = np.asarray(mats, np.float32)
    t0 = time.time()
    si = fcpw.Scene3D()
    si.set_object_count(1)
    si.set_object_vertices(V, 0)
    si.set_object_triangles(F, 0)
    si.set_object_instance_transforms(mats, 0)
    si.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    gi = fcpw.GPUScene3D(fcpw.slang_directory_path)
    gi.transfer_to_gpu(si)
    apply_opts(gi)
    t_inst_build = time.time() - t0
    # the same instances baked into one mesh (FP64 transform, rounded once)
    W = np.concatenate([V.astype(np.float64) @ M[:3, :3].T.astype(np.float64) + M[:3, 3] for M in mats])
    WF = np.concatenate([F + k * len(V) for k in range(n_inst)]).astype(np.int32)
    t0 = time.time()
    _, gf = _scene(fcpw, synth.f32(W), WF)
    t_flat_build = time.time() - t0
    lo, hi = W.min(0), W.max(0)
    q = synth.f32(synth.uniform_in_box(n_queries, lo, hi, 1.0, synth.Rng(synth.seed(6, 3))))
    sph = _spheres(torch, q, dev)
    out_i = torch.empty((n_queries, 10), dtype=torch.float32, device=dev)
    out_f = torch.empty_like(out_i)
    obj = torch.empty(n_queries, dtype=torch.int32, device=dev)

    def inst_step():
        gi.find_closest_points_device(sph, out_i, stream=stream, object_index=obj)

    def flat_step():
        gf.find_closest_points_device(sph, out_f, stream=stream)
    t_i = timed(torch, stream, inst_step, args.steps, args.warmup)
    # the instances through the one-query-per-lane kernel: same records, its own time
    first = out_i.clone()
    gi.set_option("coop", 0)
    t_i1 = timed(torch, stream, inst_step, args.steps, args.warmup)
    gi.set_option("coop", 1)
    assert torch.equal(first.view(torch.int32), out_i.view(torch.int32)), "instance kernels differ"
    # one launch with the caller's radii (no lowering from each group's first query): same records
    seed_group = gi.get_option("inst_seed")
    gi.set_option("inst_seed", 0)
    t_i0 = timed(torch, stream, inst_step, args.steps, args.warmup)
    gi.set_option("inst_seed", seed_group)
    assert torch.equal(first.view(torch.int32), out_i.view(torch.int32)), "lowered radii changed a record"
    del first
    t_f = timed(torch, stream, flat_step, args.steps, args.warmup)
    # the baked mesh through the one-query-per-lane kernel, the instanced kernels' structure
    gf.set_option("coop", 0)
    t_f1 = timed(torch, stream, flat_step, args.steps, args.warmup)
    gf.set_option("coop", 1)
    # the two scenes round the mapped geometry differently (FP32 per query vs FP64 once)
    dmax = float((out_i[:, 8].double() - out_f[:, 8].double()).abs().max())
    # rays through the instances: origins uniform in [-9, 9]^3, directions uniform on S^2, tMax = inf
    rr = synth.Rng(synth.seed(6, 4))
    ro = synth.f32(synth.uniform_in_box(n_queries, np.array([-9.0] * 3), np.array([9.0] * 3), 1.0, rr))
    rays = _rays(torch, ro, synth.f32(rr.sphere(n_queries, 3)), dev)
    t_r = timed(torch, stream, lambda: gi.intersect_device(rays, out_i, stream=stream), args.steps, args.warmup)
    first = out_i.clone()
    gi.set_option("ray_coop", 0)
    t_r1 = timed(torch, stream, lambda: gi.intersect_device(rays, out_i, stream=stream), args.steps, args.warmup)
    gi.set_option("ray_coop", 1)
    assert torch.equal(first.view(torch.int32), out_i.view(torch.int32)), "instance ray kernels differ"
    ray_hits = float((first[:, 9].view(torch.int32) != -1).float().mean())
    del first, rays
    log(f"[instances] {n_queries} rays: {t_r * 1e3:.3f} ms ({n_queries / t_r:.4e} rays/s; one ray per lane "
        f"{t_r1 * 1e3:.3f} ms), {ray_hits:.3f} hit")
    log(f"[instances] one query per lane: {t_i1 * 1e3:.3f} ms; one launch with the caller's radii "
        f"(inst_seed 0): {t_i0 * 1e3:.3f} ms")
    log(f"[instances] {n_inst} instances x {len(F)} tris: two-level {t_i * 1e3:.3f} ms "
        f"({n_queries / t_i:.4e} q/s) vs baked mesh {t_f * 1e3:.3f} ms ({n_queries / t_f:.4e} q/s; "
        f"one query per lane {t_f1 * 1e3:.3f} ms); "
        f"max |d difference| {dmax:.2e}; builds {t_inst_build:.2f}s / {t_flat_build:.2f}s")
    return {"value": n_queries / t_i, "unit": "queries/s", "ms_per_step": t_i * 1e3,
            "baked_mesh_value": n_queries / t_f, "baked_mesh_ms_per_step": t_f * 1e3,
            "baked_mesh_per_lane_ms_per_step": t_f1 * 1e3, "per_lane_ms_per_step": t_i1 * 1e3,
            "inst_seed": seed_group, "one_launch_ms_per_step": t_i0 * 1e3,
            "rays": {"value": n_queries / t_r, "unit": "rays/s", "ms_per_step": t_r * 1e3,
                     "per_lane_ms_per_step": t_r1 * 1e3, "hit_fraction": ray_hits,
                     "workload": f"{n_queries} rays, origins uniform in [-9, 9]^3, directions uniform, tMax=inf"},
            "max_abs_d_difference": dmax, "host_build_s": t_inst_build, "baked_host_build_s": t_flat_build,
            "workload": f"{n_inst} transformed instances of a {len(F)}-triangle icosphere, {n_queries} "
                        f"closest-point queries uniform in the scene box, r2=inf; baked = the same "
                        f"{n_inst * len(F)} triangles as one mesh"}


def _scene(fcpw, V, F):
    dim = V.shape[1]
    s = fcpw.Scene3D() if dim == 3 else fcpw.Scene2D()
    s.set_object_count(1)
    s.set_object_vertices(V, 0)
    (s.set_object_triangles if dim == 3 else s.set_object_line_segments)(F, 0)
    s.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    g = fcpw.GPUScene3D("") if dim == 3 else fcpw.GPUScene2D("")
    g.transfer_to_gpu(s)
    apply_opts(g)
    return s, g


def _spheres(torch, q, dev):
    sph = torch.empty((len(q), 4), dtype=torch.float32, device=dev)
    sph[:, :3] = 0.0
    sph[:, :q.shape[1]] = torch.from_numpy(q).to(dev)
    sph[:, 3] = float("inf")
    return sph


def _rays(torch, o, d, dev):
    dim = o.shape[1]
    rays = torch.zeros((len(o), 10), dtype=torch.float32, device=dev)
    rays[:, 0:dim] = torch.from_numpy(o).to(dev)
    rays[:, 3:3 + dim] = torch.from_numpy(d).to(dev)
    rays[:, 6:9] = 1.0 / rays[:, 3:6]
    rays[:, 9] = float("inf")
    return rays


def run_batch(args, torch):
    """256 independent meshes (alternately tori and geodesic spheres, triangle
    counts log-uniform in 1k-500k) with log-uniform 1k-1M queries each (half
    uniform in 1.2x the mesh box, half within 0.05 of the surface): one
    persistent launch over all of them, replayed from a HIP graph, against one
    plain launch per mesh; forward, and compact forward + adjoint."""
    import fcpw_amd as fcpw
    from fcpw_amd import autotune, synth
    from fcpw_amd.batch import BatchedScenes

    dev = torch.device("cuda:0")
    t0 = time.time()
    meshes = synth.batch_workload()
    gpus = []
    for m in meshes:
        gpus.append(_scene(fcpw, m["positions"], m["indices"])[1])
    counts = [len(m["queries"]) for m in meshes]
    q_all = np.concatenate([m["queries"] for m in meshes])
    n = len(q_all)
    n_tris = sum(len(m["indices"]) for m in meshes)
    log(f"[setup] {len(meshes)} meshes, {n_tris} triangles, {n} queries in {time.time() - t0:.1f}s")
    sph = _spheres(torch, q_all, dev)
    out = torch.empty((n, 10), dtype=torch.float32, device=dev)
    out_sep = torch.empty_like(out)
    batch = BatchedScenes(gpus)
    batch.set_queries(counts)
    t0 = time.time()
    best, _ = autotune.tune_batch(batch, sph, out, reps=1)
    log(f"[tune] {best} in {time.time() - t0:.1f}s")

    def fwd():
        batch.find_closest_points_device(sph, out)

    def sep():
        batch.separate_find_closest_points_device(sph, out_sep)
    t_eager = timed(torch, None, fwd, args.steps, args.warmup)
    graph = batch.capture(fwd)
    t_graph = timed(torch, None, graph.replay, args.steps, args.warmup)
    t_sep = timed(torch, None, sep, args.steps, args.warmup)
    same = bool(torch.equal(out.view(torch.int32), out_sep.view(torch.int32)))
    log(f"[fwd] graph {t_graph * 1e3:.2f} ms, eager {t_eager * 1e3:.2f} ms, separate {t_sep * 1e3:.2f} ms, "
        f"identical {same}")
    del out_sep

    x = sph[:, :3].contiguous()
    d = torch.empty(n, dtype=torch.float32, device=dev)
    slot = torch.empty(n, dtype=torch.int32, device=dev)
    gd = torch.ones(n, dtype=torch.float32, device=dev)
    gx = torch.empty_like(x)
    gV = torch.zeros((int(batch.vertex_offsets[-1]), 3), dtype=torch.float32, device=dev)

    def fwd_bwd():
        batch.closest_points_compact_device(x, d, slot=slot)
        gV.zero_()
        batch.distance_backward_device(x, slot, gd, gx, gV)
    graph_fb = batch.capture(fwd_bwd)
    t_fb = timed(torch, None, graph_fb.replay, args.steps, args.warmup)
    st = torch.cuda.current_stream().cuda_stream

    def fwd_bwd_sep():
        gV.zero_()
        for s_, g in enumerate(gpus):
            a, b = int(batch.offsets[s_]), int(batch.offsets[s_ + 1])
            v0 = int(batch.vertex_offsets[s_])
            if b > a:
                g._closest_points_compact_device(x[a:].data_ptr(), 0, b - a, d[a:].data_ptr(), 0, 0,
                                                 slot[a:].data_ptr(), st)
                g._distance_backward_device(x[a:].data_ptr(), slot[a:].data_ptr(), gd[a:].data_ptr(), b - a,
                                            gx[a:].data_ptr(), gV[v0:].data_ptr(), st)
    t_fb_sep = timed(torch, None, fwd_bwd_sep, args.steps, args.warmup)
    log(f"[fwd+bwd] graph {t_fb * 1e3:.2f} ms, separate {t_fb_sep * 1e3:.2f} ms")
    return {"metric": "closest-point queries/s, batched uneven workload (one persistent launch)", "value": n / t_graph,
            "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": t_graph * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{len(meshes)} meshes (tori / geodesic spheres, 1k-500k triangles log-uniform), "
                                   f"1k-1M queries each, {n_tris} triangles and {n} queries in total",
                       "meshes": len(meshes), "triangles": n_tris, "queries": n, **batch.options()},
            "eager": {"value": n / t_eager, "ms_per_step": t_eager * 1e3},
            "separate_launches": {"value": n / t_sep, "ms_per_step": t_sep * 1e3, "launches_per_step": len(meshes),
                                  "identical_results": same},
            "speedup_vs_separate": t_sep / t_graph,
            "fwd_bwd": {"value": n / t_fb, "ms_per_step": t_fb * 1e3, "ratio_to_fwd": t_fb / t_graph,
                        "separate_ms_per_step": t_fb_sep * 1e3}}


def run_star2d(args, torch):
    """262,144-segment star polygon, 8M queries uniform in [-1.5, 1.5]^2:
    closest points and rays (directions uniform on S^1); then 10 steps that
    each move a random 10% of the vertices radially by U(-0.01, 0.01), upload
    them and refit the BVH on the device, with the queries rerun after each."""
    import fcpw_amd as fcpw
    from fcpw_amd import synth

    dev = torch.device("cuda:0")
    w = synth.star_refit_workload()
    V, F = w["positions"], w["indices"]
    scene, gpu = _scene(fcpw, V, F)
    n = len(w["queries"])
    sph = _spheres(torch, w["queries"], dev)
    out = torch.empty((n, 10), dtype=torch.float32, device=dev)
    rays = _rays(torch, w["ray_o"], w["ray_d"], dev)
    rout = torch.empty((n, 10), dtype=torch.float32, device=dev)
    t_cp = timed(torch, None, lambda: gpu.find_closest_points_device(sph, out), args.steps, args.warmup)
    t_ray = timed(torch, None, lambda: gpu.intersect_device(rays, rout), args.steps, args.warmup)
    st = torch.cuda.current_stream().cuda_stream
    refit_ms, cp_ms = [], []
    Vd = torch.empty((len(V), 2), dtype=torch.float32, device=dev)
    Vk = V
    for step in range(10):
        Vk = synth.perturb_star(Vk, step)
        Vd.copy_(torch.from_numpy(Vk))
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        gpu._update_positions_device(Vd.data_ptr(), len(V), st)
        gpu._refit_device(st)
        b.record()
        b.synchronize()
        refit_ms.append(a.elapsed_time(b))
        cp_ms.append(timed(torch, None, lambda: gpu.find_closest_points_device(sph, out), 3, 0) * 1e3)
    log(f"[star2d] closest {t_cp * 1e3:.2f} ms, rays {t_ray * 1e3:.2f} ms, refit {np.median(refit_ms):.3f} ms")
    return {"metric": "closest-point queries/s, 2D segment scene with refit", "value": n / t_cp, "unit": "queries/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_cp * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"star polygon {len(F)} segments, {n} queries uniform in [-1.5,1.5]^2, r2=inf",
                       "segments": len(F), "queries": n, **{k: gpu.get_option(k) for k in ("coop", "stack", "hold_min")}},
            "rays": {"value": n / t_ray, "unit": "rays/s", "ms_per_step": t_ray * 1e3},
            "refit": {"ms_median": float(np.median(refit_ms)), "ms_max": float(np.max(refit_ms)),
                      "what": "vertex upload (device gather) + one bottom-up refit launch, 10 perturbation steps"},
            "closest_after_refit_ms": [round(v, 3) for v in cp_ms]}


def run_diff(args, torch):
    """Differentiable distance: class-I geodesic sphere of frequency 100
    (200,000 triangles, radial noise 0.01), 4M query points within 0.1 of the
    surface, loss = mean d^2 minimised over the vertex positions with Adam for
    100 steps; each step uploads the vertices, refits the BVH on the device,
    runs the compact forward and the adjoint (dL/dx and dL/dV)."""
    from fcpw_amd import synth
    from fcpw_amd.torch_ops import closest_point_distance
    import fcpw_amd as fcpw

    dev = torch.device("cuda:0")
    w = synth.geodesic_workload()
    V, F = w["positions"], w["indices"]
    scene, gpu = _scene(fcpw, V, F)
    n = len(w["queries"])
    x = torch.from_numpy(w["queries"]).to(dev).requires_grad_(True)
    Vt = torch.nn.Parameter(torch.from_numpy(V).to(dev))
    opt = torch.optim.Adam([Vt], lr=1e-4)
    losses = []

    def step():
        opt.zero_grad(set_to_none=False)
        d, p, prim = closest_point_distance(gpu, x, Vt)
        loss = (d * d).mean()
        loss.backward()
        opt.step()
        return loss
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(100):
        loss = step()
        if k % 10 == 0:
            losses.append(float(loss.detach()))
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / 100
    # the forward alone on the same queries, for the fwd+bwd ratio
    with torch.no_grad():
        t_fwd = timed(torch, None, lambda: closest_point_distance(gpu, x.detach()), args.steps, args.warmup)
    log(f"[diff] step {t_step * 1e3:.2f} ms (forward alone {t_fwd * 1e3:.2f} ms), loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    return {"metric": "differentiable closest-point queries/s (forward + adjoint + refit + Adam step)",
            "value": n / t_step, "unit": "queries/s", "n_gpus": 1, "steps": 100, "warmup": args.warmup,
            "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"geodesic sphere {len(F)} tris, {n} queries within 0.1 of the surface, "
                                   "loss mean d^2, Adam lr 1e-4, 100 steps with device refit",
                       "triangles": len(F), "queries": n},
            "forward_only": {"value": n / t_fwd, "ms_per_step": t_fwd * 1e3},
            "ratio_step_to_forward": t_step / t_fwd, "loss_every_10_steps": losses}


if __name__ == "__main__":
    main()
