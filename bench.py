"""Closest-point throughput on one MI355X.

Workload (one "step" = one pass of the closest-point kernel over the batch):
a seeded torus R=1, r=0.3 tessellated to 1,048,576 float32 triangles (1024 x 512
grid, per-vertex jitter 1e-3) and 16,777,216 closest-point queries (half uniform
in the 1.2x AABB, half within 0.05 of the surface), squared max radius inf.
Queries are resident in HBM as GPUBoundingSphere records before the timed
region; results are written as GPUInteraction records (the drop-in layout).

The printed JSON line carries
  value        queries/s of the forward closest-point query (whole job, 1 GPU)
  roofline     algorithmic bytes per query (from the reference GPU traversal's
               node and primitive counts, counted by the oracle on a 65,536-query
               sample) x queries/s against MI355X HBM bandwidth; kernel time from
               HIP events on the launch stream
  cpu_baseline the reference CPU path (oracle restatement of Scene<3>::
               findClosestPoints, std::thread-style static partition) timed on
               this host's cores over a bounded sample of the same queries
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "closest-point queries/s, 1M-tri mesh, fwd and fwd+bwd; % HBM roofline vs CPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--queries", type=int, default=1 << 24)
    p.add_argument("--grid", type=int, nargs=2, default=[1024, 512], help="torus U V (2UV triangles)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=1 << 19, help="queries timed on the CPU path")
    p.add_argument("--block", type=int, default=None)
    p.add_argument("--stack", type=int, default=None)
    p.add_argument("--sort", type=int, default=None)
    return p.parse_args()


def algorithmic_bytes(n_internal, n_prims, sphere_in=16, record_out=40, node_bytes=32, prim_bytes=40):
    """B = in + out + 32 (1 + 2 I) + 40 P per query: the root box, two child
    boxes per expanded internal node and one primitive record per primitive
    test, at the reference GPU record sizes (bvh_interop_structures.h:55-119)."""
    return sphere_in + record_out + node_bytes * (1.0 + 2.0 * n_internal) + prim_bytes * n_prims


def cpu_baseline(V, F, q, r2, sample, counters_sample=65536):
    """Reference CPU path on this host (oracle = restatement of bvh.inl)."""
    import oracle
    t0 = time.time()
    ob = oracle.Bvh(V, F)
    log(f"[cpu] oracle BVH built in {time.time() - t0:.1f}s")
    # node/primitive counts of the reference GPU traversal order, for the byte model
    idx = np.random.default_rng(2024).choice(len(q), size=min(counters_sample, len(q)), replace=False)
    _, ni, npr, _ = ob.closest_points_gpuorder(q[idx], r2[idx], counters=True)
    I, P = float(ni.mean()), float(npr.mean())
    cores = os.cpu_count() or 1
    n = min(sample, len(q))
    qs, rs = q[:n], r2[:n]
    ob.closest_points_cpu(qs[: min(n, 4096)], rs[: min(n, 4096)], nthreads=cores)  # warm up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        ob.closest_points_cpu(qs, rs, nthreads=cores)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": n / t, "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of the same queries, median of 3 runs, {cores} threads, BVH build excluded"}, I, P


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import fcpw_amd as fcpw
    from fcpw_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
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
    log(f"[setup] {len(F)} triangles, {len(q)} queries generated in {time.time() - t0:.1f}s")

    t0 = time.time()
    scene = fcpw.Scene3D()
    scene.set_object_count(1)
    scene.set_object_vertices(V, 0)
    scene.set_object_triangles(F, 0)
    scene.build(fcpw.AggregateType.Bvh_SurfaceArea, False)
    gpu = fcpw.GPUScene3D(fcpw.slang_directory_path)
    gpu.transfer_to_gpu(scene)
    for k, val in (("block", args.block), ("stack", args.stack), ("sort", args.sort)):
        if val is not None:
            gpu.set_option(k, val)
    log(f"[setup] BVH built + uploaded in {time.time() - t0:.1f}s: {gpu.info('n_records')} records, "
        f"depth {gpu.info('max_depth')}, block {gpu.get_option('block')}, stack {gpu.get_option('stack')}")

    sph = np.empty((len(q), 4), np.float32)
    sph[:, :3] = q
    sph[:, 3] = r2
    d_in = torch.from_numpy(sph).to(dev)
    d_out = torch.empty((len(q), 10), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        gpu.find_closest_points_device(d_in, d_out, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_step = wall * 1e3 / args.steps
    qps = len(q) / (wall / args.steps)
    log(f"[gpu] {ms_step:.3f} ms/step wall, kernel median {np.median(kernel_ms):.3f} ms, {qps:.3e} queries/s")

    # spot check against the oracle's GPU-order restatement on a sample
    out = d_out[:4096].cpu().numpy().view(fcpw.INTERACTION_DTYPE).ravel()
    import oracle  # checker only
    ref = oracle.Bvh(V, F).closest_points_gpuorder(q[:4096], r2[:4096]) if args.no_cpu_baseline else None

    cpu, I, P = (None, None, None)
    if not args.no_cpu_baseline:
        cpu, I, P = cpu_baseline(V, F, q, r2, args.cpu_sample)
        log(f"[cpu] {cpu['value']:.3e} queries/s on {cpu['cores']} threads; GPU-order counts I={I:.1f} P={P:.1f}")
    if ref is not None:
        assert np.array_equal(out["d"], ref["d"]), "GPU result differs from the oracle"

    roofline = None
    if I is not None:
        B = algorithmic_bytes(I, P)
        kern_s = float(np.median(kernel_ms)) / 1e3
        achieved = B * len(q) / kern_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "bytes_per_query": round(B, 1), "internal_nodes_per_query": round(I, 2),
                    "prim_tests_per_query": round(P, 2), "kernel_ms": round(float(np.median(kernel_ms)), 4)}

    res = {"metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"torus {len(F)} tris, {len(q)} closest-point queries (half uniform in 1.2x AABB, "
                                  f"half within 0.05 of the surface), r2=inf, fwd",
                      "triangles": len(F), "queries": len(q), "block": gpu.get_option("block"),
                      "stack": gpu.get_option("stack")},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
