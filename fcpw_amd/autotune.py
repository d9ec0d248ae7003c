"""Launch-configuration autotuner for the closest-point kernels.

None of the knobs changes a result bit (every configuration visits the BVH in
the same order), so tuning is a pure timing search:

``tune_scene(gpu_scene, spheres)``
    times the single-scene closest-point launch over candidate settings of
    ``layout`` (the order of the tree's records in memory: DFS or treelets),
    ``coop`` (wave-cooperative leaves or not), ``stack`` (traversal-stack
    entries kept in LDS), ``hold_min`` (held leaves before a cooperative leaf
    pass), ``xcds`` (how many ranges of the ordered queries the XCDs split
    into) and, without coop, ``block``; applies and returns the fastest.
    Candidates are grouped by layout, so the records are reordered once per
    layout.

``tune_batch(batch, spheres, out)``
    the same for a ``BatchedScenes`` persistent launch, over ``stack``,
    ``hold_min``, ``chunk`` (queries per work item) and ``waves_per_cu``.

Winners are cached per (dimension, log2 primitives, log2 queries) bucket, so
scenes of similar size reuse a result without timing again.  Timing uses HIP
events on the current stream around each launch; the median of ``reps`` runs
after one warm-up run decides.
"""
import math

import torch

__all__ = ["SCENE_CANDIDATES", "BATCH_CANDIDATES", "tune_scene", "tune_batch", "clear_cache"]

SCENE_CANDIDATES = (
    [dict(layout=L, coop=1, stack=k, hold_min=h, xcds=x) for L in (0, 1) for k in (4, 6, 8, 12) for h in (1, 2, 4)
     for x in (4, 8)]
    + [dict(layout=0, coop=0, stack=k, block=b, xcds=x) for k in (8, 16) for b in (64, 128) for x in (4, 8)]
)
BATCH_CANDIDATES = [dict(stack=k, hold_min=h, chunk=c, waves_per_cu=0)
                    for k in (6, 8, 12) for h in (1, 2) for c in (64, 256)]

_cache = {}


def clear_cache():
    _cache.clear()


def _bucket(n):
    return int(math.log2(max(int(n), 1)))


def _time(fn, reps):
    fn()  # warm-up (also sizes any workspace)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune_scene(gpu_scene, spheres, out=None, candidates=None, reps=3, sample=1 << 20, use_cache=True):
    """Pick the fastest launch settings for gpu_scene on (a sample of) spheres,
    an (n, 4) float32 CUDA tensor of GPUBoundingSphere records.  Returns
    (best settings, {settings: ms})."""
    n = min(int(spheres.shape[0]), int(sample))
    key = ("scene", gpu_scene.DIM, _bucket(gpu_scene.info("n_prims")), _bucket(n))
    if use_cache and key in _cache:
        best = _cache[key]
        for k, v in best.items():
            gpu_scene.set_option(k, v)
        return dict(best), {}
    sph = spheres[:n]
    dst = out[:n] if out is not None else torch.empty((n, 10), dtype=torch.float32, device=spheres.device)
    results = {}
    for cand in candidates or SCENE_CANDIDATES:
        for k, v in cand.items():
            gpu_scene.set_option(k, v)
        results[tuple(sorted(cand.items()))] = _time(lambda: gpu_scene.find_closest_points_device(sph, dst), reps)
    best = dict(min(results, key=results.get))
    for k, v in best.items():
        gpu_scene.set_option(k, v)
    _cache[key] = dict(best)
    return best, results


def tune_batch(batch, spheres, out, candidates=None, reps=3, use_cache=True):
    """Pick the fastest persistent-launch settings for a BatchedScenes whose
    queries are set; spheres / out span the whole batch."""
    n_prims = sum(g.info("n_prims") for g in batch.scenes)
    key = ("batch", batch.DIM, len(batch.scenes), _bucket(n_prims), _bucket(batch.n_queries))
    if use_cache and key in _cache:
        best = _cache[key]
        for k, v in best.items():
            batch.set_option(k, v)
        return dict(best), {}
    results = {}
    for cand in candidates or BATCH_CANDIDATES:
        for k, v in cand.items():
            batch.set_option(k, v)
        results[tuple(sorted(cand.items()))] = _time(lambda: batch.find_closest_points_device(spheres, out), reps)
    best = dict(min(results, key=results.get))
    for k, v in best.items():
        batch.set_option(k, v)
    _cache[key] = dict(best)
    return best, results
