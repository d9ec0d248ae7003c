"""Differentiable unsigned distance d(x, mesh) on the MI355X backend.

The reference has no gradients; this is the adjoint layered on its closest-point
query.  ``closest_point_distance(scene, x, V)`` returns d (n,), the closest
points p (n, dim) and the original primitive indices (n,).  Gradients flow to
the query points x and, when given, to the vertex positions V (caller's vertex
order):

    dd/dx   =  (x - p) / d
    dd/dV_k = -b_k (x - p) / d      (b = barycentric weights of p on its primitive)

(p moves with V, but d is a minimum over the primitive, so by the envelope
theorem only the explicit dependence through p(V) at fixed barycentrics remains).
The backward pass recomputes p and b from x and the saved leaf slot; no
traversal state is stored.  dL/dV is accumulated with FP32 atomics in HBM.
Queries with d == 0 contribute zero gradient (the distance is not
differentiable there).

Passing V makes the forward pass upload V to the scene and refit the BVH on the
device first, so optimising V in a loop keeps the BVH consistent.  The backward
pass reads the scene's current primitives, so it must see the geometry of its
own forward pass: the forward records the scene's positions version, and a
backward that finds another one (a second forward, an update of the vertices
or a new transfer in between) uploads the V of its forward pass again and
refits (the scene then holds that V), or raises RuntimeError when its forward
pass was given no V, instead of returning gradients of the wrong geometry.
"""
import torch

__all__ = ["ClosestPointDistance", "closest_point_distance"]


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(t, name, dim=None):
    if not t.is_cuda or t.dtype != torch.float32:
        raise TypeError(f"{name} must be a float32 CUDA tensor")
    if dim is not None and (t.dim() != 2 or t.shape[1] != dim):
        raise ValueError(f"{name} must have shape (n, {dim})")


class ClosestPointDistance(torch.autograd.Function):
    """forward(x, V, scene) -> (d, p, prim); gradients for x and V."""

    @staticmethod
    def forward(ctx, x, V, scene):
        dim = scene.DIM
        _check(x, "x", dim)
        x = x.contiguous()
        s = _stream()
        if V is not None:
            _check(V, "V", dim)
            V = V.contiguous()
            scene._update_positions_device(V.data_ptr(), V.shape[0], s)
            scene._refit_device(s)
        n = x.shape[0]
        d = torch.empty(n, dtype=torch.float32, device=x.device)
        p = torch.empty((n, dim), dtype=torch.float32, device=x.device)
        prim = torch.empty(n, dtype=torch.int32, device=x.device)
        slot = torch.empty(n, dtype=torch.int32, device=x.device)
        scene._closest_points_compact_device(x.data_ptr(), 0, n, d.data_ptr(), p.data_ptr(), prim.data_ptr(),
                                             slot.data_ptr(), s)
        ctx.scene = scene
        ctx.positions_version = scene.info("positions_version")
        ctx.n_vertices = None if V is None else V.shape[0]
        ctx.save_for_backward(x, slot, V if V is not None else torch.empty(0, device=x.device))
        ctx.mark_non_differentiable(p, prim)
        return d, p, prim

    @staticmethod
    def backward(ctx, grad_d, grad_p, grad_prim):
        x, slot, V = ctx.saved_tensors
        scene = ctx.scene
        dim = scene.DIM
        need_x = ctx.needs_input_grad[0]
        need_v = ctx.needs_input_grad[1] and ctx.n_vertices is not None
        if grad_d is None or not (need_x or need_v):
            return None, None, None
        if scene.info("positions_version") != ctx.positions_version:
            if ctx.n_vertices is None:
                raise RuntimeError("closest_point_distance: the scene's vertex positions changed between the forward "
                                   "pass and its backward pass, and the forward pass was given no V to restore")
            s = _stream()
            scene._update_positions_device(V.data_ptr(), V.shape[0], s)
            scene._refit_device(s)
            ctx.positions_version = scene.info("positions_version")
        grad_d = grad_d.contiguous().to(torch.float32)
        gx = torch.empty_like(x) if need_x else None
        gV = torch.zeros((ctx.n_vertices, dim), dtype=torch.float32, device=x.device) if need_v else None
        scene._distance_backward_device(x.data_ptr(), slot.data_ptr(), grad_d.data_ptr(), x.shape[0],
                                        gx.data_ptr() if gx is not None else 0,
                                        gV.data_ptr() if gV is not None else 0, _stream())
        return gx, gV, None


def closest_point_distance(scene, x, V=None):
    """d, p, prim = closest_point_distance(gpu_scene, x, V=None)

    x: (n, dim) float32 CUDA tensor of query points; V: optional (n_vertices, dim)
    float32 CUDA tensor of vertex positions in the order given to the Scene.
    Returns the unsigned distance d (differentiable in x and V), the closest
    points p and the original primitive indices (-1 where nothing was found)."""
    return ClosestPointDistance.apply(x, V, scene)
