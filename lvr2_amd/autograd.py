"""Optional PyTorch-ROCm binding of the differentiation form: the Hoppe signed distance at
query positions as a differentiable function of the point positions.

    proj = hoppe_sdf(points, queries, kn=30, kd=10, flip=(0, 0, 0), voxel=0.05)
    loss = (proj ** 2).sum(); loss.backward()      # points.grad via the hand-written adjoint

Forward: a HipSurface over ``points`` (device tensor), PCA normals with lvr2's adaptive k
(AdaptiveKSearchSurface::calculateSurfaceNormals without interpolation, tcc:158-284) and
AdaptiveKSearchSurface::distance (tcc:414-457).  Backward: lvrhip_distances_backward_device,
with neighbour sets, adaptive k and flip signs held fixed.  Gradients flow to ``points`` only.
"""
from __future__ import annotations

import torch

from .surface import HipSurface


class HoppeSDF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, queries, kn: int, kd: int, flip, voxel: float):
        s = HipSurface(points.detach())
        s.setKn(kn)
        s.setKi(0)
        s.setKd(kd)
        s.setFlippoint(*flip)
        s.calculateNormals()
        q = queries.detach().to(torch.float32).contiguous()
        proj, _, _ = s.distances(q, voxel)
        ctx.surface = s
        ctx.save_for_backward(q)
        return proj

    @staticmethod
    def backward(ctx, grad_proj):
        (q,) = ctx.saved_tensors
        g = None
        if ctx.needs_input_grad[0]:
            g = ctx.surface.distancesBackward(q, grad_proj.contiguous())
        ctx.surface = None
        return g, None, None, None, None, None


def hoppe_sdf(points: torch.Tensor, queries: torch.Tensor, kn: int = 30, kd: int = 10, flip=(0.0, 0.0, 0.0),
              voxel: float = 0.05) -> torch.Tensor:
    """proj of every query (its signed distance to the local tangent plane)."""
    return HoppeSDF.apply(points, queries, int(kn), int(kd), tuple(float(x) for x in flip), float(voxel))
