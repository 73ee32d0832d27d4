"""torch.autograd harness for the plane-cost adjoint.

``PlaneCost.apply(planes, pose_delta, backend)`` returns the per-pixel
aggregated cost of `planes` (shape (..., 4): depth, nx, ny, nz) under source
poses perturbed on the left by `pose_delta` (nsrc x 6: rotation vector w,
translation tau; Y -> Exp(w) Y + tau in source-camera coordinates), and a
validity mask.  Invalid pixels cost 0 and get no gradient.  Backward asks the
backend for d(sum_p g_p cost_p) with g = the incoming gradient.

Backends implement ``evaluate(planes, pose_delta, weights, grad)`` on numpy
arrays and return (cost, grad_plane, grad_pose), grad_pose being the gradient
of a further left perturbation at the perturbed poses (what the kernels
compute); the harness converts it to the gradient in pose_delta's own
coordinates (``delta_gradient``).  ``GpuBackend`` runs the HIP
kernels (recon_mvs_hip_plane_cost_grad), so a torch optimiser can refine
depth maps and source poses through them.
"""
from __future__ import annotations

import numpy as np
import torch

from .plane import PlaneMVS

COST_INVALID = 1e29  # costs at or above this are invalid


def so3_left_jacobian(phi):
    """J_l(phi): Exp(phi + h) = Exp(J_l(phi) h) Exp(phi) to first order in h."""
    phi = np.asarray(phi, np.float64)
    th = float(np.linalg.norm(phi))
    K = np.array([[0.0, -phi[2], phi[1]], [phi[2], 0.0, -phi[0]], [-phi[1], phi[0], 0.0]])
    if th < 1e-6:
        return np.eye(3) + 0.5 * K + (K @ K) / 6.0
    return (np.eye(3) + ((1.0 - np.cos(th)) / th**2) * K
            + ((th - np.sin(th)) / th**3) * (K @ K))


def delta_gradient(grad_left, pose_delta):
    """Gradient w.r.t. pose_delta = (w, tau) from the gradient (g_w, g_tau) of
    a further left perturbation at the perturbed pose: with Y' = Exp(w) Y0 +
    tau, dY'/dw = -[Y' - tau]x J_l(w), so g_delta_w = J_l(w)^T (g_w - tau x
    g_tau) and g_delta_tau = g_tau."""
    g = np.array(grad_left, np.float64, copy=True)
    d = np.asarray(pose_delta, np.float64)
    for k in range(g.shape[0]):
        w, tau = d[k, :3], d[k, 3:]
        gw, gt = g[k, :3], g[k, 3:]
        g[k, :3] = so3_left_jacobian(w).T @ (gw - np.cross(tau, gt))
    return g


class GpuBackend:
    """Cost and gradients of every pixel of the depth-map slot of `ref`:
    planes of shape (H, W, 4)."""

    def __init__(self, pm: PlaneMVS, ref: int, srcs, params):
        self.pm, self.ref, self.srcs, self.params = pm, int(ref), list(srcs), params

    def evaluate(self, planes, pose_delta, weights=None, grad=False):
        planes = np.asarray(planes, np.float32)
        h, w = self.pm.shapes[self.ref]
        assert planes.shape == (h, w, 4), planes.shape
        self.pm.set_state(self.ref, self.srcs, self.params, planes[..., 0], planes[..., 1:])
        cost, gp, gs = self.pm.cost_grad(self.ref, self.srcs, self.params, pose_delta=pose_delta,
                                         weights=weights, grad=grad, pose_grad=grad)
        return cost.astype(np.float64), gp, gs


class PlaneCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, planes, pose_delta, backend):
        p = planes.detach().cpu().numpy()
        d = pose_delta.detach().cpu().numpy().astype(np.float64)
        cost, _, _ = backend.evaluate(p, d, None, grad=False)
        valid = cost < COST_INVALID
        out = torch.as_tensor(np.where(valid, cost, 0.0), dtype=planes.dtype, device=planes.device)
        vmask = torch.as_tensor(valid, device=planes.device)
        ctx.backend = backend
        ctx.save_for_backward(planes, pose_delta)
        ctx.mark_non_differentiable(vmask)
        return out, vmask

    @staticmethod
    def backward(ctx, grad_cost, grad_valid):
        planes, pose_delta = ctx.saved_tensors
        w = grad_cost.detach().cpu().numpy()
        _, gp, gs = ctx.backend.evaluate(planes.detach().cpu().numpy(),
                                         pose_delta.detach().cpu().numpy().astype(np.float64),
                                         w, grad=True)
        gpl = torch.as_tensor(np.asarray(gp), dtype=planes.dtype, device=planes.device)
        gs = delta_gradient(gs, pose_delta.detach().cpu().numpy())
        gpo = torch.as_tensor(gs, dtype=pose_delta.dtype, device=pose_delta.device)
        return gpl, gpo, None


def plane_cost(planes, pose_delta, backend):
    """(cost, valid) of `planes` under `pose_delta`, differentiable in both."""
    return PlaneCost.apply(planes, pose_delta, backend)
