"""torch.autograd wrapper of the differentiable integrate (test harness only).

IntegrateSurfaceLoss(xi, depths, vol, rt0, points, weights) resets `vol`,
integrates `depths` seen from poses Exp(xi_f) T0_f with bilinear sampling
(recon_tsdf_diff_integrate), and returns the surface loss
(recon_tsdf_surface_loss).  backward runs recon_tsdf_diff_adjoint: dL/dxi
(left perturbation, exact at xi = 0, where pose refinement re-linearises
every step) and dL/ddepths."""
import torch


def _hat(w):
    z = torch.zeros_like(w[..., 0])
    return torch.stack([torch.stack([z, -w[..., 2], w[..., 1]], -1),
                        torch.stack([w[..., 2], z, -w[..., 0]], -1),
                        torch.stack([-w[..., 1], w[..., 0], z], -1)], -2)


def se3_exp(xi):
    """(F, 6) float64 xi = (rho, omega) -> (F, 3, 3) R, (F, 3) t of Exp(xi)
    (Rodrigues; series near 0)."""
    rho, om = xi[:, :3], xi[:, 3:]
    th = torch.linalg.norm(om, dim=1)[:, None, None]
    K = _hat(om)
    K2 = K @ K
    eye = torch.eye(3, dtype=xi.dtype, device=xi.device).expand_as(K)
    small = th < 1e-6
    ths = torch.where(small, torch.ones_like(th), th)
    A = torch.where(small, 1 - th ** 2 / 6, torch.sin(ths) / ths)
    B = torch.where(small, 0.5 - th ** 2 / 24, (1 - torch.cos(ths)) / ths ** 2)
    Cc = torch.where(small, 1.0 / 6 - th ** 2 / 120, (ths - torch.sin(ths)) / ths ** 3)
    R = eye + A * K + B * K2
    V = eye + B * K + Cc * K2
    return R, (V @ rho[:, :, None])[:, :, 0]


def left_update(rt0, xi):
    """(F, 12) float32 [R | t] -> Exp(xi) T0 as (F, 12) float32."""
    R0 = rt0[:, :9].double().reshape(-1, 3, 3)
    t0 = rt0[:, 9:].double()
    Re, te = se3_exp(xi.double())
    R = Re @ R0
    t = (Re @ t0[:, :, None])[:, :, 0] + te
    return torch.cat([R.reshape(-1, 9), t], 1).float().contiguous()


class IntegrateSurfaceLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, depths, vol, rt0, points, weights):
        rt = left_update(rt0, xi.detach())
        vol.reset()
        vol.diff_integrate(depths, rt)
        loss, _, ids, lam = vol.surface_loss(points, weights)
        ctx.vol = vol
        ctx.save_for_backward(depths, rt, ids, lam)
        return loss[0].clone()

    @staticmethod
    def backward(ctx, g):
        depths, rt, ids, lam = ctx.saved_tensors
        scaled = (lam.double() * g.double()).float()
        gp, gd = ctx.vol.diff_adjoint(depths, rt, ids, scaled, want_depth=ctx.needs_input_grad[1])
        return gp, gd, None, None, None, None
