"""Differentiable FMM matvec for torch.

    u = fmm_matvec(op, x, obs_locs)      # x: (R*N_src,) float64 CUDA tensor
                                         # obs_locs: (N_obs, 3) float64 CUDA tensor or None

Backward (3bem has no autograd; this is the hand-written backward of apply()):
  * dL/dx        = A^T g      -- the exact transpose of the FMM operator
                                (FMMOperator.apply_adjoint: every stage transposed,
                                source and target roles swapped);
  * dL/dobs_locs = obs_grad   -- gradient kernels for the near field and the
                                local expansions (FMMOperator.obs_grad). The
                                forward's multipole and local expansions are
                                reused when no other apply has run on the operator
                                in between, and recomputed otherwise.
The plan (trees, surfaces, interaction lists) is held fixed: moving observation
points only changes the point coordinates inside their cells, so the forward is
a smooth function of them and central differences of the frozen-plan forward
reproduce the gradient. A point that leaves its leaf cell's ball (centre,
radius |half width|) makes the forward raise (the frozen plan no longer holds it).

The operator holds one set of observation points (shared with its clones). The
backward runs at the points its forward used: when another forward has moved
them since, the backward moves them back first.
"""
import torch


class _FMMMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, obs_locs, op):
        obs = None
        if obs_locs is not None:
            obs = obs_locs.detach().to(dtype=torch.float64).contiguous()
            op.set_obs_locs(obs if obs.is_cuda else obs.numpy())
        ctx.op = op
        ctx.save_for_backward(x)
        ctx.obs = obs
        ctx.obs_version = op.obs_version()
        y = op.apply(x.contiguous())
        ctx.token = op.forward_token()
        return y

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        op = ctx.op
        g = g.contiguous()
        if op.obs_version() != ctx.obs_version:
            if ctx.obs is None:
                raise RuntimeError("the operator's observation points moved after this forward, which did not "
                                   "name them: pass obs_locs to fmm_matvec to differentiate through moves")
            op.set_obs_locs(ctx.obs if ctx.obs.is_cuda else ctx.obs.numpy())
            ctx.obs_version = op.obs_version()
        gx = op.apply_adjoint(g) if ctx.needs_input_grad[0] else None
        gobs = None
        if ctx.obs is not None and ctx.needs_input_grad[1]:
            gobs = op.obs_grad(x.contiguous(), g, token=ctx.token)
        return gx, gobs, None


def fmm_matvec(op, x, obs_locs=None):
    """Differentiable y = A(obs_locs) x on the GPU (see module docstring)."""
    return _FMMMatvec.apply(x, obs_locs, op)


def forward_model_loss(op, slip, u_obs, obs_locs=None):
    """0.5 * ||A(obs) slip - u_obs||^2 (the slip-inversion misfit)."""
    u = fmm_matvec(op, slip, obs_locs)
    return 0.5 * torch.sum((u - u_obs) ** 2)
