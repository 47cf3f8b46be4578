"""Differentiation through the collocation solve (torch.autograd).

The reference differentiates ``pde_solver`` with ``jax.grad`` through
lineax's implicit linear-solve rule and XLA's ``inv`` (as in
``demos/Laplace/10_laplace_with_diff_phys.py:89-104``): for B u = q the
cotangent of q is B^-T dL/du.  Here the solve is the augmented system
K c = [q; 0] and the adjoint reuses the LU factors of K through the
transposed triangular sweeps (``rbfx_getrs_*`` with trans=1):

    c = K^-1 b             ->  db = K^-T dc
    u = (A c)[:N]          ->  dc += A^T [du; 0]       (matrix-free, HIP)
    f_coeffs = A^-1 [f; 0] ->  df = (A^-T d)[:N]       (LU of A, transposed)
    y = E_op(c) at points  ->  dc += E_op^T dy         (rbfx_eval_t)

``checkpointed_rollout`` gives backpropagation through time with
recomputation: only every k-th state is stored in the forward pass, and each
segment is recomputed (reusing the cached factorisations) during backward.
"""
import torch

from . import _lib as L
from . import engine as E


class KSolve(torch.autograd.Function):
    """c = K^-1 b with the cached factorisation; backward solves with K^T."""

    @staticmethod
    def forward(ctx, b, fac):
        ctx.fac = fac
        return fac.solve(b, trans=False)

    @staticmethod
    def backward(ctx, g):
        return ctx.fac.solve(g.contiguous(), trans=True), None


class AValues(torch.autograd.Function):
    """u = (A c)[:N]: nodal values of the field with coefficients c
    (``updes/assembly.py:62-85`` applied matrix-free)."""

    @staticmethod
    def forward(ctx, c, prob_a):
        ctx.prob_a = prob_a
        y = E.matvec(prob_a, c, L.MAT_A, trans=False)
        return y[:prob_a.dc.N].clone()

    @staticmethod
    def backward(ctx, g):
        p = ctx.prob_a
        ge = torch.cat([g.to(torch.float64), torch.zeros(p.M, dtype=torch.float64, device=g.device)])
        return E.matvec(p, ge, L.MAT_A, trans=True), None


class ASolve(torch.autograd.Function):
    """coeffs = A^-1 f_ext (``updes/assembly.py:404-410``)."""

    @staticmethod
    def forward(ctx, f_ext, fac_a):
        ctx.fac = fac_a
        return fac_a.solve(f_ext, trans=False)

    @staticmethod
    def backward(ctx, g):
        return ctx.fac.solve(g.contiguous(), trans=True), None


class EvalOp(torch.autograd.Function):
    """Values of operator ``op`` of the field ``coeffs`` at ``points``."""

    @staticmethod
    def forward(ctx, coeffs, kind, param, centers, points, op):
        ctx.save = (kind, param, centers, points, op, coeffs.numel() - centers.shape[0])
        return E.eval_ops(kind, param, centers, coeffs, points, 1 << op)[op].clone()

    @staticmethod
    def backward(ctx, g):
        kind, param, centers, points, op, M = ctx.save
        return E.eval_t(kind, param, centers, M, points, op, g), None, None, None, None, None


class RhsAssemble(torch.autograd.Function):
    """q (N,) from the interior values q_int (Ni,) and the boundary values bc
    (N - Ni,) in node order, through rbfx_rhs (``assembly.py:434-485``); the
    backward is the restriction of dL/dq to the same rows."""

    @staticmethod
    def forward(ctx, q_int, bc, problem):
        ctx.Ni = q_int.numel()
        dc = problem.dc
        return E.rhs(problem, q_int.detach().contiguous(), bc.detach().contiguous())[:dc.N]

    @staticmethod
    def backward(ctx, g):
        return g[:ctx.Ni], g[ctx.Ni:], None


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, step, nsteps, every, *params):
        ctx.step, ctx.nsteps, ctx.every = step, nsteps, every
        ctx.params = params
        checkpoints = {0: u0.detach()}
        u = u0.detach()
        with torch.no_grad():
            for n in range(1, nsteps + 1):
                u = step(u)
                if n % every == 0 and n < nsteps:
                    checkpoints[n] = u.detach()
        ctx.checkpoints = checkpoints
        ctx.save_for_backward(*[p for p in params if torch.is_tensor(p)])
        return u.clone()

    @staticmethod
    def backward(ctx, g):
        step, nsteps, every = ctx.step, ctx.nsteps, ctx.every
        params = [p for p in ctx.params]
        pgrads = [torch.zeros_like(p) if torch.is_tensor(p) and p.requires_grad else None for p in params]
        starts = sorted(ctx.checkpoints)
        for s in reversed(starts):
            e = min(s + every, nsteps)
            with torch.enable_grad():
                u = ctx.checkpoints[s].detach().requires_grad_(True)
                v = u
                for _ in range(s, e):
                    v = step(v)
                inputs = [u] + [p for p in params if torch.is_tensor(p) and p.requires_grad]
                grads = torch.autograd.grad(v, inputs, g, allow_unused=True)
            g = grads[0] if grads[0] is not None else torch.zeros_like(u)
            k = 1
            for i, p in enumerate(params):
                if pgrads[i] is not None:
                    if grads[k] is not None:
                        pgrads[i] = pgrads[i] + grads[k]
                    k += 1
        return (g, None, None, None) + tuple(pgrads)


def checkpointed_rollout(step, u0, nsteps, every=20, params=()):
    """u_nsteps = step^nsteps(u0), differentiable w.r.t. u0 (and ``params``,
    tensors the step closes over) with O(nsteps/every) stored states."""
    if nsteps <= 0:
        return u0
    return _Rollout.apply(u0, step, int(nsteps), int(max(1, every)), *params)
