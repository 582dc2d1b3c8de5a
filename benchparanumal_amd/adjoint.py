"""Differentiable variable-coefficient BP5 solve (an extension; the
reference's BP5 coefficient is constant, okl/bp5AxHex3D.okl:415-569).

    u = A(kappa)^{-1} b,   A(kappa) = sum_e Z_e^T [D^T (kappa G) D + lambda wJ] Z_e

with kappa a nodal field on the box's node lattice. The forward pass is
libbpx's device-resident CG on the kappa-scaled operator; the backward pass
solves the adjoint system A lam = dL/du (A is symmetric) with the same CG
and forms dL/dkappa = -sum over node copies of (D lam).G(D u) with the HIP
gradient kernel, and dL/db = lam. Every step runs in libbpx; torch holds
the device buffers and the tape.
"""
import numpy as np
import torch

from . import capi


def lattice_shape(mesh):
    """(LZ, LY, LX) of the node lattice of a capi.Mesh; the flat lattice
    index X + LX*(Y + LY*Z) is the C-order index of this shape."""
    return (mesh.nz * mesh.N + 1, mesh.ny * mesh.N + 1, mesh.nx * mesh.N + 1)


def smooth_lattice_field(shape, seed, passes=3, std=0.5):
    """log-coefficient field on the lattice: g ~ N(0,1) (libbpx's seeded
    SplitMix64 + Box-Muller stream), `passes` rounds of nearest-node averaging
    (each node with its up to 6 lattice neighbours), scaled to the given
    standard deviation. Returns g (float64, shape `shape`)."""
    n = int(np.prod(shape))
    g = capi.random_normal(seed, n).reshape(shape)
    for _ in range(passes):
        acc = g.copy()
        cnt = np.ones_like(g)
        for ax in range(3):
            for sgn in (1, -1):
                sl_dst = [slice(None)] * 3
                sl_src = [slice(None)] * 3
                if sgn == 1:
                    sl_dst[ax], sl_src[ax] = slice(1, None), slice(None, -1)
                else:
                    sl_dst[ax], sl_src[ax] = slice(None, -1), slice(1, None)
                acc[tuple(sl_dst)] += g[tuple(sl_src)]
                cnt[tuple(sl_dst)] += 1
        g = acc / cnt
    g = g - g.mean()
    return g * (std / g.std())


def _solve_relative(bp, rhs, tol, maxit):
    """x = A^{-1} rhs by bp's CG with the right-hand side scaled to unit
    norm, so that tol is the relative residual: the CG's stopping test,
    ||r|| <= max(tol ||r0||, tol) (libs/core/linearSolverCG.cpp:84-88), has
    an absolute floor that a small right-hand side would hit first.
    Returns (x, iterations)."""
    nrm = float(torch.linalg.vector_norm(rhs))
    x = torch.zeros_like(rhs)
    if nrm == 0.0:
        return x, 0
    r = rhs / nrm
    its, _ = bp.solve(x, r, tol=tol, maxit=maxit, history=False)
    return x.mul_(nrm), its


class KappaSolve(torch.autograd.Function):
    """u = A(kappa)^{-1} b on a BP5 capi.BP; differentiable in kappa and b."""

    @staticmethod
    def forward(ctx, kappa, b, bp, tol, maxit):
        if kappa.dtype != torch.float64 or b.dtype != torch.float64:
            raise TypeError("KappaSolve works in float64")
        kappa = kappa.contiguous()
        bp.set_kappa(kappa)
        u, its = _solve_relative(bp, b.detach(), tol, maxit)
        if its >= maxit:
            raise RuntimeError(f"KappaSolve: CG did not reach {tol:g} in {maxit} iterations")
        ctx.bp, ctx.tol, ctx.maxit = bp, tol, maxit
        ctx.save_for_backward(kappa, u)
        ctx.iterations = its
        return u

    @staticmethod
    def backward(ctx, grad_u):
        kappa, u = ctx.saved_tensors
        bp = ctx.bp
        bp.set_kappa(kappa)  # the operator may have been reset since forward
        lam, its = _solve_relative(bp, grad_u.detach().contiguous(), ctx.tol, ctx.maxit)
        if its >= ctx.maxit:
            raise RuntimeError(f"KappaSolve: adjoint CG did not reach {ctx.tol:g} in {ctx.maxit} iterations")
        grad_kappa = torch.empty_like(kappa)
        bp.kappa_gradient(u, lam, grad_kappa)
        return grad_kappa, lam, None, None, None


def kappa_solve(bp, kappa, b, tol=1e-12, maxit=100000):
    """Differentiable u = A(kappa)^{-1} b (see KappaSolve)."""
    return KappaSolve.apply(kappa, b, bp, tol, maxit)
