"""Restarted GMRES and FP64-residual iterative refinement on device tensors.

The reference solves with scipy's GMRES in Python (acceptance_tests/dislocation.py:27-39);
this module keeps every vector on the GPU and calls the FMM through its device
entry point.

    x, info = gmres(apply, b, tol=1e-8, restart=50, maxiter=1000)
    x, info = refine(apply_fp64, apply_inner, b, tol=1e-8, inner_tol=1e-4)

gmres: GMRES(m) (Saad & Schultz 1986). The Arnoldi step orthogonalizes on the device
by classical Gram-Schmidt with one reorthogonalization (CGS2, Giraud et al. 2005: as
stable as modified Gram-Schmidt), two matrix-vector products against the basis held
as the rows of one tensor, so each iteration makes one device-to-host transfer (the
new Hessenberg column and norm) instead of one per basis vector. Givens rotations on
the small Hessenberg matrix run on the host in FP64.
refine: classic mixed-precision iterative refinement: r = b - A64 x with the
FP64 operator, correction from GMRES on the cheaper inner operator (e.g. the
FP32 FMM), x += e, until ||r|| / ||b|| <= tol.
"""
import math

import numpy as np
import torch


def gmres(apply, b, x0=None, tol=1e-8, restart=50, maxiter=1000, callback=None):
    """Solve A x = b; apply(v) -> A v on float64 device tensors.
    Returns (x, info) with info = {"iterations", "residual", "converged", "history"}."""
    b = b.contiguous()
    bnorm = float(torch.linalg.norm(b))
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    if bnorm == 0.0:
        return x, {"iterations": 0, "residual": 0.0, "converged": True, "history": [0.0]}
    hist = []
    it = 0
    r = b - apply(x) if x0 is not None else b.clone()
    beta = float(torch.linalg.norm(r))
    hist.append(beta / bnorm)
    m = restart
    V = torch.empty((m + 1, b.numel()), dtype=b.dtype, device=b.device)  # basis, one row per vector
    while it < maxiter and beta / bnorm > tol:
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        k_done = 0
        for k in range(m):
            w = apply(V[k]).contiguous()
            Vk = V[:k + 1]
            h = Vk @ w
            w = w - h @ Vk
            h2 = Vk @ w  # reorthogonalization pass
            w = w - h2 @ Vk
            col = torch.cat([h + h2, torch.linalg.norm(w).reshape(1)]).cpu().numpy()
            H[:k + 1, k] = col[:k + 1]
            hn = float(col[k + 1])
            H[k + 1, k] = hn
            for j in range(k):  # apply earlier rotations
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            k_done = k + 1
            res = abs(g[k + 1]) / bnorm
            hist.append(res)
            if callback is not None:
                callback(it, res)
            if res <= tol or it >= maxiter or hn == 0.0:
                break
            V[k + 1] = w / hn
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), g[:k_done])
        x.add_(torch.from_numpy(y).to(device=b.device, dtype=b.dtype) @ V[:k_done])
        r = b - apply(x)
        beta = float(torch.linalg.norm(r))
        hist[-1] = beta / bnorm  # true residual at the restart
        if beta == 0.0:
            break
    res = beta / bnorm
    return x, {"iterations": it, "residual": res, "converged": res <= tol, "history": hist}


def refine(apply_fp64, apply_inner, b, tol=1e-8, inner_tol=1e-4, restart=50, inner_maxiter=500,
           max_outer=30):
    """Iterative refinement with an FP64 residual around an inexact inner solver."""
    b = b.contiguous()
    bnorm = float(torch.linalg.norm(b))
    x = torch.zeros_like(b)
    outer = []
    inner_its = 0
    for k in range(max_outer):
        r = b - apply_fp64(x) if k else b.clone()
        res = float(torch.linalg.norm(r)) / bnorm if bnorm else 0.0
        outer.append(res)
        if res <= tol:
            break
        e, info = gmres(apply_inner, r, tol=inner_tol, restart=restart, maxiter=inner_maxiter)
        inner_its += info["iterations"]
        x.add_(e)
    else:
        r = b - apply_fp64(x)
        outer.append(float(torch.linalg.norm(r)) / bnorm)
    return x, {"outer": len(outer) - 1, "inner_iterations": inner_its, "residual": outer[-1],
               "converged": outer[-1] <= tol, "history": outer}


def constrained_operator(apply, cm, n_total):
    """v -> condense_vector(cm, apply(distribute_vector(cm, v, n_total))): the matvec of
    the constrained system that acceptance_tests/dislocation.py:27-39 hands to GMRES,
    here entirely on the device (cm from 3bem_amd.constraints.from_constraints)."""
    from . import constraints as C

    def mv(v):
        return C.condense_vector(cm, apply(C.distribute_vector(cm, v.contiguous(), n_total)).contiguous())
    return mv


def solve_constrained(apply, cm, b, n_total, tol=1e-8, restart=50, maxiter=1000, apply_inner=None,
                      inner_tol=1e-4):
    """Solve the constrained system condense(A distribute(x)) = b for the reduced x
    (b already condensed). With apply_inner (e.g. the mixed-precision FMM), FP64-residual
    iterative refinement around it; otherwise plain GMRES. Returns (x_reduced, info)."""
    A = constrained_operator(apply, cm, n_total)
    if apply_inner is None:
        return gmres(A, b, tol=tol, restart=restart, maxiter=maxiter)
    A_in = constrained_operator(apply_inner, cm, n_total)
    return refine(A, A_in, b, tol=tol, inner_tol=inner_tol, restart=restart)
