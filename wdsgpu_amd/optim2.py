# -*- coding: utf-8 -*-
"""First-order calibration helpers on the HIP path (dgga/optim2.py).

- ``multistart_heads``: B starting points x F training frames as B F
  scenarios of one batched forward + implicit adjoint
  (dgga/optim2.py:77-118); the cost is still B model evaluations per step,
  only the wall clock is shared.
- ``schur_nodal_diag`` / ``schur_diag_precond`` / ``gn_diag_from_S``: a
  diagonal metric for the roughness C.  With A = A21 D^-1 A12 the converged
  GGA matrix (D = diag(hgrad)), the head response to r_k solves
  A dH = (dr_k / D_kk) a_k with a_k = e_n2 - e_n1, and dH/dC_k =
  (dr/dC)_k (dr_k / D_kk) A^-1 a_k.  The Jacobi approximation A^-1 ~
  diag(A)^-1 gives M_k = mean over frames of prefac_k^2 (m1 / A_n1n1^2 + m2 /
  A_n2n2^2) without any linear solve (dgga/optim2.py:137-215); the exact
  Gauss-Newton diagonal diag(S^T S) / n comes from the sensitivity matrix.
  Both take a Tikhonov term and a relative floor.
- ``lbfgs_precond``: L-BFGS with the initial inverse Hessian gamma M^-1 and
  torch's strong-Wolfe line search, counting every function evaluation
  (dgga/optim2.py:218-322).
"""
from __future__ import annotations

import numpy as np
import torch

from .autodiff import ImplicitGGASolveGPU, ImplicitGGASolve

__all__ = ["multistart_heads", "schur_nodal_diag", "schur_diag_precond", "gn_diag_from_S",
           "lbfgs_precond", "lhs_starts"]


def multistart_heads(solver, r_starts, d_frames, rh_frames, ke=None, accuracy=1e-6,
                     max_iter=30, polish_steps=4, adjoint="gpu"):
    """Heads and flows of B starts x F frames in one forward (and, through
    autograd, one adjoint): r_starts [B,L] (may require grad), d_frames /
    rh_frames [F,N]; returns (head [B,F,N], flow [B,F,L]).  accuracy is the
    GGA stopping test before the Newton polish.  adjoint="cpu" takes B = 1
    (the reference's CPU adjoint has a batch-shared r_hw)."""
    s = solver
    dev, dt = s.device, s.dtype
    as_t = lambda x: x.to(dtype=dt, device=dev) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
    d_t, rh_t = as_t(d_frames), as_t(rh_frames)
    if d_t.dim() != 2 or rh_t.dim() != 2:
        raise ValueError("d_frames and rh_frames must be [F,N]")
    F = int(d_t.shape[0])
    r_starts = as_t(r_starts)
    if r_starts.dim() != 2 or r_starts.shape[1] != s.L:
        raise ValueError("r_starts must be [B,L] with L=%d" % s.L)
    B = int(r_starts.shape[0])
    ke_t = s.node_ke_default if ke is None else as_t(ke)
    if adjoint == "cpu":
        if B != 1:
            raise ValueError("adjoint='cpu' has a batch-shared r_hw: one start only")
        head, flow, _ = ImplicitGGASolve.apply(d_t, rh_t, ke_t, r_starts[0], s, 1e-12, 20,
                                               int(polish_steps))
        return head.reshape(1, F, -1), flow.reshape(1, F, -1)
    if adjoint != "gpu":
        raise ValueError("adjoint must be 'gpu' or 'cpu'")
    rr = r_starts.repeat_interleave(F, dim=0)
    head, flow, _ = ImplicitGGASolveGPU.apply(d_t.repeat(B, 1), rh_t.repeat(B, 1), ke_t, rr, s,
                                              accuracy, max_iter, None, "dense", "dense",
                                              int(polish_steps))
    return head.reshape(B, F, -1), flow.reshape(B, F, -1)


def lhs_starts(n, lo, hi, dim, seed):
    """Latin-hypercube starting points [n, dim] in [lo, hi]."""
    rng = np.random.default_rng(seed)
    u = np.empty((n, dim))
    for j in range(dim):
        u[:, j] = (rng.permutation(n) + rng.uniform(size=n)) / n
    return lo + (hi - lo) * u


def _ends(s):
    """Junction slots of each link's ends and whether each end is a junction."""
    n = s.net
    jidx = np.full(s.N, -1, dtype=np.int64)
    jidx[s.junc_nodes] = np.arange(s.Nj)
    j1 = jidx[np.asarray(n.link_n1, dtype=np.int64)]
    j2 = jidx[np.asarray(n.link_n2, dtype=np.int64)]
    return j1, j2, j1 >= 0, j2 >= 0


def schur_nodal_diag(solver, q, r_hw, status=None):
    """Diagonal [Nj] of the converged GGA matrix A = A21 D^-1 A12 at flows q
    [L] (dgga/optim2.py:121-134): A_ii = sum of the conductances 1/hgrad of
    the links at junction i, from the kernels' link coefficients.  Returns
    (Adiag, P, dr) with P = 1/hgrad [L] and dr = d hloss / d r [L]."""
    s = solver
    P, _, dr = s.link_coeffs(q, r_hw, status)
    P, dr = P[0].cpu().numpy(), dr[0].cpu().numpy()
    j1, j2, m1, m2 = _ends(s)
    Adiag = np.bincount(j1[m1], weights=P[m1], minlength=s.Nj) + \
        np.bincount(j2[m2], weights=P[m2], minlength=s.Nj)
    return Adiag, P, dr


def schur_diag_precond(solver, q_frames, r_hw, C_free, free_idx, w_reg=0.0, floor_rel=1e-6,
                       status=None):
    """Jacobi-Schur metric M [n_free] in C space, no linear solve
    (dgga/optim2.py:137-199): M_k = mean_f prefac_k^2 (m1 / A_n1n1^2 + m2 /
    A_n2n2^2) + w_reg with prefac_k = (Hexp r_k / C_k) dr_k / hgrad_k (0 on
    the RQtol clamp), then floored at floor_rel max(M).  q_frames [F,L]
    converged flows, r_hw [L], C_free [n_free], free_idx [n_free] links.
    Returns (M, info)."""
    s = solver
    q_frames = np.atleast_2d(np.asarray(q_frames, dtype=np.float64))
    f_idx = np.asarray(free_idx, dtype=np.int64)
    C_free = np.asarray(C_free, dtype=np.float64)
    r = np.asarray(r_hw.detach().cpu().numpy() if isinstance(r_hw, torch.Tensor) else r_hw,
                   dtype=np.float64)
    j1, j2, m1, m2 = _ends(s)
    j1, j2, m1, m2 = j1[f_idx], j2[f_idx], m1[f_idx], m2[f_idx]
    acc = np.zeros(f_idx.size)
    for b in range(q_frames.shape[0]):
        Adiag, P, dr = schur_nodal_diag(s, q_frames[b], r, status)
        inv1 = np.where(m1, 1.0 / np.maximum(Adiag[np.where(m1, j1, 0)], 1e-300), 0.0)
        inv2 = np.where(m2, 1.0 / np.maximum(Adiag[np.where(m2, j2, 0)], 1e-300), 0.0)
        prefac = (s.hexp * r[f_idx] / C_free) * (dr[f_idx] * P[f_idx])
        acc += prefac ** 2 * (inv1 ** 2 + inv2 ** 2)
    M = acc / q_frames.shape[0] + w_reg
    mx = float(M.max())
    fl = floor_rel * mx
    n_floored = int((M < fl).sum())
    M = np.maximum(M, fl)
    return M, dict(kind="schur_jacobi", max=mx, min=float(M.min()), floor_rel=floor_rel,
                   n_floored=n_floored, cond=float(M.max() / M.min()), w_reg=w_reg)


def gn_diag_from_S(S, n_obs, w_reg=0.0, floor_rel=1e-6):
    """Exact Gauss-Newton diagonal diag(S^T S) / n_obs + w_reg, floored at
    floor_rel max (dgga/optim2.py:202-215).  Returns (M, info)."""
    S = np.asarray(S, dtype=np.float64)
    M = np.einsum("ij,ij->j", S, S) / float(n_obs) + w_reg
    mx = float(M.max())
    fl = floor_rel * mx
    n_floored = int((M < fl).sum())
    M = np.maximum(M, fl)
    return M, dict(kind="gn_exact", max=mx, min=float(M.min()), floor_rel=floor_rel,
                   n_floored=n_floored, cond=float(M.max() / M.min()), w_reg=w_reg)


def lbfgs_precond(fun, x0, minv=None, max_iter=100, max_eval=None, history=30, tol_grad=1e-13,
                  tol_change=1e-18, c1=1e-4, c2=0.9, max_ls=25, callback=None, f0=None,
                  g0=None):
    """Diagonally preconditioned L-BFGS (dgga/optim2.py:218-322): the
    two-loop recursion with H0 = gamma M^-1 (gamma = s.y / y.M^-1 y), torch's
    strong-Wolfe line search, every evaluation counted (line-search trials
    included).  fun(x) -> (f, g) on float64 numpy; minv = diag(M^-1) [n] or
    None (plain L-BFGS).  Returns dict(x, f, n_eval, n_iter, stop, f_hist)."""
    from torch.optim.lbfgs import _strong_wolfe

    x = np.asarray(x0, dtype=np.float64).copy()
    minv = None if minv is None else np.asarray(minv, dtype=np.float64)
    max_eval = max_eval if max_eval is not None else max_iter * 5
    n_eval = [0]

    def ev(xx):
        f, g = fun(np.asarray(xx, dtype=np.float64))
        n_eval[0] += 1
        return float(f), np.asarray(g, dtype=np.float64).copy()

    if f0 is None or g0 is None:
        f, g = ev(x)
    else:
        f, g = float(f0), np.asarray(g0, dtype=np.float64).copy()
    f_hist = [f]
    Ss, Ys, rhos = [], [], []
    stop = "max_iter"
    n_iter = 0
    apply_minv = (lambda v: v * minv) if minv is not None else (lambda v: v)
    for it in range(max_iter):
        if np.max(np.abs(g)) <= tol_grad:
            stop = "tol_grad"
            break
        if n_eval[0] + max_ls > max_eval:
            stop = "max_eval"
            break
        qv = g.copy()
        alpha = []
        for sv, yv, rho in zip(reversed(Ss), reversed(Ys), reversed(rhos)):
            a = rho * float(sv @ qv)
            alpha.append(a)
            qv -= a * yv
        alpha.reverse()
        if Ss:
            den = float(Ys[-1] @ apply_minv(Ys[-1]))
            gamma = float(Ss[-1] @ Ys[-1]) / den if den > 0 else 1.0
        else:
            gamma = 1.0
        z = gamma * apply_minv(qv)
        for sv, yv, rho, a in zip(Ss, Ys, rhos, alpha):
            z += sv * (a - rho * float(yv @ z))
        d = -z
        gtd = float(g @ d)
        if gtd > -1e-300:
            d = -apply_minv(g)
            Ss, Ys, rhos = [], [], []
            gtd = float(g @ d)
            if gtd > -1e-300:
                stop = "no_descent"
                break
        t0 = 1.0 if (it > 0 or minv is not None or Ss) else \
            min(1.0, 1.0 / max(float(np.abs(g).sum()), 1e-300))

        def obj(xx, tt, dd):
            fn, gn = ev((xx + tt * dd).numpy())
            return fn, torch.from_numpy(gn)

        f_new, g_new, t, _ = _strong_wolfe(obj, torch.from_numpy(x), t0, torch.from_numpy(d), f,
                                           torch.from_numpy(g), gtd, c1=c1, c2=c2,
                                           tolerance_change=1e-12, max_ls=max_ls)
        f_new = float(f_new)
        g_new = np.asarray(g_new.detach().cpu().numpy(), dtype=np.float64)
        x_new = x + float(t) * d
        s_vec, y_vec = x_new - x, g_new - g
        ys = float(y_vec @ s_vec)
        if ys > 1e-10 * float(np.linalg.norm(s_vec) * np.linalg.norm(y_vec) + 1e-300):
            if len(Ss) == history:
                Ss.pop(0)
                Ys.pop(0)
                rhos.pop(0)
            Ss.append(s_vec)
            Ys.append(y_vec)
            rhos.append(1.0 / ys)
        moved = float(np.max(np.abs(s_vec)))
        x, f, g = x_new, f_new, g_new
        f_hist.append(f)
        n_iter += 1
        if callback is not None:
            callback(it, x, f, n_eval[0])
        if moved <= tol_change:
            stop = "tol_change"
            break
        if n_eval[0] >= max_eval:
            stop = "max_eval"
            break
    return dict(x=x, f=f, n_eval=n_eval[0], n_iter=n_iter, stop=stop, f_hist=f_hist)
