# -*- coding: utf-8 -*-
"""Roughness calibration helpers on the HIP path (dgga/calib.py).

- ``hw_resistance`` / ``dr_dC``: the C -> r map and its chain factor
  (dgga/calib.py:64-108), here the elementwise HIP kernels of
  ``autodiff.hw_resistance``;
- ``RoughnessParam``: the box-constrained parametrisation
  C = C_lo + (C_hi - C_lo) sigmoid(theta) of the pipes under calibration,
  whose ``r_hw()`` feeds ``implicit_solve`` (dgga/calib.py:116-258);
- ``clamped_mask``, ``dead_branch_mask``, ``identifiability``: which pipes a
  sensor layout cannot calibrate and why (dgga/calib.py:264-437), from the
  HIP forward's converged flows and the HIP sensitivity matrix.
"""
from __future__ import annotations

import numpy as np
import torch

from . import autodiff
from .network import CVPIPE, PIPE

__all__ = ["hw_resistance", "RoughnessParam", "dr_dC", "clamped_mask", "dead_branch_mask",
           "identifiability", "REASONS"]

# CVPIPE and PIPE carry a Hazen-Williams resistance (dgga/calib.py:53-55)
_PIPE_TYPES = (CVPIPE, PIPE)
REASONS = ("informative", "clamped", "dead_branch", "closed", "non_pipe", "unobservable")


def _dev_tensor(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)


def hw_resistance(C, len_ft, diam_ft, device=None):
    """r = 4.727 len / C^1.852 / diam^4.871 (dgga/calib.py:64-86),
    differentiable with respect to C, on the HIP device: C [..., L] with
    len_ft, diam_ft [L] (array-likes go to ``device``, default the first
    tensor's device or cuda:0).  CPU tensors are refused."""
    dev = device
    if dev is None:
        dev = next((x.device for x in (C, len_ft, diam_ft) if isinstance(x, torch.Tensor)),
                   torch.device("cuda", 0))
    Ct, lt, dt = (_dev_tensor(x, dev) for x in (C, len_ft, diam_ft))
    lt, dt = lt.reshape(-1), dt.reshape(-1)
    if Ct.dim() == 0:
        Ct = Ct.expand(lt.numel())
    if Ct.shape[-1] != lt.numel() or dt.numel() != lt.numel():
        raise ValueError("C [..., L] with len_ft, diam_ft [L]")
    return autodiff.hw_resistance(Ct, lt, dt)


def dr_dC(solver, r_hw=None):
    """Chain factor dr/dC = -Hexp r / C ([L] numpy, 0 outside H-W pipes with
    C > 0), dgga/calib.py:89-108."""
    s = solver
    if getattr(s, "headloss_form", "H-W") != "H-W":
        raise NotImplementedError("wrt='C' needs Hazen-Williams headloss; use wrt='r'")
    r = np.asarray(s.r_hw.detach().cpu().numpy() if r_hw is None else r_hw, dtype=np.float64)
    C = np.asarray(s.net.roughness, dtype=np.float64)
    lt = np.asarray(s.net.link_type)
    on = np.isin(lt, _PIPE_TYPES) & (C > 0.0)
    out = np.zeros(s.L)
    out[on] = -s.hexp * r[on] / C[on]
    return out


def _net_of(obj):
    return getattr(obj, "net", obj)


class RoughnessParam(torch.nn.Module):
    """Unconstrained parametrisation of the calibrated pipes' C
    (dgga/calib.py:116-258): C(theta) = C_lo + (C_hi - C_lo) sigmoid(theta),
    strictly inside (C_lo, C_hi) with dC/dtheta > 0 for every finite theta,
    so a zero gradient can only come from the hydraulics.

    net: a Net or a GGASolver (whose r_hw is then the base); pipe_idx: the
    calibrated links (None: every H-W pipe including CV pipes); init: the
    initial C (None: the network's roughness); r_base: [L] resistances kept
    at the other links.  ``r_hw()`` is the full [L] resistance, the free
    pipes recomputed by the HIP map and every other link passed through
    bit for bit, for ``implicit_solve(..., r_hw=...)``."""

    def __init__(self, net, pipe_idx=None, C_lo=40.0, C_hi=160.0, init=None, r_base=None,
                 device=None):
        super().__init__()
        n = _net_of(net)
        L = int(n.L)
        lt = np.asarray(n.link_type, dtype=np.int64)
        rough = np.asarray(n.roughness, dtype=np.float64)
        if str(n.meta.get("headloss", "H-W")).upper() != "H-W":
            raise NotImplementedError("RoughnessParam parametrises Hazen-Williams C only")
        if not (C_hi > C_lo > 0.0):
            raise ValueError("box (C_lo=%r, C_hi=%r) is not 0 < C_lo < C_hi" % (C_lo, C_hi))
        is_pipe = np.isin(lt, _PIPE_TYPES) & (rough > 0.0)
        if pipe_idx is None:
            idx = np.nonzero(is_pipe)[0]
        else:
            idx = np.asarray(pipe_idx, dtype=np.int64).ravel()
            if idx.size and (idx.min() < 0 or idx.max() >= L):
                raise ValueError("pipe_idx out of range")
            bad = idx[~is_pipe[idx]]
            if bad.size:
                raise ValueError("pipe_idx holds links without a Hazen-Williams C: %s"
                                 % bad.tolist()[:10])
            if np.unique(idx).size != idx.size:
                raise ValueError("pipe_idx has repeated links")
        self.n_free = int(idx.size)
        C0 = rough[idx].copy() if init is None else np.broadcast_to(
            np.asarray(init, dtype=np.float64), (self.n_free,)).copy()
        if np.any(C0 <= C_lo) or np.any(C0 >= C_hi):
            raise ValueError("initial C must lie strictly inside (%g, %g)" % (C_lo, C_hi))
        if device is None:
            device = getattr(net, "device", None) or torch.device("cuda", 0)
        if r_base is None:
            r0 = getattr(net, "r_hw", None)
            r0 = n.r_hw if r0 is None else r0
        else:
            r0 = r_base
        r0 = np.asarray(r0.detach().cpu().numpy() if isinstance(r0, torch.Tensor) else r0,
                        dtype=np.float64).ravel()
        if r0.shape != (L,):
            raise ValueError("r_base must be [%d]" % L)
        self.L = L
        self.C_lo, self.C_hi = float(C_lo), float(C_hi)
        self.theta = torch.nn.Parameter(torch.as_tensor(self.theta_of_C(C0, C_lo, C_hi),
                                                        dtype=torch.float64, device=device))
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=device)
        self.register_buffer("idx", torch.as_tensor(idx, dtype=torch.int64, device=device))
        self.register_buffer("len_free", f64(np.asarray(n.len_ft)[idx]))
        self.register_buffer("diam_free", f64(np.asarray(n.diam_ft)[idx]))
        self.register_buffer("r_base", f64(r0))

    @staticmethod
    def theta_of_C(C, C_lo=40.0, C_hi=160.0):
        """theta = logit((C - C_lo) / (C_hi - C_lo)), the exact inverse."""
        u = (np.asarray(C, dtype=np.float64) - C_lo) / (C_hi - C_lo)
        if np.any(u <= 0.0) or np.any(u >= 1.0):
            raise ValueError("C must lie strictly inside (C_lo, C_hi)")
        return np.log(u / (1.0 - u))

    def C(self):
        """[n_free] roughness inside (C_lo, C_hi), differentiable in theta."""
        return self.C_lo + (self.C_hi - self.C_lo) * torch.sigmoid(self.theta)

    def r_hw(self, C=None):
        """[L] resistances: the free pipes from the HIP map, the rest r_base."""
        Cv = self.C() if C is None else C
        if self.n_free == 0:
            return self.r_base.clone()
        r_free = autodiff.hw_resistance(Cv, self.len_free, self.diam_free)
        return self.r_base.index_copy(0, self.idx, r_free)

    def C_full(self):
        """[L] numpy: the current C at the free pipes, 0 elsewhere."""
        out = np.zeros(self.L)
        out[self.idx.cpu().numpy()] = self.C().detach().cpu().numpy()
        return out

    def extra_repr(self):
        return "n_free=%d, L=%d, box=(%g, %g)" % (self.n_free, self.L, self.C_lo, self.C_hi)


def _links(s):
    n = s.net
    return (np.asarray(n.link_n1, dtype=np.int64), np.asarray(n.link_n2, dtype=np.int64),
            np.asarray(n.link_type, dtype=np.int64), np.asarray(n.init_status) == 0)


def clamped_mask(solver, sol, frames=None, margin=10.0):
    """Pipes on the RQtol clamp (dgga/calib.py:264-300): a pipe, not
    closed, whose friction gradient Hexp r |q|^(Hexp-1) does not exceed
    margin * RQtol in a frame (margin 1: the solver's own branch).  Over
    several frames a pipe counts as clamped only if it is clamped in every
    frame.  Returns dict(mask [L], per_frame [T,L], n_per_frame [T], frames,
    margin)."""
    s = solver
    r = np.asarray(sol["r_hw"], dtype=np.float64)
    q_all = np.atleast_2d(np.asarray(sol["q"], dtype=np.float64))
    fr = list(range(q_all.shape[0])) if frames is None else [int(b) for b in frames]
    _, _, lt, closed = _links(s)
    base = np.isin(lt, _PIPE_TYPES) & ~closed
    thr = margin * s.rqtol
    per = np.zeros((len(fr), s.L), dtype=bool)
    for i, b in enumerate(fr):
        rb = r if r.ndim == 1 else r[b]
        hg = s.hexp * rb * np.abs(q_all[b]) ** (s.hexp - 1.0)
        per[i] = base & ~(hg > thr)
    return dict(mask=per.all(axis=0) if fr else base.copy(), per_frame=per,
                n_per_frame=per.sum(axis=1), frames=fr, margin=float(margin))


def dead_branch_mask(solver, demand=None, ke=None, dem_atol=0.0):
    """Dead branches (dgga/calib.py:303-340): open links of leaf chains with
    no source, demand or emitter, found by pruning degree-1 nodes that are
    not fixed-head, have no demand above dem_atol and no emitter; their flow
    is 0 whatever r is.  Returns [L] bool."""
    s = solver
    n1, n2, _, closed = _links(s)
    keep = np.zeros(s.N, dtype=bool)
    keep[s.fixed_nodes] = True
    if demand is not None:
        d = demand.detach().cpu().numpy() if isinstance(demand, torch.Tensor) else demand
        keep |= np.abs(np.atleast_2d(np.asarray(d, dtype=np.float64))).max(axis=0) > dem_atol
    if ke is not None:
        k = ke.detach().cpu().numpy() if isinstance(ke, torch.Tensor) else ke
        keep |= np.abs(np.atleast_2d(np.asarray(k, dtype=np.float64))).max(axis=0) > 0.0
    dead = np.zeros(s.L, dtype=bool)
    alive = ~closed
    while True:
        deg = np.bincount(n1[alive], minlength=s.N) + np.bincount(n2[alive], minlength=s.N)
        leaf = (deg == 1) & ~keep
        cut = alive & (leaf[n1] | leaf[n2])
        if not cut.any():
            break
        dead |= cut
        alive &= ~cut
    return dead


def identifiability(solver, demand, res_head, sensor_nodes, r_hw=None, ke=None, weights=None,
                    wrt="C", frames=None, atol=0.0, clamp_margin=10.0, accuracy=1e-12,
                    max_iter=200, polish_steps=3, return_S=False, **solve_kw):
    """Which pipes the sensors cannot calibrate, and why
    (dgga/calib.py:343-437).  One HIP forward and one multi-right-hand-side
    HIP adjoint give S = dp_sensor/dtheta ([T m, L]); the gradient of the
    weighted sum w^T S and S's zero columns are then attributed, in this
    order, to non_pipe, closed, dead_branch, clamped and unobservable; the
    rest are informative.  Returns the reference's dict."""
    from .sensitivity import sensitivity_matrix

    if polish_steps < 3:
        raise ValueError("polish_steps must be at least 3")
    s = solver
    S, info = sensitivity_matrix(s, demand, res_head, r_hw, sensor_nodes, wrt=wrt,
                                 frames=frames, ke=ke, accuracy=accuracy, max_iter=max_iter,
                                 polish_steps=polish_steps, return_info=True, **solve_kw)
    m = int(np.asarray(info["sensors"]).size)
    T = S.shape[0] // m
    if weights is None:
        w = np.ones(S.shape[0])
    else:
        w = np.asarray(weights, dtype=np.float64).ravel()
        if w.size == 1:
            w = np.full(S.shape[0], float(w[0]))
        elif w.size == m and S.shape[0] != m:
            w = np.tile(w, T)
        if w.shape != (S.shape[0],):
            raise ValueError("weights of length %d do not match the %d rows of S"
                             % (w.size, S.shape[0]))
    grad = w @ S
    sol = info["solution"]
    cm = clamped_mask(s, sol, frames=info["frames"], margin=clamp_margin)
    cm1 = clamped_mask(s, sol, frames=info["frames"], margin=1.0)
    dead = dead_branch_mask(s, demand=demand, ke=ke)
    _, _, lt, closed = _links(s)
    pipe = np.isin(lt, _PIPE_TYPES)
    zero_col = ~(np.max(np.abs(S), axis=0) > atol)
    zero_grad = ~(np.abs(grad) > atol)
    reason = np.empty(s.L, dtype=object)
    reason[:] = "informative"
    reason[~pipe] = "non_pipe"
    reason[pipe & closed] = "closed"
    reason[pipe & ~closed & dead] = "dead_branch"
    rest = pipe & ~closed & ~dead
    reason[rest & cm["mask"]] = "clamped"
    reason[rest & ~cm["mask"] & zero_col] = "unobservable"
    counts = {k: int(np.sum(reason == k)) for k in REASONS}
    out = dict(clamped_mask=cm["mask"], clamped_strict=cm1["mask"],
               clamped_per_frame=cm["per_frame"], n_clamped_per_frame=cm["n_per_frame"],
               dead_mask=dead, closed_mask=closed, pipe_mask=pipe, zero_grad_mask=zero_grad,
               zero_col_mask=zero_col, grad=grad, reason=reason, counts=counts,
               n_informative=counts["informative"], n_pipe=int(pipe.sum()),
               sensors=info["sensors"], frames=info["frames"], wrt=wrt, info=info)
    if return_S:
        out["S"] = S
    return out
