# -*- coding: utf-8 -*-
"""Implicit adjoint of the batched GGA solve as a torch.autograd.Function.

Mirrors dgga/autodiff.py ``ImplicitGGASolveGPU`` (1157-1312) and
``implicit_solve`` (1315-1351): the forward runs the fused HIP Newton loop,
the reduced full-Newton polish and a terminal factorisation whose factor the
backward reuses (zero new factorisations); the backward is one HIP launch
that solves the node adjoint system (Woodbury row replacement for ACTIVE
PRVs) and contracts it into gradients with respect to demands, reservoir
heads, emitter coefficients and pipe resistances.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .network import LPS_PER_CFS, M_PER_FT


class ImplicitGGASolveGPU(torch.autograd.Function):
    """(demand, res_head, ke, r_hw, solver, accuracy, max_iter,
    status_machine, assemble, linear_solver, polish_steps) ->
    (head [B,N], flow [B,L], emitter [B,N])."""

    @staticmethod
    def forward(ctx, demand, res_head, ke, r_hw, solver, accuracy=1e-12, max_iter=200,
                status_machine=None, assemble="dense", linear_solver="dense",
                polish_steps=2):
        s = solver
        if s.dtype != torch.float64:
            raise NotImplementedError("adjoint='gpu' is float64 only")
        if s.headloss_form != "H-W" or s.n_pumps:
            raise NotImplementedError(
                "adjoint='gpu' covers Hazen-Williams networks without pumps "
                "(pump theta and Darcy-Weisbach gradients are not on this path)")
        if s.n_pcv:
            j2 = np.asarray(s.prv_n2)
            j1 = np.asarray(s.prv_n1)
            if np.unique(j2).size < j2.size:
                raise NotImplementedError(
                    "adjoint='gpu': several PRVs share a downstream junction")
            if np.intersect1d(j2, j1).size:
                raise NotImplementedError("adjoint='gpu': cascaded PRVs")
        dims = (demand.dim(), res_head.dim(), ke.dim(), r_hw.dim())
        sm = bool(s.dense_status_machine) if status_machine is None else bool(status_machine)
        s._validate(assemble, linear_solver if linear_solver != "dense" else "hip", sm,
                    None, None)
        d = demand.detach()
        rh = res_head.detach()
        ket = ke.detach()
        rt = r_hw.detach()
        if d.dim() == 1:
            d = d.unsqueeze(0)
        if rh.dim() == 1:
            rh = rh.unsqueeze(0)
        B = max(d.shape[0], rh.shape[0], ket.shape[0] if ket.dim() == 2 else 1)
        d = d.expand(B, -1)
        rh = rh.expand(B, -1)
        a, out, keep, single, B = s._args(d, rh, ket, None, None, max_iter, accuracy, sm,
                                          r_hw=rt, polish_steps=int(polish_steps),
                                          want_factor=True)
        with torch.cuda.device(s.device):
            _lib.check(s._timed("k_solve", lambda: s._lib.wds_solve(s._plan, s._stream(),
                                                                   C.byref(a)),
                                B, out["iters"]), "wds_solve")
        info = out["info"]
        if bool((info != 0).any()):
            bad = torch.nonzero(info != 0).flatten().tolist()
            raise RuntimeError("adjoint='gpu': samples %s failed in the solve or the polish "
                               "(info %s)" % (bad[:8], info[bad[:8]].tolist()))
        maxtrials = s.maxtrials(max_iter, sm)
        unconv = out["iters"].to(torch.int64) > maxtrials
        if bool(unconv.any()):
            bad = torch.nonzero(unconv).flatten().tolist()
            raise RuntimeError(
                "adjoint='gpu': scenarios %s did not converge (relerr %s > accuracy %g); "
                "the adjoint is undefined there" % (bad[:8], out["relerr"][bad[:4]].tolist(),
                                                   float(accuracy)))
        ctx.solver = s
        ctx.keep = keep
        ctx.out = out
        ctx.dims = dims
        ctx.in_devs = (demand.device, res_head.device, ke.device, r_hw.device)
        ctx.B = B
        s._last_resid = out["resid"]
        head, flow, emit = out["head"], out["flow"], out["emitter"]
        if all(x == 1 for x in dims[:3]):
            head, flow, emit = head[0], flow[0], emit[0]
        return head, flow, emit

    @staticmethod
    def backward(ctx, gH, gQ, gE):
        if torch.is_grad_enabled():
            raise NotImplementedError(
                "adjoint='gpu' is first order only (create_graph=True is not supported)")
        s = ctx.solver
        out = ctx.out
        B = ctx.B

        def _g(g, ref):
            if g is None:
                return torch.zeros_like(ref)
            g = g.detach().to(dtype=s.dtype, device=s.device)
            return (g.unsqueeze(0).expand_as(ref) if g.dim() == 1 else g).contiguous()

        gHt = _g(gH, out["head"])
        gQt = _g(gQ, out["flow"])
        gEt = _g(gE, out["emitter"])
        dkeep, rh, ke, r = ctx.keep[0], ctx.keep[1], ctx.keep[2], ctx.keep[3]
        dev = s.device
        gd = torch.empty(B, s.N, dtype=s.dtype, device=dev)
        grh = torch.empty_like(gd)
        gke = torch.empty_like(gd)
        gr = torch.empty(B, s.L, dtype=s.dtype, device=dev)
        info = torch.zeros(B, dtype=torch.int32, device=dev)
        ws = s._workspace(B)
        a = _lib.AdjointArgs(
            B=B, demand=_lib.ptr(dkeep), ke=_lib.ptr(ke),
            ke_stride=s.N if ke.dim() == 2 else 0, r_hw=_lib.ptr(r),
            r_stride=s.L if r.dim() == 2 else 0, head=_lib.ptr(out["head"]),
            flow=_lib.ptr(out["flow"]), emitter=_lib.ptr(out["emitter"]),
            status=_lib.ptr(out["status"]), factor=_lib.ptr(out["factor"]),
            gH=_lib.ptr(gHt), gQ=_lib.ptr(gQt), gE=_lib.ptr(gEt),
            n_refine=int(s.dense_refine), g_demand=_lib.ptr(gd), g_res_head=_lib.ptr(grh),
            g_ke=_lib.ptr(gke), g_r=_lib.ptr(gr), info=_lib.ptr(info), workspace=_lib.ptr(ws))
        with torch.cuda.device(dev):
            _lib.check(s._timed("k_adjoint", lambda: s._lib.wds_adjoint(s._plan, s._stream(),
                                                                       C.byref(a)), B),
                       "wds_adjoint")
        if bool((info != 0).any()):
            bad = torch.nonzero(info != 0).flatten().tolist()
            raise RuntimeError("adjoint='gpu': the node adjoint failed for samples %s" % bad[:8])
        dm, dv = ctx.dims, ctx.in_devs

        def _fit(g, dim, dev_):
            g = g.to(device=dev_)
            return g.sum(0) if dim == 1 else g

        return (_fit(gd, dm[0], dv[0]), _fit(grh, dm[1], dv[1]), _fit(gke, dm[2], dv[2]),
                _fit(gr, dm[3], dv[3]), None, None, None, None, None, None, None)


def implicit_solve(solver, demand, res_head, ke=None, r_hw=None, accuracy=1e-12,
                   max_iter=200, polish_steps=3, speed=None, pump_h0=None, pump_r=None,
                   status=None, adjoint="gpu", status_machine=None, assemble="dense",
                   linear_solver="dense"):
    """dgga/autodiff.py:1315-1351.  adjoint must be 'gpu' (the scalar
    SuperLU reference adjoint is not part of the MI355X path); pump theta
    arguments are refused as on the reference's GPU adjoint."""
    if adjoint == "cpu":
        raise NotImplementedError(
            "adjoint='cpu' (scipy SuperLU per scenario) is the reference's CPU path; "
            "this package runs adjoint='gpu'")
    if adjoint != "gpu":
        raise ValueError("adjoint must be 'cpu' or 'gpu', got %r" % (adjoint,))
    if speed is not None or pump_h0 is not None or pump_r is not None or status is not None:
        raise NotImplementedError("adjoint='gpu' does not take speed/pump_h0/pump_r/status")
    dt, dev = solver.dtype, solver.device
    if ke is None:
        ke = solver.node_ke_default
    if r_hw is None:
        r_hw = solver.r_hw
    t = lambda x: x.to(dtype=dt, device=dev) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
    return ImplicitGGASolveGPU.apply(t(demand), t(res_head), t(ke), t(r_hw), solver,
                                     accuracy, max_iter, status_machine, assemble,
                                     linear_solver, int(polish_steps))


def hw_resistance(C_hw, len_ft, diam_ft):
    """Differentiable Hazen-Williams r = 4.727 L / C^1.852 / D^4.871
    (dgga/calib.py:64-86, same operation order)."""
    return 4.727 * len_ft / C_hw ** 1.852 / diam_ft ** 4.871


def emitter_ke(c_user, qexp=2.0, ucf_flow=LPS_PER_CFS, ucf_pressure=M_PER_FT):
    """Differentiable emitter coefficient Ke = (Ucf_Q^Qexp / Ucf_P) / C^Qexp
    (dgga/parse.py:557-568); zero where C == 0."""
    pos = c_user > 0
    safe = torch.where(pos, c_user, torch.ones_like(c_user))
    return torch.where(pos, (ucf_flow ** qexp / ucf_pressure) / safe ** qexp,
                       torch.zeros_like(c_user))
