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


def _prv_topology(s):
    """(several PRVs share a downstream junction, a PRV feeds another PRV's
    upstream node) for the solver's network, computed once per solver."""
    t = getattr(s, "_adj_prv_topology", None)
    if t is None:
        j2 = np.asarray(s.prv_n2)
        j1 = np.asarray(s.prv_n1)
        t = (bool(np.unique(j2).size < j2.size), bool(np.intersect1d(j2, j1).size))
        s._adj_prv_topology = t
    return t


def _check_forward(info, iters, maxtrials, out, accuracy):
    """The forward's errors (dgga/autodiff.py:1206-1213): a CONST_HP pump on
    its clamp branch, a failed solve or polish, an unconverged scenario."""
    if (info == -300).any():
        bad = np.nonzero(info == -300)[0].tolist()
        raise NotImplementedError(
            "adjoint='gpu': samples %s have a CONST_HP pump on its CBIG/RQtol clamp "
            "branch at the converged state (the reduced system is not the "
            "Jacobian's Schur complement there, dgga/autodiff.py:808-823)" % bad[:8])
    if (info != 0).any():
        bad = np.nonzero(info != 0)[0].tolist()
        raise RuntimeError("adjoint='gpu': samples %s failed in the solve or the polish "
                           "(info %s)" % (bad[:8], info[bad[:8]].tolist()))
    if (iters > maxtrials).any():
        bad = np.nonzero(iters > maxtrials)[0].tolist()
        raise RuntimeError(
            "adjoint='gpu': scenarios %s did not converge (relerr %s > accuracy %g); "
            "the adjoint is undefined there" % (bad[:8], out["relerr"][bad[:4]].tolist(),
                                               float(accuracy)))


def _check_step(binfo, fwd, out, B):
    """The forward's deferred check (if any), then the node adjoint's."""
    if fwd is not None:
        buf, maxtrials, acc = fwd
        _check_forward(buf[:B].numpy(), buf[B:].numpy(), maxtrials, out, acc)
    if (binfo.numpy() != 0).any():
        bad = np.nonzero(binfo.numpy() != 0)[0].tolist()
        raise RuntimeError("adjoint='gpu': the node adjoint failed for samples %s" % bad[:8])


def _forward(ctx, demand, res_head, ke, r_hw, s, accuracy, max_iter, sm, polish_steps,
             defer_check, theta, status0, single_dims):
    """The implicit solve's forward: the fused HIP solve (status machine,
    polish, terminal factor) with the pump parameters theta (or the
    network's), its status checked now or at the backward's synchronisation.
    single_dims: the input ranks that de-batch the outputs when all are 1."""
    if s.dtype != torch.float64:
        raise NotImplementedError("the implicit adjoint is float64 only")
    if s.n_pcv:
        shared, cascaded = _prv_topology(s)
        if shared:
            raise NotImplementedError(
                "implicit adjoint: several PRVs share a downstream junction")
        if cascaded:
            raise NotImplementedError("implicit adjoint: cascaded PRVs")
    d = demand.detach()
    rh = res_head.detach()
    ket = ke.detach()
    rt = r_hw.detach()
    if d.dim() == 1:
        d = d.unsqueeze(0)
    if rh.dim() == 1:
        rh = rh.unsqueeze(0)
    B = max(d.shape[0], rh.shape[0], ket.shape[0] if ket.dim() == 2 else 1,
            rt.shape[0] if rt.dim() == 2 else 1)
    if d.shape[0] not in (1, B):
        raise ValueError("demand has %d scenarios, another input %d" % (d.shape[0], B))
    d = d.expand(B, -1)
    if s._pending:
        s.poll()
    s.set_pumps(theta)
    try:
        a, out, keep, single, B = s._args(d, rh, ket, None, None, max_iter, accuracy, sm,
                                          r_hw=rt, polish_steps=int(polish_steps),
                                          want_factor=True, status0=status0)
        with torch.cuda.device(s.device):
            _lib.check(s._timed("k_solve", lambda: s._lib.wds_solve(s._plan, s._stream(),
                                                                   C.byref(a)),
                                B, out["iters"]), "wds_solve")
    finally:
        s.set_pumps(None)
    maxtrials = s.maxtrials(max_iter, sm)
    ctx.fwd_check = None
    if defer_check:
        # a backward will follow: the solve's status goes to pinned host
        # memory asynchronously and is checked at the backward's one
        # synchronisation, so the loss and the backward queue behind the
        # forward without a host round trip
        buf = torch.empty(2 * B, dtype=torch.int32, pin_memory=True)
        buf.copy_(out["st"], non_blocking=True)
        ctx.fwd_check = (buf, maxtrials, float(accuracy))
    else:
        info, iters = s._host_status(out, B)
        _check_forward(info, iters, maxtrials, out, accuracy)
    # outputs that do not reach the loss arrive in backward as None and are
    # passed to the kernel as null pointers (zero cotangents)
    ctx.set_materialize_grads(False)
    ctx.solver = s
    ctx.keep = keep
    ctx.out = out
    ctx.theta = theta
    ctx.dims = (demand.dim(), res_head.dim(), ke.dim(), r_hw.dim())
    ctx.in_devs = (demand.device, res_head.device, ke.device, r_hw.device)
    ctx.B = B
    s._last_resid = out["resid"]
    head, flow, emit = out["head"], out["flow"], out["emitter"]
    if all(x == 1 for x in single_dims):
        head, flow, emit = head[0], flow[0], emit[0]
    return head, flow, emit


def _backward(ctx, gH, gQ, gE, want_pump):
    """One wds_adjoint launch on the forward's terminal factor; returns the
    gradients (demand, res_head, ke, r_hw) fitted to the inputs' shapes and,
    with want_pump, d loss / d (speed, pump_h0, pump_r) [L] summed over the
    batch (batch-shared parameters, dgga/autodiff.py:713-718)."""
    if torch.is_grad_enabled():
        raise NotImplementedError(
            "the implicit adjoint is first order only (create_graph=True is not supported)")
    s = ctx.solver
    out = ctx.out
    B = ctx.B

    def _g(g, ref):
        if g is None:
            return None
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
    gp = torch.empty(B, 3, s.L, dtype=s.dtype, device=dev) if want_pump else None
    info = torch.empty(B, dtype=torch.int32, device=dev)
    ws = s._workspace(B)
    a = _lib.AdjointArgs(
        B=B, demand=_lib.ptr(dkeep), ke=_lib.ptr(ke),
        ke_stride=s.N if ke.dim() == 2 else 0, r_hw=_lib.ptr(r),
        r_stride=s.L if r.dim() == 2 else 0, head=_lib.ptr(out["head"]),
        flow=_lib.ptr(out["flow"]), emitter=_lib.ptr(out["emitter"]),
        status=_lib.ptr(out["status"]), factor=_lib.ptr(out["factor"]),
        gH=_lib.ptr(gHt), gQ=_lib.ptr(gQt), gE=_lib.ptr(gEt),
        n_refine=int(s.cudss_grad_refine), g_demand=_lib.ptr(gd), g_res_head=_lib.ptr(grh),
        g_ke=_lib.ptr(gke), g_r=_lib.ptr(gr), info=_lib.ptr(info), workspace=_lib.ptr(ws),
        g_pump=_lib.ptr(gp))
    # the adjoint's link coefficients use the forward's pump parameters
    s.set_pumps(ctx.theta)
    try:
        with torch.cuda.device(dev):
            _lib.check(s._timed("k_adjoint", lambda: s._lib.wds_adjoint(s._plan, s._stream(),
                                                                       C.byref(a)), B),
                       "wds_adjoint")
    finally:
        s.set_pumps(None)
    binfo = torch.empty(B, dtype=torch.int32, pin_memory=True) if s.lazy_checks \
        else s._pinned_info(B)
    binfo.copy_(info, non_blocking=True)
    dm, dv = ctx.dims, ctx.in_devs
    # the batch reductions (inputs shared by the batch) are queued behind the
    # adjoint, in one launch, before the host waits for its status
    per = [gd, grh, gke, gr]
    sums = [i for i in range(4) if dm[i] == 1]
    outs = {}
    src = [per[i] for i in sums] + ([gp.reshape(B, -1)] if want_pump else [])
    if src:
        dst = [torch.empty(t.shape[1:], dtype=s.dtype, device=dev) for t in src]
        arr = C.c_void_p * len(src)
        n = (C.c_int64 * len(src))(*[t.shape[1] for t in src])
        with torch.cuda.device(dev):
            _lib.check(s._lib.wds_batch_sums(s._stream(), B, len(src),
                                              arr(*[_lib.ptr(t) for t in src]), n,
                                              arr(*[_lib.ptr(t) for t in dst])),
                       "wds_batch_sums")
        outs = dict(zip(sums, dst))
    grads = tuple((outs[i] if i in outs else per[i]).to(device=dv[i]) for i in range(4))
    pump = tuple(dst[-1].view(3, s.L).unbind(0)) if want_pump else (None, None, None)
    fwd = ctx.fwd_check
    ctx.fwd_check = None
    if s.lazy_checks:
        # checked at the next call into the solver or at s.check()
        s._defer(_check_step, binfo, fwd, out, B)
    else:
        torch.cuda.current_stream(dev).synchronize()
        _check_step(binfo, fwd, out, B)
    return grads, pump


class ImplicitGGASolveGPU(torch.autograd.Function):
    """(demand, res_head, ke, r_hw, solver, accuracy, max_iter,
    status_machine, assemble, linear_solver, polish_steps) ->
    (head [B,N], flow [B,L], emitter [B,N])."""

    @staticmethod
    def forward(ctx, demand, res_head, ke, r_hw, solver, accuracy=1e-12, max_iter=200,
                status_machine=None, assemble="dense", linear_solver="dense",
                polish_steps=2, defer_check=False):
        s = solver
        sm = bool(s.dense_status_machine) if status_machine is None else bool(status_machine)
        s._validate(assemble, linear_solver if linear_solver != "dense" else "hip", sm,
                    None, None)
        defer = bool(defer_check) and any(ctx.needs_input_grad[:4])
        dims = (demand.dim(), res_head.dim(), ke.dim(), r_hw.dim())
        return _forward(ctx, demand, res_head, ke, r_hw, s, accuracy, max_iter, sm,
                        polish_steps, defer, None, None, dims)

    @staticmethod
    def backward(ctx, gH, gQ, gE):
        grads, _ = _backward(ctx, gH, gQ, gE, False)
        return grads + (None, None, None, None, None, None, None, None)


class ImplicitGGASolve(torch.autograd.Function):
    """The reference's CPU-adjoint route (dgga/autodiff.py:609-732, the
    default adjoint='cpu' of implicit_solve) on the same HIP kernels:
    (demand, res_head, ke, r_hw, solver, accuracy, max_iter, polish_steps,
    speed [L], pump_h0 [L], pump_r [L], status [L]) -> (head, flow,
    emitter).  speed / pump_h0 / pump_r are batch-shared pump parameters
    (read at the pump links), differentiable when passed as tensors;
    status (internal codes) is the solve's initial link status.  The
    statuses at which the adjoint is taken are those the status machine
    converges to (the reference freezes the given, or initial, statuses
    there; the two agree when status is the converged one, the reference's
    recommended use)."""

    @staticmethod
    def forward(ctx, demand, res_head, ke, r_hw, solver, accuracy=1e-12, max_iter=200,
                polish_steps=3, speed=None, pump_h0=None, pump_r=None, status=None,
                defer_check=False):
        s = solver
        sm = bool(s.dense_status_machine)
        s._validate("dense", "hip", sm, status, None)
        pumps = (speed, pump_h0, pump_r)
        theta = None
        if any(x is not None for x in pumps):
            if not s.n_pumps:
                raise ValueError("speed / pump_h0 / pump_r given for a network without pumps")
            theta = s.pump_theta(*pumps)
        ctx.pump_in = tuple(isinstance(x, torch.Tensor) for x in pumps)
        ctx.pump_devs = tuple(x.device if isinstance(x, torch.Tensor) else None for x in pumps)
        defer = bool(defer_check) and (any(ctx.needs_input_grad[:4]) or
                                       any(ctx.needs_input_grad[8:11]))
        dims = (demand.dim(), res_head.dim(), ke.dim())
        return _forward(ctx, demand, res_head, ke, r_hw, s, accuracy, max_iter, sm,
                        polish_steps, defer, theta, status, dims)

    @staticmethod
    def backward(ctx, gH, gQ, gE):
        want = any(ctx.pump_in) and any(ctx.needs_input_grad[8:11])
        grads, pump = _backward(ctx, gH, gQ, gE, want)
        pg = tuple(g.to(device=d) if (g is not None and inp) else None
                   for g, inp, d in zip(pump, ctx.pump_in, ctx.pump_devs))
        return grads + (None, None, None, None) + pg + (None, None)


def implicit_solve(solver, demand, res_head, ke=None, r_hw=None, accuracy=1e-12,
                   max_iter=200, polish_steps=3, speed=None, pump_h0=None, pump_r=None,
                   status=None, adjoint="cpu", status_machine=None, assemble="dense",
                   linear_solver="dense"):
    """dgga/autodiff.py:1315-1351 with the reference's defaults.  Both
    routes run the HIP forward and the HIP implicit adjoint:
    adjoint="cpu" (ImplicitGGASolve) takes the pump parameters speed /
    pump_h0 / pump_r ([L], differentiable) and status; adjoint="gpu"
    (ImplicitGGASolveGPU) takes status_machine / assemble / linear_solver
    and refuses the pump arguments, as the reference's GPU route does."""
    dt, dev = solver.dtype, solver.device
    if ke is None:
        ke = solver.node_ke_default
    if r_hw is None:
        r_hw = solver.r_hw
    t = lambda x: x.to(dtype=dt, device=dev) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
    args = (t(demand), t(res_head), t(ke), t(r_hw))
    if adjoint == "cpu":
        tn = lambda x: None if x is None else t(x)
        pumps = (tn(speed), tn(pump_h0), tn(pump_r))
        defer = torch.is_grad_enabled() and any(
            x is not None and x.requires_grad for x in args + pumps)
        return ImplicitGGASolve.apply(*args, solver, accuracy, max_iter, int(polish_steps),
                                      *pumps, status, defer)
    if adjoint != "gpu":
        raise ValueError("adjoint must be 'cpu' or 'gpu', got %r" % (adjoint,))
    if speed is not None or pump_h0 is not None or pump_r is not None or status is not None:
        raise NotImplementedError(
            "adjoint='gpu' does not take speed/pump_h0/pump_r/status (the reference's GPU "
            "route refuses them too); use adjoint='cpu' for pump-parameter gradients")
    # with a backward to come, the forward's status check waits for the
    # backward's synchronisation
    defer = torch.is_grad_enabled() and any(x.requires_grad for x in args)
    return ImplicitGGASolveGPU.apply(*args, solver, accuracy, max_iter, status_machine,
                                     assemble, linear_solver, int(polish_steps), defer)


class _State:
    """What a forward keeps for its adjoints, outside autograd."""

    def set_materialize_grads(self, value):
        pass


def forward_state(solver, demand, res_head, ke=None, r_hw=None, accuracy=1e-12,
                  max_iter=200, polish_steps=3, speed=None, pump_h0=None, pump_r=None,
                  status=None):
    """The implicit solve's forward outside autograd (the forward of
    ImplicitGGASolve, dgga/autodiff.py:539-603 + 1171-1243): the fused HIP
    solve with the status machine, the polish and the terminal factor, its
    status checked now.  Returns the state ``adjoint_rows`` takes."""
    s = solver
    dt, dev = s.dtype, s.device
    t = lambda x: x.detach().to(dtype=dt, device=dev) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
    ke = s.node_ke_default if ke is None else ke
    r_hw = s.r_hw if r_hw is None else r_hw
    d, rh, ket, rt = t(demand), t(res_head), t(ke), t(r_hw)
    sm = bool(s.dense_status_machine)
    s._validate("dense", "hip", sm, status, None)
    theta = None
    if speed is not None or pump_h0 is not None or pump_r is not None:
        if not s.n_pumps:
            raise ValueError("speed / pump_h0 / pump_r given for a network without pumps")
        theta = s.pump_theta(speed, pump_h0, pump_r)
    st = _State()
    with torch.no_grad():
        _forward(st, d, rh, ket, rt, s, accuracy, max_iter, sm, polish_steps, False, theta,
                 status, (2,))
    return st


def adjoint_rows(st, rows, gH=None, gQ=None, gE=None, want_pump=False):
    """One wds_adjoint launch over the forward scenarios ``rows`` (1-D
    indices into the forward's batch, repeats allowed) with cotangents
    gH/gE [R,N], gQ [R,L] (None: zero), each row reusing its scenario's
    terminal factor (no new factorisation; the reference solves one
    factorised Jacobian for many right-hand sides the same way,
    dgga/sensitivity.py:59-160).  Returns per-row (g_demand [R,N],
    g_res_head [R,N], g_ke [R,N], g_r [R,L]) and, with want_pump,
    g_pump [R,3,L]; the status of the launch is checked before returning."""
    s = st.solver
    dev = s.device
    rows = torch.as_tensor(rows, dtype=torch.int64, device=dev).reshape(-1)
    R = int(rows.numel())
    if R < 1 or int(rows.min()) < 0 or int(rows.max()) >= st.B:
        raise ValueError("rows must index the forward's %d scenarios" % st.B)
    out, keep = st.out, st.keep
    pick = lambda x: x.index_select(0, rows).contiguous()
    dk, ke, r = keep[0], keep[2], keep[3]
    ke = pick(ke) if ke.dim() == 2 and ke.shape[0] == st.B and st.B > 1 else ke
    r = pick(r) if r.dim() == 2 and r.shape[0] == st.B and st.B > 1 else r
    dk = pick(dk.expand(st.B, -1))
    head, flow, emit = pick(out["head"]), pick(out["flow"]), pick(out["emitter"])
    status, factor = pick(out["status"]), pick(out["factor"])

    def _g(g, n):
        if g is None:
            return None
        g = g.detach().to(dtype=s.dtype, device=dev).contiguous()
        if tuple(g.shape) != (R, n):
            raise ValueError("cotangent must be [%d,%d]" % (R, n))
        return g

    gHt, gQt, gEt = _g(gH, s.N), _g(gQ, s.L), _g(gE, s.N)
    gd = torch.empty(R, s.N, dtype=s.dtype, device=dev)
    grh = torch.empty_like(gd)
    gke = torch.empty_like(gd)
    gr = torch.empty(R, s.L, dtype=s.dtype, device=dev)
    gp = torch.empty(R, 3, s.L, dtype=s.dtype, device=dev) if want_pump else None
    info = torch.empty(R, dtype=torch.int32, device=dev)
    ws = s._workspace(R)
    a = _lib.AdjointArgs(
        B=R, demand=_lib.ptr(dk), ke=_lib.ptr(ke), ke_stride=s.N if ke.dim() == 2 else 0,
        r_hw=_lib.ptr(r), r_stride=s.L if r.dim() == 2 else 0, head=_lib.ptr(head),
        flow=_lib.ptr(flow), emitter=_lib.ptr(emit), status=_lib.ptr(status),
        factor=_lib.ptr(factor), gH=_lib.ptr(gHt), gQ=_lib.ptr(gQt), gE=_lib.ptr(gEt),
        n_refine=int(s.cudss_grad_refine), g_demand=_lib.ptr(gd), g_res_head=_lib.ptr(grh),
        g_ke=_lib.ptr(gke), g_r=_lib.ptr(gr), info=_lib.ptr(info), workspace=_lib.ptr(ws),
        g_pump=_lib.ptr(gp))
    s.set_pumps(st.theta)
    try:
        with torch.cuda.device(dev):
            _lib.check(s._timed("k_adjoint", lambda: s._lib.wds_adjoint(s._plan, s._stream(),
                                                                       C.byref(a)), R),
                       "wds_adjoint")
    finally:
        s.set_pumps(None)
    bad = torch.nonzero(info != 0).flatten().tolist()
    if bad:
        raise RuntimeError("adjoint: the node adjoint failed for rows %s" % bad[:8])
    res = (gd, grh, gke, gr)
    return res + (gp,) if want_pump else res


def solve_polished(solver, d, rh, ke=None, r_hw=None, accuracy=1e-12, max_iter=200,
                   polish_steps=3, speed=None, status=None, pump_h0=None, pump_r=None,
                   assemble="dense", linear_solver="dense"):
    """dgga/autodiff.py:539-603: converged and polished states as numpy
    (q [B,L], e_j [B,Nj], head [B,N], emitter [B,N], d, rh, ke [B,N], r_hw
    [L], resid_inf [B], speed / status / h0p / rp [L] as resolved, iters
    [B]), computed by the HIP forward (status machine, reduced Newton
    polish).  ``status`` holds the links' initial internal statuses."""
    s = solver
    s._validate(assemble, linear_solver if linear_solver != "dense" else "hip",
                bool(s.dense_status_machine), status, None)
    st = forward_state(s, d, rh, ke, r_hw, accuracy, max_iter, polish_steps, speed, pump_h0,
                       pump_r, status)
    out, keep = st.out, st.keep
    B = st.B
    np_ = lambda x: x.detach().cpu().numpy()
    head = np_(out["head"])
    em = np_(out["emitter"])
    kk = np.broadcast_to(np.atleast_2d(np_(keep[2])), (B, s.N)).copy()
    th = s.pump_theta(*(st.theta if st.theta is not None else (None, None, None)))
    return dict(q=np_(out["flow"]), e_j=em[:, s.junc_nodes].copy(), head=head, emitter=em,
                d=np.broadcast_to(np.atleast_2d(np_(keep[0])), (B, s.N)).copy(),
                rh=np.broadcast_to(np.atleast_2d(np_(keep[1])), (B, s.N)).copy(), ke=kk,
                r_hw=np_(keep[3]).reshape(-1, s.L)[0].copy() if keep[3].dim() == 2
                and keep[3].shape[0] == 1 else np_(keep[3]),
                resid_inf=np_(out["resid"]), speed=th[0], status=np_(out["status"]),
                h0p=th[1], rp=th[2], iters=np_(out["iters"]).astype(np.int64))


def _stream_of(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _hip_tensor(t, what):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
        raise NotImplementedError("%s runs on a HIP device in float64 only" % what)
    return t.contiguous()


class _HWResistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C_hw, len_ft, diam_ft, mask):
        Ct = _hip_tensor(C_hw, "hw_resistance")
        L = int(len_ft.numel())
        lt = _hip_tensor(len_ft, "hw_resistance")
        dt = _hip_tensor(diam_ft, "hw_resistance")
        if Ct.shape[-1] != L or dt.numel() != L:
            raise ValueError("hw_resistance: C [..., L] with len/diam [L]")
        if mask is None:
            mk = None
        elif mask.dtype == torch.bool and mask.device == Ct.device:
            mk = mask.contiguous().view(torch.int8)  # same bytes, 0/1: no conversion kernel
        else:
            mk = mask.to(device=Ct.device, dtype=torch.int8).contiguous()
        r = torch.empty_like(Ct)
        lib = _lib.load()
        with torch.cuda.device(Ct.device):
            _lib.check(lib.wds_hw_resistance(_stream_of(Ct), Ct.numel(), L, _lib.ptr(Ct),
                                             _lib.ptr(lt), _lib.ptr(dt), _lib.ptr(mk),
                                             _lib.ptr(r)), "wds_hw_resistance")
        ctx.save_for_backward(Ct, r)
        return r

    @staticmethod
    def backward(ctx, g_r):
        Ct, r = ctx.saved_tensors
        g = g_r.contiguous()
        gC = torch.empty_like(Ct)
        lib = _lib.load()
        with torch.cuda.device(Ct.device):
            _lib.check(lib.wds_hw_resistance_grad(_stream_of(Ct), Ct.numel(), _lib.ptr(Ct),
                                                  _lib.ptr(r), _lib.ptr(g), _lib.ptr(gC)),
                       "wds_hw_resistance_grad")
        return gC, None, None, None


class _EmitterKe(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c_user, qexp, coef):
        Ct = _hip_tensor(c_user, "emitter_ke")
        ke = torch.empty_like(Ct)
        lib = _lib.load()
        with torch.cuda.device(Ct.device):
            _lib.check(lib.wds_emitter_ke(_stream_of(Ct), Ct.numel(), _lib.ptr(Ct),
                                          float(qexp), float(coef), _lib.ptr(ke)),
                       "wds_emitter_ke")
        ctx.save_for_backward(Ct, ke)
        ctx.qexp = float(qexp)
        return ke

    @staticmethod
    def backward(ctx, g_ke):
        Ct, ke = ctx.saved_tensors
        g = g_ke.contiguous()
        gC = torch.empty_like(Ct)
        lib = _lib.load()
        with torch.cuda.device(Ct.device):
            _lib.check(lib.wds_emitter_ke_grad(_stream_of(Ct), Ct.numel(), _lib.ptr(Ct),
                                               _lib.ptr(ke), ctx.qexp, _lib.ptr(g),
                                               _lib.ptr(gC)), "wds_emitter_ke_grad")
        return gC, None, None


def hw_resistance(C_hw, len_ft, diam_ft, mask=None):
    """Differentiable Hazen-Williams r = 4.727 L / C^1.852 / D^4.871
    (dgga/calib.py:64-86, same operation order; gradient dr/dC = -1.852 r / C,
    calib.py:89-108) as one HIP kernel each way.  C is [L] or [B, L]; mask
    [L] (optional) zeroes r on links that are not pipes."""
    return _HWResistance.apply(C_hw, len_ft, diam_ft, mask)


def emitter_ke(c_user, qexp=2.0, ucf_flow=LPS_PER_CFS, ucf_pressure=M_PER_FT):
    """Differentiable emitter coefficient Ke = (Ucf_Q^Qexp / Ucf_P) / C^Qexp
    (dgga/parse.py:557-568); zero where C == 0.  One HIP kernel each way."""
    return _EmitterKe.apply(c_user, qexp, ucf_flow ** qexp / ucf_pressure)
