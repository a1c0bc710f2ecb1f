# -*- coding: utf-8 -*-
"""Differentiable batched extended-period simulation: a fixed-step EPS whose
every step is the fused GGA solve on the GPU, and its discrete-time adjoint
as a torch.autograd.Function.

The reference's EPS (dgga/eps.py ``EpsDriver`` 40-414) is forward only and
event-driven; this model keeps its per-step semantics on a fixed step grid
of ``dt`` seconds, for B scenarios at once.  Step t:

1. simple controls on the statuses carried from step t-1
   (dgga/eps.py:171-218): LOWLEVEL / HILEVEL on tank volumes with the
   tolerance vplus = |net tank inflow of step t-1|, TIMER and TIMEOFDAY at
   the step's time; a constant-horsepower pump that a control reopens
   restarts from its design flow;
2. demands d_t = demand * f_t, f_t the pattern factor of period
   (t dt + Pstart) // Pstep (dgga/eps.py:125-147);
3. fixed heads: reservoirs from ``res_head``, tanks at h_t (h_0 =
   ``tank_h0``);
4. the fused solve (Newton loop with the status machine, polish, terminal
   factor) warm-started from step t-1's flows, emitter flows and statuses;
5. tank update (dgga/eps.py:278-291): net inflow qnet_j (+Q over links
   with n2 = tank, then -Q over links with n1 = tank, in link order),
   V += qnet dt, V = Vmax if V + qnet >= Vmax, else Vmin if V - qnet <= Vmin,
   h = hmin + (V - Vmin) / A; V_0 = A (h_0 - El).

A control therefore acts at the first step boundary after its threshold is
crossed (the event-driven driver inserts a sub-step there instead).

The backward sweeps the steps in reverse with statuses, control decisions
and tank clamps frozen.  With nu_t = dLoss/dV_t: the flow cotangent of step
t is sign dt nu_{t+1} on the links into each unclamped tank, the step's
implicit adjoint (wds_adjoint, reusing the step's terminal factor) gives
the gradients of the step's inputs, and nu_t = g_head_t[tank] / A +
nu_{t+1} (unclamped); dLoss/dh_0 = g_head_0[tank] + A nu_1.  Parameter
gradients are summed over the steps.

Memory: with ``checkpoint_every = k`` the forward keeps the full state
(heads, flows, emitter flows, statuses and terminal factor) of every k-th
step only; the backward recomputes the other steps of a segment from its
checkpoint (the solve is deterministic, so the recomputed state has the
forward's bits) just before their adjoints, so at most T/k + k - 1 step
states are held at once.  k = 1 keeps every step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .network import ST_CLOSED, ST_OPEN
from .solver import GGASolver

SEC_PER_DAY = 86400


def demand_factor(net, t_sec):
    """Per-node demand multiplier at time t: the node's pattern factor for
    period (t + Pstart) // Pstep (1 without a pattern) times the global
    demand multiplier."""
    pstep = int(net.meta.get("pat_step_sec", 3600)) or 3600
    p = (int(t_sec) + int(net.meta.get("pat_start_sec", 0))) // pstep
    f = np.ones(net.N)
    dp = np.asarray(net.dem_pat)
    for i, pat in enumerate(net.patterns):
        f[dp == i] = float(pat[p % len(pat)])
    return f * float(net.meta.get("demand_multiplier", 1.0))


class EpsModel:
    """B-scenario fixed-step EPS on one network (see the module docstring).

    ``simulate(demand, res_head, ke, r_hw, tank_h0, observe)`` returns the
    heads at the ``observe`` nodes for every step, [T, B, n_obs], and is
    differentiable with respect to demand [B,N], res_head ([N] or [B,N];
    reservoir entries), ke ([N] or [B,N]), r_hw ([L] or [B,L]) and tank_h0
    ([nt] or [B,nt]).  Observing a tank node gives its level head h_t."""

    def __init__(self, net, device="cuda", steps=None, dt=3600, accuracy=1e-6,
                 polish_steps=3, max_iter=200, checkpoint_every=4, solver=None):
        self.net = net
        self.s = solver if solver is not None else GGASolver(net, device=device,
                                                             dense_status_machine=True)
        s = self.s
        self.device = s.device
        self.dt = int(dt)
        self.T = int(steps) if steps is not None else int(net.meta["duration_sec"]) // self.dt
        if self.T < 1:
            raise ValueError("the simulation needs at least one step")
        self.accuracy = float(accuracy)
        self.polish_steps = int(polish_steps)
        self.max_iter = int(max_iter)
        self.k = int(checkpoint_every)
        if self.k < 1:
            raise ValueError("checkpoint_every must be >= 1")
        dev = self.device
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        self.factors = f64(np.stack([demand_factor(net, t * self.dt) for t in range(self.T)]))
        tanks = np.asarray(net.tank_node, dtype=np.int64)
        self.nt = tanks.size
        self.tanks = torch.as_tensor(tanks, device=dev)
        vmin, vmax, _ = net.tank_volumes()
        self.vmin_np, self.hmin_np = np.asarray(vmin, np.float64), np.asarray(net.tank_hmin,
                                                                             np.float64)
        self.area_np = np.asarray(net.tank_area, dtype=np.float64)
        self.vmin, self.vmax = f64(vmin), f64(vmax)
        self.area, self.hmin = f64(self.area_np), f64(self.hmin_np)
        self.el = f64(np.asarray(net.elev_ft)[tanks])
        # links into each tank: +1 where n2 is the tank, then -1 where n1 is
        n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
        self.inc = [[(int(k), 1.0) for k in np.nonzero(n2 == n)[0]] +
                    [(int(k), -1.0) for k in np.nonzero(n1 == n)[0]] for n in tanks]
        tlist = tanks.tolist()
        self.controls = []
        for i in range(len(net.ctl_link)):
            n = int(net.ctl_node[i])
            if n >= 0 and n not in tlist:
                raise NotImplementedError("level controls on a reservoir are not supported")
            j = tlist.index(n) if n >= 0 else -1
            v2 = None
            if j >= 0:
                v2 = float(self.vmin_np[j] + (float(net.ctl_grade[i]) - self.hmin_np[j])
                           * self.area_np[j])
            self.controls.append((int(net.ctl_link[i]), j, int(net.ctl_type[i]), v2,
                                  int(net.ctl_time[i]), int(net.ctl_status[i])))
        self.chp_q0 = {int(k): float(net.pump_q0[j]) for j, k in enumerate(net.pump_link)
                       if int(net.pump_ptype[j]) == 0}
        self.tstart = int(net.meta.get("tstart_sec", 0))
        self.init_flow = f64(s.init_flow_np)
        self.init_status = torch.as_tensor(np.asarray(s.init_status_int_np, dtype=np.int8),
                                           device=dev)
        self.solves = 0
        # record_decisions: keep every step's input statuses and tank clamps
        # of the last forward in ``decisions`` (S_in [T,B,L] int8, clamp
        # [T,B,nt] bool), the discrete decisions the adjoint freezes
        self.record_decisions = False
        self.decisions = None

    # ---- one step -------------------------------------------------------
    def _controls(self, t, S, q, h, qnet):
        """In place on S [B,L] int8 and q [B,L]."""
        t_sec = t * self.dt
        for k, j, ctype, v2, ctime, s2 in self.controls:
            if j >= 0:
                if ctype not in (0, 1):
                    continue
                vplus = qnet[:, j].abs()
                v1 = self.vmin[j] + (h[:, j] - self.hmin[j]) * self.area[j]
                reset = v1 <= v2 + vplus if ctype == 0 else v1 >= v2 - vplus
            elif ctype == 2:
                if ctime != t_sec:
                    continue
                reset = torch.ones(S.shape[0], dtype=torch.bool, device=S.device)
            elif ctype == 3:
                if (t_sec + self.tstart) % SEC_PER_DAY != ctime:
                    continue
                reset = torch.ones(S.shape[0], dtype=torch.bool, device=S.device)
            else:
                continue
            sk = S[:, k]
            closed1 = sk <= ST_CLOSED
            if k in self.chp_q0 and s2 == ST_OPEN:
                q[:, k] = torch.where(reset & closed1, torch.full_like(q[:, k], self.chp_q0[k]),
                                      q[:, k])
            s1 = torch.where(closed1, torch.full_like(sk, ST_CLOSED), torch.full_like(sk, ST_OPEN))
            chg = reset & (s1 != s2)
            S[:, k] = torch.where(chg, torch.full_like(sk, s2), sk)

    def _solve(self, t, inp, q, e, S, h, want_factor):
        """The warm-started fused solve of step t; returns its output dict
        and the step's demand tensor."""
        s = self.s
        demand, rh, ke, r = inp
        d = demand * self.factors[t]
        rht = rh.expand(demand.shape[0], -1).clone()
        rht[:, self.tanks] = h
        a, out, keep, _, _ = s._args(d, rht, ke, q, e, self.max_iter, self.accuracy, True,
                                     r_hw=r, polish_steps=self.polish_steps,
                                     want_factor=want_factor, status0=S,
                                     status0_trusted=True)
        with torch.cuda.device(self.device):
            _lib.check(s._timed("k_solve", lambda: s._lib.wds_solve(s._plan, s._stream(),
                                                                   C.byref(a)),
                                d.shape[0], out["iters"]), "wds_solve")
        self.solves += 1
        return out, keep

    def _tank_update(self, V, Q):
        """(V_next, clamped, qnet, h_next) from V [B,nt] and step flows Q."""
        B = V.shape[0]
        qnet = torch.zeros(B, self.nt, dtype=V.dtype, device=V.device)
        for j, links in enumerate(self.inc):
            for k, sg in links:
                qnet[:, j] = qnet[:, j] + sg * Q[:, k]
        Vn = V + qnet * float(self.dt)
        full = Vn + qnet >= self.vmax
        empty = ~full & (Vn - qnet <= self.vmin)
        Vn = torch.where(full, self.vmax.expand_as(Vn), torch.where(empty, self.vmin.expand_as(Vn),
                                                                    Vn))
        h = self.hmin + (Vn - self.vmin) / self.area
        return Vn, full | empty, qnet, h

    def _advance(self, t, inp, st, want_factor):
        """Controls, solve and tank update of step t from the carried state
        st = (q, e, S_out, h, qnet, V); returns (out, new state, clamp,
        input statuses)."""
        q, e, S, h, qnet, V = st
        q = q.clone()
        S = S.clone()
        self._controls(t, S, q, h, qnet)
        out, _ = self._solve(t, inp, q, e, S, h, want_factor)
        Vn, clamp, qn, hn = self._tank_update(V, out["flow"])
        return out, (out["flow"], out["emitter"], out["status"], hn, qn, Vn), clamp, S

    def initial_state(self, B, tank_h0):
        h0 = tank_h0.expand(B, -1).contiguous()
        V0 = self.area * (h0 - self.el)
        q = self.init_flow.expand(B, -1).contiguous()
        S = self.init_status.expand(B, -1).contiguous()
        z = torch.zeros(B, self.nt, dtype=torch.float64, device=self.device)
        return (q, None, S, h0, z, V0)

    def simulate(self, demand, res_head, ke=None, r_hw=None, tank_h0=None, observe=None):
        s = self.s
        dev, dt = self.device, torch.float64
        t = lambda x: x.to(dtype=dt, device=dev) if isinstance(x, torch.Tensor) \
            else torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
        ke = s.node_ke_default if ke is None else ke
        r_hw = s.r_hw if r_hw is None else r_hw
        tank_h0 = self.net.tank_h0 if tank_h0 is None else tank_h0
        if observe is None:
            observe = self.tanks
        obs = torch.as_tensor(np.asarray(observe.cpu() if isinstance(observe, torch.Tensor)
                                         else observe), dtype=torch.int64)
        if obs.dim() != 1 or obs.unique().numel() != obs.numel() or \
                (obs.numel() and (int(obs.min()) < 0 or int(obs.max()) >= s.N)):
            raise ValueError("observe must list distinct node indices in [0, N)")
        args = (t(demand), t(res_head), t(ke), t(r_hw), t(tank_h0))
        if args[0].dim() != 2 or args[0].shape[1] != s.N:
            raise ValueError("demand must be [B,N] with N=%d" % s.N)
        B = args[0].shape[0]
        for x, n, what in ((args[1], s.N, "res_head"), (args[2], s.N, "ke"),
                           (args[3], s.L, "r_hw"), (args[4], self.nt, "tank_h0")):
            if x.shape[-1] != n or x.dim() not in (1, 2) or (x.dim() == 2 and
                                                             x.shape[0] not in (1, B)):
                raise ValueError("%s must be [%d] or [B,%d]" % (what, n, n))
        return _EpsFunction.apply(*args, self, obs.to(dev))


def _check_steps(s, sts, max_iter, what):
    """sts: the status words (info, iters) of every step's solve."""
    if not sts:
        return
    st = torch.stack(sts).cpu().numpy()
    B = st.shape[1] // 2
    info, iters = st[:, :B], st[:, B:]
    bad = np.argwhere((info != 0) | (iters > s.maxtrials(max_iter, True)))
    if bad.size:
        t, b = bad[0]
        raise RuntimeError("EPS %s: step %d scenario %d failed or did not converge "
                           "(info %d, iters %d); %d (step, scenario) pairs in all"
                           % (what, t, b, info[t, b], iters[t, b], len(bad)))


class _EpsFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, demand, res_head, ke, r_hw, tank_h0, model, obs):
        m = model
        B = demand.shape[0]
        rh = res_head if res_head.dim() == 2 else res_head.unsqueeze(0)
        inp = (demand.contiguous(), torch.nan_to_num(rh).contiguous(), ke, r_hw)
        st = m.initial_state(B, tank_h0 if tank_h0.dim() == 2 else tank_h0.unsqueeze(0))
        H_obs = torch.empty(m.T, B, obs.numel(), dtype=torch.float64, device=m.device)
        ckpt, clamps, outs, s_in = {}, [], [], []
        for t in range(m.T):
            keep = t % m.k == 0
            if keep:
                ckpt[t] = st
            out, st, clamp, S = m._advance(t, inp, st, want_factor=keep)
            H_obs[t] = out["head"].index_select(1, obs)
            clamps.append(clamp)
            outs.append(out["st"])
            if m.record_decisions:
                s_in.append(S)
            if keep:
                ckpt[t] = (ckpt[t], out)
        _check_steps(m.s, outs, m.max_iter, "forward")
        del outs
        m.decisions = (torch.stack(s_in), torch.stack(clamps)) if m.record_decisions else None
        ctx.model, ctx.obs, ctx.inp, ctx.ckpt, ctx.clamps = m, obs, inp, ckpt, clamps
        ctx.shapes = (res_head.dim(), ke.dim(), r_hw.dim(), tank_h0.dim())
        return H_obs

    @staticmethod
    def backward(ctx, g_obs):
        if torch.is_grad_enabled():
            raise NotImplementedError("the EPS adjoint is first order only")
        m = ctx.model
        s = m.s
        demand, rh, ke, r = ctx.inp
        B, dev = demand.shape[0], m.device
        g_d = torch.zeros_like(demand)
        g_rh = torch.zeros(B, s.N, dtype=torch.float64, device=dev)
        g_ke = torch.zeros_like(g_rh)
        g_r = torch.zeros(B, s.L, dtype=torch.float64, device=dev)
        g_h0 = None
        nu = torch.zeros(B, m.nt, dtype=torch.float64, device=dev)
        ws = s._workspace(B)
        infos, recomputed = [], []
        starts = sorted(ctx.ckpt)
        for c in reversed(starts):
            end = min(c + m.k, m.T)
            st_in, out_c = ctx.ckpt.pop(c)
            states = {c: out_c}
            st = None
            for t in range(c + 1, end):
                if st is None:
                    # the carried state after step c, rebuilt from its output
                    Vc = st_in[5]
                    Vn, _, qn, hn = m._tank_update(Vc, out_c["flow"])
                    st = (out_c["flow"], out_c["emitter"], out_c["status"], hn, qn, Vn)
                out, st, _, _ = m._advance(t, ctx.inp, st, want_factor=True)
                states[t] = out
                recomputed.append(out["st"])
            for t in range(end - 1, c - 1, -1):
                out = states.pop(t)
                w = torch.where(ctx.clamps[t], torch.zeros_like(nu), nu)
                gQ = torch.zeros(B, s.L, dtype=torch.float64, device=dev)
                for j, links in enumerate(m.inc):
                    for k, sg in links:
                        gQ[:, k] += (sg * float(m.dt)) * w[:, j]
                gH = torch.zeros(B, s.N, dtype=torch.float64, device=dev)
                if g_obs is not None:
                    gH[:, ctx.obs] = g_obs[t]
                d = demand * m.factors[t]
                gd = torch.empty(B, s.N, dtype=torch.float64, device=dev)
                grh, gke = torch.empty_like(gd), torch.empty_like(gd)
                gr = torch.empty(B, s.L, dtype=torch.float64, device=dev)
                info = torch.empty(B, dtype=torch.int32, device=dev)
                a = _lib.AdjointArgs(
                    B=B, demand=_lib.ptr(d), ke=_lib.ptr(ke),
                    ke_stride=s.N if ke.dim() == 2 else 0, r_hw=_lib.ptr(r),
                    r_stride=s.L if r.dim() == 2 else 0, head=_lib.ptr(out["head"]),
                    flow=_lib.ptr(out["flow"]), emitter=_lib.ptr(out["emitter"]),
                    status=_lib.ptr(out["status"]), factor=_lib.ptr(out["factor"]),
                    gH=_lib.ptr(gH), gQ=_lib.ptr(gQ), gE=None,
                    n_refine=int(s.cudss_grad_refine), g_demand=_lib.ptr(gd),
                    g_res_head=_lib.ptr(grh), g_ke=_lib.ptr(gke), g_r=_lib.ptr(gr),
                    info=_lib.ptr(info), workspace=_lib.ptr(ws))
                with torch.cuda.device(dev):
                    _lib.check(s._timed("k_adjoint",
                                        lambda: s._lib.wds_adjoint(s._plan, s._stream(),
                                                                   C.byref(a)), B),
                               "wds_adjoint")
                infos.append(info)
                g_d += m.factors[t] * gd
                g_ke += gke
                g_r += gr
                gtank = grh.index_select(1, m.tanks)
                grh[:, m.tanks] = 0.0
                g_rh += grh
                if t > 0:
                    nu = gtank / m.area + w
                else:
                    g_h0 = gtank + m.area * w
                del out
        _check_steps(s, recomputed, m.max_iter, "recompute")
        info = torch.stack(infos).cpu().numpy()
        if (info != 0).any():
            raise RuntimeError("EPS adjoint: the node adjoint failed at %d (step, scenario) "
                               "pairs" % int((info != 0).sum()))
        d_rh, d_ke, d_r, d_h0 = ctx.shapes
        fit = lambda g, dim: g.sum(0) if dim == 1 else g
        ctx.ckpt = ctx.clamps = ctx.inp = None
        return (g_d, fit(g_rh, d_rh), fit(g_ke, d_ke), fit(g_r, d_r), fit(g_h0, d_h0),
                None, None)
