# -*- coding: utf-8 -*-
"""Oracle: the fixed-step differentiable extended-period model and its
discrete adjoint.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The model (the one wdsgpu_amd.eps_adjoint runs on the GPU) takes T steps of
dt seconds.  Step t, for every scenario:
  1. simple controls on the statuses carried from step t-1 (EpsDriver
     ``_controls``, dgga/eps.py:171-218: LOWLEVEL / HILEVEL on tank volumes
     with the tolerance vplus = |net tank inflow of step t-1|, TIMER,
     TIMEOFDAY; a constant-horsepower pump reopened restarts from its design
     flow);
  2. demands d_t = demand * f_t, f_t the pattern factor of period
     (t dt + Pstart) // Pstep (dgga/eps.py:125-147);
  3. fixed heads: reservoirs from res_head, tanks at h_t (h_0 = tank_h0);
  4. the polished GGA solve warm-started from step t-1's flows, emitter
     flows and statuses (solve_polished);
  5. tank update (EpsDriver ``_tanklevels``, dgga/eps.py:278-291): net
     inflow qnet_j = sum of +Q over links with n2 = tank, then -Q over links
     with n1 = tank (link order); V += qnet dt; V = Vmax if V + qnet >= Vmax,
     else V = Vmin if V - qnet <= Vmin (the one-second look-ahead clamp);
     h = hmin + (V - Vmin) / A.  V_0 = A (h_0 - El).
Unlike EpsDriver the step is fixed (no tank-fill or control-crossing
sub-steps): a control acts at the first step boundary after its threshold
is crossed.

The adjoint sweeps the steps backwards with statuses, control decisions and
tank clamps frozen.  With nu_t = dLoss/dV_t (t >= 1):
  gQ_t[k]  = sign_jk dt nu_{t+1}[j]          (tank j not clamped at t -> t+1)
  step t's implicit adjoint (cpu_adjoint) with cotangents (gH_t, gQ_t)
  nu_t     = g_res_head_t[tank_j] / A_j + nu_{t+1}[j]  (not clamped)
  dLoss/dh_0 = g_res_head_0[tank_j] + A_j nu_1[j]      (not clamped at 0 -> 1)
and the parameter gradients sum over the steps: demand f_t g_d_t,
reservoir heads, emitter Ke and pipe resistance r.  Pinned by central finite
differences of the whole T-step loss (tests/test_eps_adjoint.py) and, for
the forward, by equality with the event-driven driver (oracle/eps.py
run_eps) on a network whose EPS takes only whole-hour steps.
"""
from __future__ import annotations

import numpy as np

from wdsgpu_amd.network import ST_CLOSED, ST_OPEN

from . import adjoint as adj
from . import gga

SEC_PER_DAY = 86400


def demand_factor(net, t_sec):
    """Per-node demand multiplier at time t (pattern factor times the
    global demand multiplier; 1 for nodes without a pattern)."""
    pstep = int(net.meta.get("pat_step_sec", 3600)) or 3600
    p = (int(t_sec) + int(net.meta.get("pat_start_sec", 0))) // pstep
    f = np.ones(net.N)
    dp = np.asarray(net.dem_pat)
    for i, pat in enumerate(net.patterns):
        f[dp == i] = float(pat[p % len(pat)])
    return f * float(net.meta.get("demand_multiplier", 1.0))


def tank_incidence(net):
    """Per tank: [(link, +1) for links with n2 = tank] + [(link, -1) for n1 =
    tank], each part in link order."""
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
    return [[(int(k), 1.0) for k in np.nonzero(n2 == n)[0]] +
            [(int(k), -1.0) for k in np.nonzero(n1 == n)[0]] for n in net.tank_node]


def _controls(net, t_sec, S, q, h_tank, qnet, vmin, hmin, area, chp_q0):
    """In place on S [L], q [L] of one scenario."""
    tanks = list(np.asarray(net.tank_node))
    tstart = int(net.meta.get("tstart_sec", 0))
    for i in range(len(net.ctl_link)):
        k, n, ctype = int(net.ctl_link[i]), int(net.ctl_node[i]), int(net.ctl_type[i])
        reset = False
        if n >= 0:
            j = tanks.index(n)
            vplus = abs(qnet[j])
            v1 = vmin[j] + (h_tank[j] - hmin[j]) * area[j]
            v2 = vmin[j] + (float(net.ctl_grade[i]) - hmin[j]) * area[j]
            if ctype == 0 and v1 <= v2 + vplus:
                reset = True
            if ctype == 1 and v1 >= v2 - vplus:
                reset = True
        if ctype == 2 and int(net.ctl_time[i]) == t_sec:
            reset = True
        if ctype == 3 and (t_sec + tstart) % SEC_PER_DAY == int(net.ctl_time[i]):
            reset = True
        if reset:
            s1 = ST_CLOSED if S[k] <= ST_CLOSED else ST_OPEN
            s2 = int(net.ctl_status[i])
            if s1 == ST_CLOSED and s2 == ST_OPEN and k in chp_q0:
                q[k] = chp_q0[k]
            if s1 != s2:
                S[k] = s2


def forward(net, demand, res_head, ke, r_hw, tank_h0, steps, dt=3600, accuracy=1e-6,
            polish_steps=3, max_iter=200, linsolve="sparse"):
    """demand [B,N] (base, before patterns); res_head [N] (reservoir heads;
    tank entries unused); ke [N]; r_hw [L]; tank_h0 [B,nt].  Returns dict:
    sols (per step: solve_polished's dict), d (per step demands), V [T+1,B,nt],
    clamp [T,B,nt] bool, S_in [T,B,L], H [T,B,N], qnet [T,B,nt]."""
    m = gga.GGAModel(net)
    demand = np.atleast_2d(np.asarray(demand, dtype=np.float64))
    B, N, L = demand.shape[0], net.N, net.L
    tank_h0 = np.broadcast_to(np.atleast_2d(tank_h0), (B, len(net.tank_node))).astype(
        np.float64)
    vmin, vmax, _ = net.tank_volumes()
    area = np.asarray(net.tank_area, dtype=np.float64)
    hmin = np.asarray(net.tank_hmin, dtype=np.float64)
    el = np.asarray(net.elev_ft)[net.tank_node]
    inc = tank_incidence(net)
    chp_q0 = {int(k): float(net.pump_q0[j]) for j, k in enumerate(net.pump_link)
              if int(net.pump_ptype[j]) == 0}
    tanks = np.asarray(net.tank_node)
    nt = tanks.size
    rh0 = np.nan_to_num(np.asarray(res_head, dtype=np.float64))
    S = np.broadcast_to(m.init_status_int, (B, L)).astype(np.int8).copy()
    q = np.broadcast_to(m.init_flow(), (B, L)).copy()
    e = None
    V = np.zeros((steps + 1, B, nt))
    V[0] = area * (tank_h0 - el)
    h = tank_h0.copy()
    qnet = np.zeros((B, nt))
    out = dict(sols=[], d=[], V=V, clamp=np.zeros((steps, B, nt), dtype=bool),
               S_in=np.zeros((steps, B, L), dtype=np.int8), H=np.zeros((steps, B, N)),
               qnet=np.zeros((steps, B, nt)))
    for t in range(steps):
        ts = t * dt
        for b in range(B):
            _controls(net, ts, S[b], q[b], h[b], qnet[b], vmin, hmin, area, chp_q0)
        out["S_in"][t] = S
        d = demand * demand_factor(net, ts)
        rh = np.broadcast_to(rh0, (B, N)).copy()
        rh[:, tanks] = h
        sol = adj.solve_polished(m, d, rh, ke, r_hw, accuracy=accuracy, max_iter=max_iter,
                                 polish_steps=polish_steps, linsolve=linsolve, q0=q, e0=e,
                                 status0=S)
        out["sols"].append(sol)
        out["d"].append(d)
        q = sol["q"].copy()
        e = sol["emitter"].copy()
        S = sol["status"].astype(np.int8).copy()
        out["H"][t] = sol["head"]
        qnet = np.zeros((B, nt))
        for j in range(nt):
            for k, sg in inc[j]:
                qnet[:, j] = qnet[:, j] + sg * q[:, k]
        out["qnet"][t] = qnet
        Vn = V[t] + qnet * dt
        full = Vn + qnet >= vmax
        empty = ~full & (Vn - qnet <= vmin)
        Vn = np.where(full, vmax, np.where(empty, vmin, Vn))
        V[t + 1] = Vn
        out["clamp"][t] = full | empty
        h = hmin + (Vn - vmin) / area
    return out


def adjoint(net, fw, gH, dt=3600):
    """gH [T,B,N]: dLoss/dH_t (heads at every node; at a tank node the head
    is h_t).  Returns (g_demand [B,N], g_res_head [B,N] (reservoirs),
    g_ke [B,N], g_r [B,L], g_tank_h0 [B,nt])."""
    m = gga.GGAModel(net)
    T = len(fw["sols"])
    B = fw["V"].shape[1]
    area = np.asarray(net.tank_area, dtype=np.float64)
    tanks = np.asarray(net.tank_node)
    nt = tanks.size
    inc = tank_incidence(net)
    g_d = np.zeros((B, net.N))
    g_rh = np.zeros((B, net.N))
    g_ke = np.zeros((B, net.N))
    g_r = np.zeros((B, net.L))
    nu_next = np.zeros((B, nt))
    for t in range(T - 1, -1, -1):
        free = ~fw["clamp"][t]
        w = np.where(free, nu_next, 0.0)
        gQ = np.zeros((B, net.L))
        for j in range(nt):
            for k, sg in inc[j]:
                gQ[:, k] += sg * dt * w[:, j]
        gd, grh, gke, gr = adj.cpu_adjoint(m, fw["sols"][t], gH[t], gQ, np.zeros((B, net.N)))
        f = demand_factor(net, t * dt)
        g_d += f * gd
        g_ke += gke
        g_r += gr
        gtank = grh[:, tanks]
        grh[:, tanks] = 0.0
        g_rh += grh
        if t > 0:
            nu_next = gtank / area + w
        else:
            g_h0 = gtank + area * w
    return g_d, g_rh, g_ke, g_r, g_h0
