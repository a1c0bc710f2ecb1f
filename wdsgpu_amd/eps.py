# -*- coding: utf-8 -*-
"""Batched extended-period simulation on the GPU solve.

``EpsDriver`` follows dgga/eps.py ``EpsDriver`` (40-414) for B scenarios at
once: every scenario keeps its own clock, tank volumes, link statuses and
warm-start flows, and each round solves every unfinished scenario's current
frame in one fused launch (``GGASolver.solve`` with the previous frame's
flows, emitter flows and statuses as the warm start).  Between frames the
host applies, per scenario:

- demands: base demand times the node's pattern factor for period
  (t + Pstart) // Pstep (demands, hydraul.c:465-504), times an optional
  per-scenario multiplier;
- simple controls (controls, hydraul.c:538-619): LOWLEVEL / HILEVEL on tank
  volume with the one-second tolerance, TIMER, TIMEOFDAY; statuses only;
- timestep (hydraul.c:622-659): hydraulic step, pattern boundary, report
  time, tanktimestep (:662-707) and controltimestep (:710-783);
- tanklevels (:998-1035): explicit Euler with the one-second look-ahead
  clamp, cylindrical tanks;
- nexthyd (:225-275): a frame that does not converge halts the scenario.

Out of scope as in the reference's batched path: rules, pump speed
patterns and setting changes, variable-head reservoirs, volume curves.
"""
from __future__ import annotations

import numpy as np
import torch

from .network import QZERO, ST_CLOSED, ST_OPEN
from .solver import GGASolver

SEC_PER_DAY = 86400


def _round_c(x):
    """ROUND of types.h:103, elementwise: (int)(x + .5) or (int)(x - .5)."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x >= 0.0, np.trunc(x + 0.5), np.trunc(x - 0.5)).astype(np.int64)


class EpsDriver:
    def __init__(self, net, device="cuda", solver=None, accuracy=None):
        self.net = net
        self.s = solver if solver is not None else GGASolver(net, device=device,
                                                             dense_status_machine=True)
        m = net.meta
        self.dur = int(m["duration_sec"])
        self.pstep = int(m.get("pat_step_sec", 3600)) or 3600
        self.rstep = int(m.get("report_step_sec", 0)) or self.pstep
        self.hstep = min(int(m.get("hyd_step_sec", 3600)) or 3600, self.pstep, self.rstep)
        self.pstart = int(m.get("pat_start_sec", 0))
        self.tstart = int(m.get("tstart_sec", 0))
        self.hacc = self.s.hacc_default if accuracy is None else float(accuracy)
        self.tanks = np.asarray(net.tank_node, dtype=np.int64)
        self.vmin, self.vmax, self.v0 = net.tank_volumes()
        self.area = np.asarray(net.tank_area, dtype=np.float64)
        self.hmin = np.asarray(net.tank_hmin, dtype=np.float64)
        self.hmax = np.asarray(net.tank_hmax, dtype=np.float64)
        self.n1 = np.asarray(net.link_n1, dtype=np.int64)
        self.n2 = np.asarray(net.link_n2, dtype=np.int64)
        self.ctl_tank = np.array([int(np.nonzero(self.tanks == n)[0][0]) if n >= 0 else -1
                                  for n in np.asarray(net.ctl_node)], dtype=np.int64)
        if np.any((np.asarray(net.ctl_node) >= 0) & (self.ctl_tank < 0)):
            raise NotImplementedError("level controls on a reservoir are not supported")
        # design flows of the constant-horsepower pumps (resetpumpflow)
        self.chp_q0 = {int(k): float(net.pump_q0[j]) for j, k in enumerate(net.pump_link)
                       if int(net.pump_ptype[j]) == 0}

    def _tank_vol(self, j, h):
        return self.vmin[j] + (h - self.hmin[j]) * self.area[j]

    def run(self, max_frames=100000):
        """One scenario (the reference's EpsDriver.run); frames dict."""
        return self.run_batch(1, max_frames=max_frames)[0]

    def run_batch(self, B, demand_scale=None, max_frames=100000):
        """B scenarios; demand_scale [B] or [B,N] multiplies the demands.
        Returns one frames dict per scenario (t_sec, head_ft, flow_cfs,
        status, iterations, relerr, tank_head_ft)."""
        net, s = self.net, self.s
        N, L, nt = net.N, net.L, len(self.tanks)
        dev = s.device
        scale = np.ones((B, N))
        if demand_scale is not None:
            ds = np.asarray(demand_scale, dtype=np.float64)
            scale = np.broadcast_to(ds[:, None] if ds.ndim == 1 else ds, (B, N)).copy()
        # inithyd
        S = np.broadcast_to(s.init_status_int_np, (B, L)).copy()
        q = np.broadcast_to(s.init_flow_np, (B, L)).copy()
        e = None
        V = np.broadcast_to(self.v0, (B, nt)).copy()
        H = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), (B, N)).copy()
        fixed_dem = np.zeros((B, N))
        htime = np.zeros(B, dtype=np.int64)
        rtime = np.full(B, self.rstep, dtype=np.int64)
        halt = np.zeros(B, dtype=bool)
        active = np.ones(B, dtype=bool)
        frames = [{k: [] for k in ("t_sec", "head_ft", "flow_cfs", "status", "iterations",
                                   "relerr", "tank_head_ft")} for _ in range(B)]
        ke = torch.as_tensor(net.node_ke, dtype=torch.float64, device=dev)
        while active.any():
            idx = np.nonzero(active)[0]
            n = idx.size
            d = np.stack([net.demand_cfs_at(int(htime[b])) for b in idx]) * scale[idx]
            self._controls(idx, htime, H, V, fixed_dem, S, q)
            out = s.solve(torch.as_tensor(d, device=dev), torch.as_tensor(H[idx], device=dev),
                          ke_int=ke, q0=torch.as_tensor(q[idx], device=dev),
                          e0=None if e is None else torch.as_tensor(e[idx], device=dev),
                          accuracy=self.hacc, status_machine=True,
                          link_status=torch.as_tensor(S[idx], device=dev))
            q[idx] = out["flow_cfs"].cpu().numpy().reshape(n, L)
            S[idx] = out["status"].cpu().numpy().reshape(n, L)
            if e is None:
                e = np.zeros((B, N))
            e[idx] = out["emitter_cfs"].cpu().numpy().reshape(n, N)
            H[idx] = out["head_ft"].cpu().numpy().reshape(n, N)
            relerr = out["relerr"].cpu().numpy().reshape(n)
            iters = out["iters"].cpu().numpy().reshape(n)
            fd = np.zeros((n, N))
            np.add.at(fd, (slice(None), self.n2), q[idx])
            np.add.at(fd, (slice(None), self.n1), -q[idx])
            fixed_dem[idx] = fd
            for i, b in enumerate(idx):
                f = frames[b]
                f["t_sec"].append(int(htime[b]))
                f["head_ft"].append(H[b].copy())
                f["flow_cfs"].append(q[b].copy())
                f["status"].append(S[b].copy())
                f["iterations"].append(int(iters[i]))
                f["relerr"].append(float(relerr[i]))
                f["tank_head_ft"].append(H[b, self.tanks].copy())
                if len(f["t_sec"]) >= max_frames:
                    raise RuntimeError("too many frames for scenario %d" % b)
            # nexthyd
            halt[idx] |= (relerr > self.hacc) & (s.extra_iter == -1)
            htime[idx] = np.where(halt[idx], self.dur, htime[idx])
            for b in idx:
                tstep = 0
                if htime[b] < self.dur:
                    tstep = self._timestep(b, htime, rtime, H, V, fixed_dem, S)
                    htime[b] += tstep
                    if htime[b] >= rtime[b]:
                        rtime[b] += self.rstep
                else:
                    htime[b] += 1
                if tstep == 0:
                    active[b] = False
        return [{k: np.asarray(v) for k, v in f.items()} for f in frames]

    def _controls(self, idx, htime, H, V, fixed_dem, S, q):
        net = self.net
        for i in range(len(net.ctl_link)):
            k, n, ctype = int(net.ctl_link[i]), int(net.ctl_node[i]), int(net.ctl_type[i])
            reset = np.zeros(idx.size, dtype=bool)
            if n >= 0:
                j = self.ctl_tank[i]
                vplus = np.abs(fixed_dem[idx, n])
                v1 = self._tank_vol(j, H[idx, n])
                v2 = self._tank_vol(j, float(net.ctl_grade[i]))
                if ctype == 0:
                    reset |= v1 <= v2 + vplus
                if ctype == 1:
                    reset |= v1 >= v2 - vplus
            if ctype == 2:
                reset |= int(net.ctl_time[i]) == htime[idx]
            if ctype == 3:
                reset |= (htime[idx] + self.tstart) % SEC_PER_DAY == int(net.ctl_time[i])
            s1 = np.where(S[idx, k] <= ST_CLOSED, ST_CLOSED, ST_OPEN)
            s2 = int(net.ctl_status[i])
            # a constant-horsepower pump reopened by a control restarts from
            # its design flow (resetpumpflow, dgga/eps.py:205-209)
            if k in self.chp_q0 and s2 == ST_OPEN:
                reopen = reset & (s1 == ST_CLOSED)
                q[idx[reopen], k] = self.chp_q0[k]
            chg = reset & (s1 != s2)
            S[idx[chg], k] = s2

    def _timestep(self, b, htime, rtime, H, V, fixed_dem, S):
        net = self.net
        t_now = int(htime[b])
        tstep = self.hstep
        nb = (t_now + self.pstart) // self.pstep + 1
        t = nb * self.pstep - t_now
        if 0 < t < tstep:
            tstep = t
        t = int(rtime[b]) - t_now
        if 0 < t < tstep:
            tstep = t
        qn = fixed_dem[b, self.tanks]
        h = H[b, self.tanks]
        for j in range(len(self.tanks)):
            if abs(qn[j]) <= QZERO:
                continue
            if qn[j] > 0.0 and h[j] < self.hmax[j]:
                v = self.vmax[j] - V[b, j]
            elif qn[j] < 0.0 and h[j] > self.hmin[j]:
                v = self.vmin[j] - V[b, j]
            else:
                continue
            t = int(_round_c(v / qn[j]))
            if 0 < t < tstep:
                tstep = t
        for i in range(len(net.ctl_link)):
            t = 0
            n, ctype = int(net.ctl_node[i]), int(net.ctl_type[i])
            if n >= 0:
                j = self.ctl_tank[i]
                qj = fixed_dem[b, n]
                if abs(qj) > QZERO:
                    grade = float(net.ctl_grade[i])
                    if (H[b, n] < grade and ctype == 1 and qj > 0.0) or \
                            (H[b, n] > grade and ctype == 0 and qj < 0.0):
                        t = int(_round_c((self._tank_vol(j, grade) - V[b, j]) / qj))
            if ctype == 2 and int(net.ctl_time[i]) > t_now:
                t = int(net.ctl_time[i]) - t_now
            if ctype == 3:
                t1 = (t_now + self.tstart) % SEC_PER_DAY
                t2 = int(net.ctl_time[i])
                t = t2 - t1 if t2 >= t1 else SEC_PER_DAY - t1 + t2
            if 0 < t < tstep and S[b, int(net.ctl_link[i])] != int(net.ctl_status[i]):
                tstep = t
        # tanklevels
        for j, n in enumerate(self.tanks):
            V[b, j] += fixed_dem[b, n] * tstep
            if V[b, j] + fixed_dem[b, n] >= self.vmax[j]:
                V[b, j] = self.vmax[j]
            elif V[b, j] - fixed_dem[b, n] <= self.vmin[j]:
                V[b, j] = self.vmin[j]
            H[b, n] = self.hmin[j] + (V[b, j] - self.vmin[j]) / self.area[j]
        return tstep
