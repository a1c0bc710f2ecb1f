# -*- coding: utf-8 -*-
"""Oracle: scalar extended-period driver over the oracle's GGA solve.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates dgga/eps.py ``EpsDriver`` (40-414) for one scenario: inithyd
(85-179), demands (125-147, junction patterns), simple controls (161-217:
LOWLEVEL / HILEVEL on tank volumes with the one-second tolerance vplus,
TIMER, TIMEOFDAY; statuses only), the warm-started hydraulic solve (flows,
emitter flows and statuses carried between frames), timestep (292-310:
hydraulic step, pattern boundary, report time, tanktimestep 220-240,
controltimestep 242-273), tanklevels (275-290, explicit Euler with the
one-second clamp) and nexthyd (333-352, Haltflag on an unconverged frame).
"""
from __future__ import annotations

import numpy as np

from wdsgpu_amd.network import QZERO, ST_CLOSED

from . import gga

SEC_PER_DAY = 86400


def _round_c(x):
    return int(x + 0.5) if x >= 0.0 else int(x - 0.5)


def run_eps(net, demand_scale=None, accuracy=None, max_frames=100000, pow_jitter=None,
            extended=False):
    """One scenario; returns frames dict of arrays (t_sec, head_ft, flow_cfs,
    status, iterations, relerr, tank_head_ft).  ``pow_jitter`` (a numpy
    Generator) perturbs the power laws by one ulp, which measures the
    trajectory's own sensitivity to rounding.  ``extended`` solves every
    frame with the correctly rounded linear solves (gga.solve)."""
    m = gga.GGAModel(net)
    m.pow_jitter = pow_jitter
    meta = net.meta
    dur = int(meta["duration_sec"])
    pstep = int(meta.get("pat_step_sec", 3600)) or 3600
    rstep = int(meta.get("report_step_sec", 0)) or pstep
    hstep = int(meta.get("hyd_step_sec", 3600)) or 3600
    hstep = min(hstep, pstep, rstep)
    pstart = int(meta.get("pat_start_sec", 0))
    tstart = int(meta.get("tstart_sec", 0))
    hacc = m.hacc_default if accuracy is None else accuracy
    tanks = [int(n) for n in net.tank_node]
    vmin, vmax, v0 = net.tank_volumes()
    area = np.asarray(net.tank_area)
    hmin = np.asarray(net.tank_hmin)
    hmax = np.asarray(net.tank_hmax)
    scale = np.ones(net.N) if demand_scale is None else np.asarray(demand_scale)
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)

    # inithyd
    S = m.init_status_int.copy()
    q = m.init_flow()
    e = None
    V = v0.astype(np.float64).copy()
    H = np.nan_to_num(net.reservoir_head_ft())
    fixed_dem = np.zeros(net.N)
    htime, rtime, halt = 0, rstep, False
    frames = {k: [] for k in ("t_sec", "head_ft", "flow_cfs", "status", "iterations",
                              "relerr", "tank_head_ft")}

    def tank_vol(j, h):
        return vmin[j] + (h - hmin[j]) * area[j]

    # design flows of the constant-horsepower pumps (hydraul.c resetpumpflow)
    pump_q0 = {int(k): float(net.pump_q0[j]) for j, k in enumerate(net.pump_link)
               if int(net.pump_ptype[j]) == 0}

    while True:
        # demands
        d = net.demand_cfs_at(htime) * scale
        # controls (statuses only)
        for i in range(len(net.ctl_link)):
            k, n, ctype = int(net.ctl_link[i]), int(net.ctl_node[i]), int(net.ctl_type[i])
            reset = False
            if n >= 0:
                j = tanks.index(n)
                vplus = abs(fixed_dem[n])
                v1 = tank_vol(j, H[n])
                v2 = tank_vol(j, float(net.ctl_grade[i]))
                if ctype == 0 and v1 <= v2 + vplus:
                    reset = True
                if ctype == 1 and v1 >= v2 - vplus:
                    reset = True
            if ctype == 2 and int(net.ctl_time[i]) == htime:
                reset = True
            if ctype == 3 and (htime + tstart) % SEC_PER_DAY == int(net.ctl_time[i]):
                reset = True
            if reset:
                s1 = ST_CLOSED if S[k] <= ST_CLOSED else 3
                s2 = int(net.ctl_status[i])
                # a constant-horsepower pump reopened by a control restarts
                # from its design flow (resetpumpflow, dgga/eps.py:205-209)
                if s1 == ST_CLOSED and s2 == 3 and k in pump_q0:
                    q[k] = pump_q0[k]
                if s1 != s2:
                    S[k] = s2
        # hydraulic solve, warm started
        out = gga.solve(m, d, H, ke=net.node_ke, q0=q, e0=e, accuracy=hacc,
                        status_machine=True, status0=S, extended=extended)
        q, S = out["flow_cfs"], out["status"]
        e = out["emitter_cfs"]
        H = out["head_ft"].copy()
        relerr = float(out["relerr"])
        fixed_dem = np.zeros(net.N)
        np.add.at(fixed_dem, n2, q)
        np.add.at(fixed_dem, n1, -q)
        frames["t_sec"].append(htime)
        frames["head_ft"].append(H.copy())
        frames["flow_cfs"].append(q.copy())
        frames["status"].append(S.copy())
        frames["iterations"].append(int(out["iters"]))
        frames["relerr"].append(relerr)
        frames["tank_head_ft"].append(H[tanks].copy())
        # nexthyd
        if relerr > hacc and m.extra_iter == -1:
            halt = True
        if halt:
            htime = dur
        tstep = 0
        if htime < dur:
            tstep = hstep
            nb = (htime + pstart) // pstep + 1
            t = nb * pstep - htime
            if 0 < t < tstep:
                tstep = t
            t = rtime - htime
            if 0 < t < tstep:
                tstep = t
            # tanktimestep
            for j, n in enumerate(tanks):
                qn = fixed_dem[n]
                if abs(qn) <= QZERO:
                    continue
                if qn > 0.0 and H[n] < hmax[j]:
                    v = vmax[j] - V[j]
                elif qn < 0.0 and H[n] > hmin[j]:
                    v = vmin[j] - V[j]
                else:
                    continue
                t = _round_c(v / qn)
                if 0 < t < tstep:
                    tstep = t
            # controltimestep
            for i in range(len(net.ctl_link)):
                t = 0
                n, ctype = int(net.ctl_node[i]), int(net.ctl_type[i])
                if n >= 0:
                    j = tanks.index(n)
                    qn = fixed_dem[n]
                    if abs(qn) > QZERO:
                        grade = float(net.ctl_grade[i])
                        if (H[n] < grade and ctype == 1 and qn > 0.0) or \
                                (H[n] > grade and ctype == 0 and qn < 0.0):
                            t = _round_c((tank_vol(j, grade) - V[j]) / qn)
                if ctype == 2 and int(net.ctl_time[i]) > htime:
                    t = int(net.ctl_time[i]) - htime
                if ctype == 3:
                    t1 = (htime + tstart) % SEC_PER_DAY
                    t2 = int(net.ctl_time[i])
                    t = t2 - t1 if t2 >= t1 else SEC_PER_DAY - t1 + t2
                if 0 < t < tstep and S[int(net.ctl_link[i])] != int(net.ctl_status[i]):
                    tstep = t
            # tanklevels
            for j, n in enumerate(tanks):
                V[j] += fixed_dem[n] * tstep
                if V[j] + fixed_dem[n] >= vmax[j]:
                    V[j] = vmax[j]
                elif V[j] - fixed_dem[n] <= vmin[j]:
                    V[j] = vmin[j]
                H[n] = hmin[j] + (V[j] - vmin[j]) / area[j]
            htime += tstep
            if htime >= rtime:
                rtime += rstep
        else:
            htime += 1
        if tstep == 0:
            break
        if len(frames["t_sec"]) >= max_frames:
            raise RuntimeError("too many frames")
    return {k: np.asarray(v) for k, v in frames.items()}
