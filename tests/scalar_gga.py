# -*- coding: utf-8 -*-
"""Brute-force scalar GGA for networks of a few links.

Written for the tests from the EPANET 2.2 equations (hydcoeffs.c pipe,
pump and emitter coefficients; hydsolver.c hydsolve / newflows; hydstatus.c
linkstatus, cvstatus, pumpstatus, tankstatus) as the reference quotes them
(dgga/solver.py:697-739, 2176-2326, 2368-2377, 3092-3179), one scenario,
Python floats and loops, one dense numpy solve per iteration.  It shares no
code with oracle/: it exists to check the oracle's iteration counts, status
sequences and branch formulas independently.

Network description (EPANET internal units, ft and cfs):
  nodes: list of dicts kind 'j' (elev, demand, ke), 'r' (head) or
         't' (head, hmin, hmax, overflow)
  links: list of dicts n1, n2, kind 'pipe' | 'cv' | 'pump', r (H-W
         resistance), diam (ft), and for pumps ptype 'chp' | 'pf' | 'lin' |
         'nocurve' with h0, r, n, q0, hmax (shutoff head for pumpstatus)
"""
import math

import numpy as np

CSMALL, CBIG, RQTOL, QZERO, BIG = 1e-6, 1e8, 1e-7, 1e-6, 1e10
HTOL, QTOL = 0.0005, 0.0001
XHEAD, TEMPCLOSED, CLOSED, OPEN = 0, 1, 2, 3
PI = 3.141592654


def cvstatus(s, dh, q):
    """hydstatus.c cvstatus: a big head difference decides by its sign (and
    reverse flow); a small one closes only on reverse flow."""
    if abs(dh) > HTOL:
        return CLOSED if (dh < -HTOL or q < -QTOL) else OPEN
    return CLOSED if q < -QTOL else s


def link_coeffs(lk, q, status):
    """(P, Y) of one link at flow q and status."""
    if status <= CLOSED:
        return 1.0 / CBIG, q
    if lk["kind"] in ("pipe", "cv"):
        qa = abs(q)
        hg = 1.852 * lk["r"] * max(qa, 1e-30) ** 0.852
        if hg < RQTOL:
            hg = RQTOL
            hl = hg * q
        else:
            hl = lk["r"] * qa ** 1.852 * (1.0 if q >= 0 else -1.0)
        return 1.0 / hg, hl / hg
    t = lk["ptype"]
    if t == "nocurve":
        return 1.0 / CSMALL, q
    qa = abs(q)
    if t == "chp":
        # hloss = r / q with r < 0 (a head gain of 8.814 P / q); hgrad = -r / q^2
        hg = -lk["r"] / max(qa, 1e-30) ** 2
        if hg > CBIG or hg < RQTOL:
            hg = CBIG if hg > CBIG else RQTOL
            return 1.0 / hg, -q
        return 1.0 / hg, (lk["r"] / q) / hg
    if t == "lin":
        return 1.0 / lk["r"], (lk["h0"] + lk["r"] * q) / lk["r"]
    n = lk["n"]
    hg = n * lk["r"] * (qa ** (n - 1.0) if qa > 0.0 else 0.0)
    if hg < RQTOL:
        hg = RQTOL
        return 1.0 / hg, (lk["h0"] + hg * q) / hg
    return 1.0 / hg, (lk["h0"] + lk["r"] * qa ** (n - 1.0) * q) / hg


def emitter(e, ke, qexp=2.0):
    hg = qexp * max(ke, CSMALL) * max(abs(e), 1e-30) ** (qexp - 1.0)
    if hg < RQTOL:
        return RQTOL * e, RQTOL
    return hg * e / qexp, hg


def linkstatus(nodes, links, S, H, q):
    """hydstatus.c linkstatus: reopen XHEAD/TEMPCLOSED, cvstatus, pumpstatus,
    then tankstatus; returns the new statuses."""
    S = list(S)
    for k, lk in enumerate(links):
        s = OPEN if S[k] in (XHEAD, TEMPCLOSED) else S[k]
        dh = H[lk["n1"]] - H[lk["n2"]]
        if lk["kind"] == "cv":
            s = cvstatus(s, dh, q[k])
        if lk["kind"] == "pump" and s >= OPEN:
            s = XHEAD if -dh > lk["hmax"] + HTOL else OPEN
        S[k] = s
    for k, lk in enumerate(links):
        t = next((n for n in (lk["n1"], lk["n2"]) if nodes[n]["kind"] in ("r", "t")), None)
        if t is None or nodes[t]["kind"] != "t" or S[k] <= CLOSED:
            continue
        other = lk["n2"] if t == lk["n1"] else lk["n1"]
        h = H[t] - H[other]
        q_out = q[k] if t == lk["n1"] else -q[k]   # flow out of the tank
        nt = nodes[t]
        full = H[t] >= nt["hmax"] - HTOL and not nt.get("overflow", False)
        empty = H[t] <= nt["hmin"] + HTOL
        if lk["kind"] == "pump":
            if (full and lk["n2"] == t) or (empty and lk["n1"] == t):
                S[k] = TEMPCLOSED
        elif (full and cvstatus(OPEN, h, q_out) == CLOSED) or \
                (empty and cvstatus(CLOSED, h, q_out) == OPEN):
            S[k] = TEMPCLOSED
    return S


def solve(nodes, links, accuracy, max_iter=200, checkfreq=2, maxcheck=10, qexp=2.0,
          status0=None):
    """hydsolve for one scenario.  Returns dict(iters, status, head, flow,
    emitter, events) with events = [(iteration, link, old, new, branch)],
    branch 'conv' (convergence check) or 'per' (CheckFreq cadence); iters is
    maxtrials + 1 when unconverged."""
    jn = [i for i, n in enumerate(nodes) if n["kind"] == "j"]
    row = {n: r for r, n in enumerate(jn)}
    H = [0.0 if n["kind"] == "j" else n["head"] for n in nodes]
    q = []
    S = list(status0) if status0 is not None else [OPEN] * len(links)
    for k, lk in enumerate(links):
        if lk["kind"] == "pump":
            q.append(lk.get("q0", 1.0))
        else:
            q.append(PI * lk["diam"] ** 2 / 4.0)
        if S[k] <= CLOSED:
            q[k] = QZERO
    e = {i: 1.0 for i in jn if nodes[i].get("ke", 0.0) > 0.0}
    nextcheck = checkfreq
    events = []
    for it in range(1, max_iter + 1):
        nj = len(jn)
        A = np.zeros((nj, nj))
        F = [0.0] * nj
        PY = [link_coeffs(lk, q[k], S[k]) for k, lk in enumerate(links)]
        for k, lk in enumerate(links):
            P, Y = PY[k]
            a, b = lk["n1"], lk["n2"]
            if a in row:
                A[row[a], row[a]] += P
                F[row[a]] += Y - q[k]
            if b in row:
                A[row[b], row[b]] += P
                F[row[b]] += -Y + q[k]
            if a in row and b in row:
                A[row[a], row[b]] -= P
                A[row[b], row[a]] -= P
            elif b in row:
                F[row[b]] += P * H[a]
            elif a in row:
                F[row[a]] += P * H[b]
        em = {}
        for i, ei in e.items():
            hl, hg = emitter(ei, nodes[i]["ke"], qexp)
            em[i] = (hl, hg)
            A[row[i], row[i]] += 1.0 / hg
            F[row[i]] += (hl + nodes[i]["elev"]) / hg - ei
        for i in jn:
            F[row[i]] -= nodes[i]["demand"]
        x = np.linalg.solve(A, np.asarray(F))
        for i in jn:
            H[i] = float(x[row[i]])
        dqs, qs = [], []
        for k, lk in enumerate(links):
            P, Y = PY[k]
            dq = Y - P * (H[lk["n1"]] - H[lk["n2"]])
            if lk["kind"] == "pump" and lk["ptype"] == "chp" and dq > q[k]:
                dq = q[k] / 2.0
            q[k] -= dq
            dqs.append(abs(dq))
            qs.append(abs(q[k]))
        for i in e:
            hl, hg = em[i]
            dq = (hl - (H[i] - nodes[i]["elev"])) / hg
            e[i] -= dq
            dqs.append(abs(dq))
            qs.append(abs(e[i]))
        dqsum, qsum = math.fsum(dqs), math.fsum(qs)
        relerr = dqsum / qsum if qsum > accuracy else dqsum
        if relerr <= accuracy:
            new = linkstatus(nodes, links, S, H, q)
            chg = [k for k in range(len(links)) if new[k] != S[k]]
            events += [(it, k, S[k], new[k], "conv") for k in chg]
            S = new
            if not chg:
                return dict(iters=it, status=S, head=H, flow=q, emitter=e, events=events,
                            relerr=relerr)
            nextcheck = it + checkfreq
        elif it <= maxcheck and it == nextcheck:
            new = linkstatus(nodes, links, S, H, q)
            events += [(it, k, S[k], new[k], "per") for k in range(len(links))
                       if new[k] != S[k]]
            S = new
            nextcheck += checkfreq
    return dict(iters=max_iter + 1, status=S, head=H, flow=q, emitter=e, events=events,
                relerr=relerr)


def to_net(nodes, links, accuracy=1e-3, checkfreq=2, maxcheck=10, trials=200):
    """The same network as a wdsgpu_amd Net (CFS units, Hazen-Williams)."""
    from wdsgpu_amd.network import (Net, JUNCTION, RESERVOIR, TANK, PIPE, CVPIPE, PUMP,
                                    PUMP_CONST_HP, PUMP_POWER_FUNC, PUMP_NOCURVE)
    N, L = len(nodes), len(links)
    nt = np.array([{"j": JUNCTION, "r": RESERVOIR, "t": TANK}[n["kind"]] for n in nodes],
                  dtype=np.int8)
    elev = np.array([n.get("elev", n.get("head", 0.0)) if n["kind"] != "t"
                     else n["hmin"] - 1.0 for n in nodes])
    lt = np.array([{"pipe": PIPE, "cv": CVPIPE, "pump": PUMP}[lk["kind"]] for lk in links],
                  dtype=np.int8)
    net = Net(node_id=["N%d" % i for i in range(N)], node_type=nt, elev_ft=elev,
              node_ke=np.array([n.get("ke", 0.0) for n in nodes]),
              demand_cfs=np.array([n.get("demand", 0.0) for n in nodes]),
              link_id=["L%d" % k for k in range(L)], link_type=lt,
              link_n1=np.array([lk["n1"] for lk in links], dtype=np.int32),
              link_n2=np.array([lk["n2"] for lk in links], dtype=np.int32),
              diam_ft=np.array([lk.get("diam", 0.0) for lk in links]),
              len_ft=np.array([lk.get("len", 0.0) for lk in links]),
              roughness=np.array([lk.get("C", 0.0) for lk in links]),
              km_int=np.zeros(L), r_hw=np.array([lk.get("r", 0.0) if lk["kind"] != "pump"
                                                 else 0.0 for lk in links]),
              init_status=np.ones(L, dtype=np.int8),
              valve_setting_user=np.array([1.0 if lk["kind"] == "pump" else 0.0
                                           for lk in links]),
              meta=dict(flow_units="CFS", headloss="H-W", accuracy=accuracy, trials=trials,
                        checkfreq=checkfreq, maxcheck=maxcheck))
    tanks = [i for i, n in enumerate(nodes) if n["kind"] == "t"]
    if tanks:
        net.tank_node = np.array(tanks, dtype=np.int32)
        net.tank_h0 = np.array([nodes[i]["head"] for i in tanks])
        net.tank_hmin = np.array([nodes[i]["hmin"] for i in tanks])
        net.tank_hmax = np.array([nodes[i]["hmax"] for i in tanks])
        net.tank_area = np.ones(len(tanks))
        net.tank_overflow = np.array([int(nodes[i].get("overflow", False)) for i in tanks],
                                     dtype=np.int8)
    pumps = [k for k, lk in enumerate(links) if lk["kind"] == "pump"]
    if pumps:
        code = {"chp": PUMP_CONST_HP, "pf": PUMP_POWER_FUNC, "lin": PUMP_POWER_FUNC,
                "nocurve": PUMP_NOCURVE}
        net.pump_link = np.array(pumps, dtype=np.int32)
        net.pump_ptype = np.array([code[links[k]["ptype"]] for k in pumps], dtype=np.int8)
        net.pump_h0 = np.array([links[k].get("h0", 0.0) for k in pumps])
        net.pump_r = np.array([links[k].get("r", 0.0) for k in pumps])
        net.pump_n = np.array([links[k].get("n", 1.0) for k in pumps])
        net.pump_q0 = np.array([links[k].get("q0", 1.0) for k in pumps])
        net.pump_qmax = np.full(len(pumps), BIG)
        net.pump_hmax = np.array([links[k].get("hmax", BIG) for k in pumps])
        net.pump_curve_ptr = np.zeros(len(pumps) + 1, dtype=np.int32)
    return net


def fixed_heads(nodes):
    return np.array([n["head"] if n["kind"] in ("r", "t") else 0.0 for n in nodes])


def hw_r(length, diam, C):
    """Hazen-Williams resistance in EPANET internal units."""
    return 4.727 * length / C ** 1.852 / diam ** 4.871
