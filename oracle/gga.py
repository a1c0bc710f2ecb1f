# -*- coding: utf-8 -*-
"""Oracle: float64 numpy restatement of dgga's batched dense GGA.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``GGAModel(net)`` restates the host precompute of ``GGASolver.__init__``
(dgga/solver.py:160-495, 1272-1390, 646-832, 950-1011) for
mode="dense", dense_status_machine=True; ``solve`` restates
``GGASolver.solve`` (dgga/solver.py:2703-3212) with
assemble="dense"/"csr", linear_solver="dense": batched link coefficients,
scatter assembly in the reference's source order, dense Cholesky with two
refinement steps, newflows/relerr, per-scenario freezing and the batched
status machine (PRV every iteration; CV/pump/tank on convergence or on the
CheckFreq cadence).  Extensions the reference's dense path refuses but its
scalar EPANET replica supports are restated from that replica:
Darcy-Weisbach (dgga/solver.py:1404-1450).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .detmath import det_log, det_pow

from wdsgpu_amd.network import (
    CSMALL, CBIG, BIG, TINY, MISSING, QZERO, HEXP_HW, PI_EPANET,
    DW_A1, DW_A2, DW_A8, DW_A9, DW_AB, DW_AC,
    CVPIPE, PIPE, PUMP, PRV, PSV, PBV, FCV, TCV,
    ST_XHEAD, ST_TEMPCLOSED, ST_CLOSED, ST_OPEN, ST_ACTIVE, ST_XFCV, ST_XPRESSURE,
    PUMP_CONST_HP, PUMP_POWER_FUNC, PUMP_CUSTOM, PUMP_NOCURVE, JUNCTION)


class GGAModel:
    """Static per-topology data (dgga/solver.py:160-495 and helpers)."""

    def __init__(self, net):
        self.net = net
        lt = np.asarray(net.link_type, dtype=np.int64)
        nt = np.asarray(net.node_type)
        allowed = {CVPIPE, PIPE, TCV, PUMP, PRV, PSV, FCV}
        bad = set(lt.tolist()) - allowed
        if bad:
            raise NotImplementedError(f"link types {sorted(bad)} not on the dense path")
        self.headloss = str(net.meta.get("headloss", "H-W")).upper()
        if self.headloss not in ("H-W", "D-W"):
            raise NotImplementedError(self.headloss)
        self.N, self.L = net.N, net.L
        is_junc = nt == JUNCTION
        self.junc_nodes = np.where(is_junc)[0]
        self.fixed_nodes = np.where(~is_junc)[0]
        self.Nj = int(is_junc.sum())
        Nj = self.Nj
        junc_row = np.full(self.N, -1, dtype=np.int64)
        junc_row[self.junc_nodes] = np.arange(Nj)
        self.junc_row = junc_row
        n1 = np.asarray(net.link_n1, dtype=np.int64)
        n2 = np.asarray(net.link_n2, dtype=np.int64)
        self.n1, self.n2 = n1, n2
        self.lt = lt
        self.is_tcv = lt == TCV
        self.closed = np.asarray(net.init_status) == 0
        self.r_hw = np.asarray(net.r_hw, dtype=np.float64)
        self.km_pipe = np.asarray(net.km_int, dtype=np.float64)
        self.diam = np.asarray(net.diam_ft, dtype=np.float64)
        self.kc = np.asarray(net.roughness, dtype=np.float64)
        st = np.asarray(net.init_status)
        setting = np.asarray(net.valve_setting_user, dtype=np.float64)
        d4 = self.diam ** 2 * self.diam ** 2
        self.km_tcv = np.where(st == 2, np.divide(0.02517 * setting, d4,
                                                  out=np.zeros_like(setting),
                                                  where=d4 > 0), self.km_pipe)
        self.km_valve_ml = self.km_pipe.copy()
        # ---- scatter maps (dgga/solver.py:283-334) ----
        j1, j2 = junc_row[n1], junc_row[n2]
        m1, m2 = is_junc[n1], is_junc[n2]
        both = m1 & m2
        idx_off1 = j1[both] * Nj + j2[both]
        idx_off2 = j2[both] * Nj + j1[both]
        lo = np.minimum(j1[both], j2[both])
        hi = np.maximum(j1[both], j2[both])
        pkey = lo * Nj + hi
        uq, inv, cnt = np.unique(pkey, return_inverse=True, return_counts=True)
        n_fwd = np.zeros(uq.size, dtype=np.int64)
        np.add.at(n_fwd, inv, (j1[both] == lo).astype(np.int64))
        trig = ((cnt >= 3) & (n_fwd != 0) & (n_fwd != cnt))[inv]
        idx_off1 = np.where(trig, lo * Nj + hi, idx_off1)
        idx_off2 = np.where(trig, hi * Nj + lo, idx_off2)
        idx_d1 = j1[m1] * Nj + j1[m1]
        idx_d2 = j2[m2] * Nj + j2[m2]
        self.A_idx = np.concatenate([idx_off1, idx_off2, idx_d1, idx_d2])
        self.lk_both = np.where(both)[0]
        self.lk_m1 = np.where(m1)[0]
        self.lk_m2 = np.where(m2)[0]
        self.f_idx1 = j1[m1]
        self.f_idx2 = j2[m2]
        g1 = (~m1) & m2
        g2 = (~m2) & m1
        self.lk_g1, self.g1_row, self.g1_src = np.where(g1)[0], j2[g1], n1[g1]
        self.lk_g2, self.g2_row, self.g2_src = np.where(g2)[0], j1[g2], n2[g2]
        self.m1, self.m2, self.j1, self.j2 = m1, m2, j1, j2
        # CSR pattern (dgga/solver.py:336-369)
        diag_flat = np.arange(Nj, dtype=np.int64) * (Nj + 1)
        csr_pos = np.unique(np.concatenate([self.A_idx, diag_flat]))
        self.csr_pos = csr_pos
        self.csr_scatter = np.searchsorted(csr_pos, self.A_idx)
        self.csr_diag = np.searchsorted(csr_pos, diag_flat)
        self.csr_row = csr_pos // Nj
        self.csr_col = csr_pos % Nj
        self.el_junc = np.asarray(net.elev_ft, dtype=np.float64)[self.junc_nodes]
        self.elev = np.asarray(net.elev_ft, dtype=np.float64)
        self.node_ke = np.asarray(net.node_ke, dtype=np.float64)
        self.hexp = HEXP_HW if self.headloss == "H-W" else 2.0
        self.viscos = float(net.meta.get("viscosity", 1.1e-5))
        self.qexp = float(net.meta["qexp"])
        self.rqtol = float(net.meta.get("rqtol", 1e-7))
        self.max_iter_default = int(net.meta["trials"])
        self.hacc_default = min(max(float(net.meta["accuracy"]), 1e-5), 1e-1)
        # ---- status encoding, settings, tolerances (1272-1325) ----
        map_init = {0: ST_CLOSED, 1: ST_OPEN, 2: ST_ACTIVE}
        self.init_status_int = np.asarray([map_init[int(s)] for s in st], dtype=np.int8)
        self.is_valve = lt >= PRV
        ucf_p, ucf_q = net.ucf_pressure(), net.ucf_flow()
        self.ucf_flow = ucf_q
        self.ucf_head = 0.3048 if str(net.meta["flow_units"]).upper() in \
            ("LPS", "LPM", "MLD", "CMH", "CMD") else 1.0
        init_setting = setting.copy()
        for k in np.where(self.is_valve)[0]:
            if self.init_status_int[k] != ST_ACTIVE:
                init_setting[k] = MISSING
            elif lt[k] in (PRV, PSV, PBV):
                init_setting[k] = init_setting[k] / ucf_p
            elif lt[k] == FCV:
                init_setting[k] = init_setting[k] / ucf_q
        self.init_setting = init_setting
        self.htol = float(net.meta.get("htol", 0.0005))
        self.qtol = float(net.meta.get("qtol", 0.0001))
        self.checkfreq = int(net.meta.get("checkfreq", 2))
        self.maxcheck = int(net.meta.get("maxcheck", 10))
        self.extra_iter = int(net.meta.get("extra_iter", -1))
        # ---- pumps (1327-1348, 950-1011) ----
        self.is_pump = lt == PUMP
        pl = np.asarray(net.pump_link, dtype=np.int64)
        self.n_pumps = len(pl)
        self.pump_links = pl
        L = self.L
        self.pl_ptype = np.full(L, -1, dtype=np.int64)
        self.pl_h0 = np.zeros(L)
        self.pl_r = np.zeros(L)
        self.pl_n = np.zeros(L)
        self.pl_q0 = np.zeros(L)
        self.pl_hmax = np.zeros(L)
        if self.n_pumps:
            self.pl_ptype[pl] = np.asarray(net.pump_ptype)
            self.pl_h0[pl] = net.pump_h0
            self.pl_r[pl] = net.pump_r
            self.pl_n[pl] = net.pump_n
            self.pl_q0[pl] = net.pump_q0
            self.pl_hmax[pl] = net.pump_hmax
        # CUSTOM curves, points in user units (dgga/solver.py:1346-1370)
        self.pump_curve = {}
        cptr = np.asarray(net.pump_curve_ptr, dtype=np.int64)
        for j, k in enumerate(pl):
            if int(net.pump_ptype[j]) == PUMP_CUSTOM:
                a, b = int(cptr[j]), int(cptr[j + 1])
                if b - a < 2:
                    raise RuntimeError("CUSTOM pump without curve points")
                self.pump_curve[int(k)] = (np.asarray(net.pump_curve_q[a:b], dtype=np.float64),
                                           np.asarray(net.pump_curve_h[a:b], dtype=np.float64))
        sp = init_setting
        self.pump_speed = sp.copy()
        n_np = self.pl_n.copy()
        n_np[np.abs(n_np - 1.0) < TINY] = 1.0
        self.pump_n = n_np
        self.pump_lin = (n_np == 1.0) & self.is_pump
        self.pump_pf = self.pl_ptype == PUMP_POWER_FUNC
        self.pump_chp = self.pl_ptype == PUMP_CONST_HP
        self.pump_nocurve = self.pl_ptype == PUMP_NOCURVE
        self.closed_dense = self.closed | (self.is_pump & (sp == 0.0))
        # ---- tanks (1350-1368, 536-568) ----
        self.is_fixed = nt != JUNCTION
        tn = np.asarray(net.tank_node, dtype=np.int64)
        self.is_tank = np.zeros(self.N, dtype=bool)
        self.tn_hmax = np.zeros(self.N)
        self.tn_hmin = np.zeros(self.N)
        self.tn_overflow = np.zeros(self.N, dtype=bool)
        if tn.size:
            self.is_tank[tn] = True
            self.tn_hmax[tn] = net.tank_hmax
            self.tn_hmin[tn] = net.tank_hmin
            self.tn_overflow[tn] = np.asarray(net.tank_overflow) != 0
        fx = self.is_fixed
        use_n2 = (~fx[n1]) & fx[n2]
        cand = np.where(fx[n1], n1, np.where(fx[n2], n2, -1))
        sel = (cand >= 0) & self.is_tank[np.where(cand >= 0, cand, 0)]
        k = np.where(sel)[0]
        self.tl_k = k
        self.tl_tank = cand[k]
        self.tl_other = np.where(use_n2[k], n1[k], n2[k])
        self.tl_qsign = np.where(use_n2[k], -1.0, 1.0)
        self.tl_hmax = self.tn_hmax[self.tl_tank]
        self.tl_hmin = self.tn_hmin[self.tl_tank]
        self.tl_ovf = self.tn_overflow[self.tl_tank]
        self.tl_is_pump = self.is_pump[k]
        self.tl_pump_in = self.is_pump[k] & (n2[k] == self.tl_tank)
        self.tl_pump_out = self.is_pump[k] & (n1[k] == self.tl_tank)
        # ---- status-machine statics (668-682) ----
        self.is_cv = lt == CVPIPE
        self.sm_zero_speed = self.is_pump & (sp == 0.0)
        self.sm_pump_live = self.is_pump & (sp > 0.0)
        self.sm_pump_hmax = np.where(self.pump_chp, BIG, sp * sp * self.pl_hmax)
        # ---- control valves: PRVs (745-832), PSVs and FCVs (the scalar
        # replica's valvecoeffs, 1616-1688), in valve (file) order ----
        prv_all = np.where((lt == PRV) | (lt == PSV) | (lt == FCV))[0]
        pcv = prv_all[init_setting[prv_all] != MISSING]
        fixedv = prv_all[init_setting[prv_all] == MISSING]
        self.n_pcv = int(pcv.size)
        self.prv_fixed = np.zeros(L, dtype=bool)
        self.prv_fixed[fixedv] = True
        self.pcv = np.zeros(L, dtype=bool)
        self.pcv[pcv] = True
        self.prv_k = pcv
        if pcv.size:
            pn1, pn2 = n1[pcv], n2[pcv]
            if self.is_fixed[pn1].any() or self.is_fixed[pn2].any():
                raise NotImplementedError("PRV with a fixed-head end")
            self.prv_n1, self.prv_n2 = pn1, pn2
            self.prv_j1, self.prv_j2 = junc_row[pn1], junc_row[pn2]
            # 0 PRV, 1 PSV, 2 FCV; hset: PRV El[n2] + setting (prvcoeff),
            # PSV El[n1] + setting (psvcoeff), FCV the flow setting (fcvcoeff)
            self.prv_kind = np.where(lt[pcv] == PRV, 0, np.where(lt[pcv] == PSV, 1, 2))
            self.prv_hset = np.where(self.prv_kind == 0, self.elev[pn2] + init_setting[pcv],
                                     np.where(self.prv_kind == 1,
                                              self.elev[pn1] + init_setting[pcv],
                                              init_setting[pcv]))
            self.prv_km = self.km_valve_ml[pcv].copy()
            pj1, pj2 = self.prv_j1, self.prv_j2
            # per valve, in the order the terms are added: the off-diagonal
            # pair, d1 (+ P, then + CBIG of an ACTIVE PSV), d2 (+ P, then
            # + CBIG of an ACTIVE PRV)
            self.prv_A_idx = np.stack([pj1 * Nj + pj2, pj2 * Nj + pj1, pj1 * Nj + pj1,
                                       pj1 * Nj + pj1, pj2 * Nj + pj2, pj2 * Nj + pj2],
                                      axis=1).reshape(-1)
            self.prv_A_csr = np.searchsorted(csr_pos, self.prv_A_idx)
            self.prv_F_idx = np.stack([pj1, pj2], axis=1).reshape(-1)

    # ------------------------------------------------------------- K1 pieces
    def init_flow(self):
        """dgga/solver.py:2163-2174."""
        q = PI_EPANET * self.diam ** 2 / 4.0
        if self.n_pumps:
            q = np.where(self.is_pump, self.init_setting * self.pl_q0, q)
        return np.where(self.closed, QZERO, q)

    # Power laws and logarithms use the deterministic pow / log that the HIP
    # kernels share (oracle/detmath.py): the reference's libm calls
    # (dgga/solver.py:58-69) agree with it to one ulp, and the two paths of
    # this build then compute identical coefficients.
    # When set to a numpy Generator, every power-law and logarithm evaluation
    # (pipe, Darcy-Weisbach friction, pump, emitter) is perturbed by -1, 0 or
    # +1 ulp at random: the oracle's own sensitivity to one
    # rounding of the transcendental (the reference's self-jitter acceptance
    # practice, data/sparse_gpu_plan.md:404-445).
    pow_jitter = None
    # When set to a numpy Generator, every linear solve's heads are perturbed
    # by -1, 0 or +1 ulp at random: the oracle's own sensitivity to one
    # rounding of each solve.  An intermediate Newton system with
    # kappa * eps near 1 (a junction held only through closed links, whose
    # conductance is 1/CBIG) has no reproducible solution in float64: the
    # final heads then move by far more than a rounding under such a
    # perturbation, and a comparison is held to a multiple of that motion.
    solve_jitter = None

    def _pow(self, x, y):
        p = det_pow(x, y)
        if self.pow_jitter is not None:
            u = self.pow_jitter.integers(-1, 2, size=np.shape(p))
            p = p * (1.0 + u * 2.0 ** -52)
        return p

    def _log(self, x):
        p = det_log(x)
        if self.pow_jitter is not None:
            u = self.pow_jitter.integers(-1, 2, size=np.shape(p))
            p = p * (1.0 + u * 2.0 ** -52)
        return p

    def pipe_PY(self, q, r_hw):
        """H-W pipe coefficients, dgga/solver.py:2176-2201."""
        sgn = np.where(q < 0, -1.0, 1.0)
        qa = q * sgn
        hgrad = self.hexp * r_hw * self._pow(np.maximum(qa, 1e-30), self.hexp - 1.0)
        lin = hgrad < self.rqtol
        hgrad = np.where(lin, self.rqtol, hgrad)
        hloss = np.where(lin, hgrad * qa, hgrad * qa / self.hexp)
        ml = self.km_pipe > 0.0
        hloss = np.where(ml, hloss + self.km_pipe * qa * qa, hloss)
        hgrad = np.where(ml, hgrad + 2.0 * self.km_pipe * qa, hgrad)
        hloss = hloss * sgn
        return 1.0 / hgrad, hloss / hgrad

    def dw_PY(self, q, r_hw):
        """D-W pipe coefficients, vectorised restatement of
        dgga/solver.py:1404-1450 (laminar, Swamee-Jain, Dunlop cubic).
        Returns (P, Y, hloss, hgrad, dhloss/dr) on pipes (zeros elsewhere)."""
        pipe = self.lt <= PIPE
        qa = np.abs(q)
        s = self.viscos * self.diam
        e = np.divide(self.kc, self.diam, out=np.zeros_like(self.kc), where=self.diam > 0)
        lam = qa <= DW_A2 * s
        r16 = 16.0 * PI_EPANET * s
        hl_l = q * (r16 * r_hw + self.km_pipe * qa)
        hg_l = r16 * r_hw + 2.0 * self.km_pipe * qa
        dr_l = r16 * q
        with np.errstate(divide="ignore", invalid="ignore"):
            w = qa / np.where(s > 0, s, 1.0)
            sj = w >= DW_A1
            y1 = DW_A8 / self._pow(np.where(sj, w, 1.0), 0.9)
            y2 = e / 3.7 + y1
            y3 = DW_A9 * self._log(y2)
            f_sj = 1.0 / (y3 * y3)
            dfdq_sj = 1.8 * f_sj * y1 * DW_A9 / y2 / y3 / np.where(qa > 0, qa, 1.0)
            y2d = e / 3.7 + DW_AB
            y3d = DW_A9 * self._log(y2d)
            fa = 1.0 / (y3d * y3d)
            fb = (2.0 + DW_AC / (y2d * y3d)) * fa
            r2 = w / DW_A2
            x1 = 7.0 * fa - fb
            x2 = 0.128 - 17.0 * fa + 2.5 * fb
            x3 = -0.128 + 13.0 * fa - (fb + fb)
            x4 = 0.032 - 3.0 * fa + 0.5 * fb
            f_d = x1 + r2 * (x2 + r2 * (x3 + r2 * x4))
            dfdq_d = (x2 + r2 * (2.0 * x3 + r2 * 3.0 * x4)) / np.where(s > 0, s, 1.0) / DW_A2
        f = np.where(sj, f_sj, f_d)
        dfdq = np.where(sj, dfdq_sj, dfdq_d)
        r1 = f * r_hw + self.km_pipe
        hl_t = r1 * qa * q
        hg_t = (2.0 * r1 * qa) + (dfdq * r_hw * qa * qa)
        dr_t = f * qa * q
        hl = np.where(lam, hl_l, hl_t)
        hg = np.where(lam, hg_l, hg_t)
        dr = np.where(lam, dr_l, dr_t)
        hl = np.where(pipe, hl, 0.0)
        hg = np.where(pipe, hg, 1.0)
        dr = np.where(pipe, dr, 0.0)
        return 1.0 / hg, hl / hg, hl, hg, dr

    def tcv_PY(self, q):
        """dgga/solver.py:2203-2219."""
        km = self.km_tcv
        qa = np.abs(q)
        hgrad = 2.0 * km * qa
        lin = hgrad < self.rqtol
        hgrad = np.where(lin, self.rqtol, hgrad)
        hloss = np.where(lin, q * hgrad, q * hgrad / 2.0)
        P = np.where(km > 0.0, 1.0 / hgrad, 1.0 / CSMALL)
        Y = np.where(km > 0.0, hloss / hgrad, q)
        return P, Y

    def curvecoeff(self, k, q):
        """Local linear model of pump link k's CUSTOM curve at flow q (cfs,
        already divided by the speed): dgga/solver.py:1452-1474 (EPANET
        curvecoeff, hydcoeffs.c:794-831).  The right end of the segment is
        the first point with x >= q; below the first point the first segment
        and above the last point the last segment extend.  Returns (h0, r)
        in internal units."""
        q = q * self.ucf_flow
        x, y = self.pump_curve[k]
        npts = len(x)
        k2 = 0
        while k2 < npts and x[k2] < q:
            k2 += 1
        if k2 == 0:
            k2 += 1
        elif k2 == npts:
            k2 -= 1
        k1 = k2 - 1
        r = (y[k2] - y[k1]) / (x[k2] - x[k1])
        h0 = y[k1] - r * x[k1]
        h0 = h0 / self.ucf_head
        r = r * self.ucf_flow / self.ucf_head
        return h0, r

    def custom_hl_hg(self, k, q):
        """hloss, hgrad of CUSTOM pump link k at flows q (any shape):
        H0 = -h0, R = -r, N = 1, hgrad = R w, hloss = H0 w^2 + hgrad q
        (dgga/solver.py:1548-1558, hydcoeffs.c:716-733)."""
        sp = max(float(self.pump_speed[k]), 1e-12)
        qf = np.asarray(q, dtype=np.float64)
        hl = np.empty_like(qf)
        hg = np.empty_like(qf)
        for i, qk in np.ndenumerate(qf):
            h0c, rc = self.curvecoeff(k, abs(float(qk)) / sp)
            g = -rc * sp
            hg[i] = g
            hl[i] = -h0c * (sp * sp) + g * float(qk)
        return hl, hg

    def pump_PY(self, q):
        """Power-function / linear / constant-HP / CUSTOM / no-curve pumps,
        dgga/solver.py:2221-2326."""
        sgn = np.where(q < 0, -1.0, 1.0)
        qa = q * sgn
        spc = np.maximum(self.pump_speed, 1e-12)
        h0e = spc * spc * self.pl_h0
        re = self.pl_r * self._pow(spc, 2.0 - self.pump_n)
        qa_c = np.maximum(qa, 1e-30)
        hg_chp_raw = -re / qa_c / qa_c
        big_c = hg_chp_raw > CBIG
        sml_c = hg_chp_raw < self.rqtol
        hg_chp = np.where(big_c, CBIG, np.where(sml_c, self.rqtol, hg_chp_raw))
        q_c = sgn * qa_c
        hl_chp = np.where(big_c | sml_c, -hg_chp * q, re / q_c)
        q_zero = qa == 0.0
        nm1 = self.pump_n - 1.0
        pw0 = np.where(nm1 > 0.0, 0.0, np.where((nm1 < 0.0) & self.pump_pf, np.inf, 1.0))
        qa_pf = np.where(q_zero, 1.0, qa)
        hg_pf_raw = self.pump_n * re * np.where(q_zero, pw0, self._pow(qa_pf, nm1))
        sml_p = hg_pf_raw < self.rqtol
        hg_pf = np.where(sml_p, self.rqtol, hg_pf_raw)
        hl_pf = np.where(sml_p, h0e + hg_pf * q, h0e + hg_pf * q / np.maximum(self.pump_n, 1e-30))
        hg_pf = np.where(self.pump_lin, re, hg_pf)
        hl_pf = np.where(self.pump_lin, h0e + re * q, hl_pf)
        hg_p = np.where(self.pump_chp, hg_chp, hg_pf)
        hl_p = np.where(self.pump_chp, hl_chp, hl_pf)
        if self.pump_curve:
            hg_p = np.array(hg_p, dtype=np.float64)
            hl_p = np.array(hl_p, dtype=np.float64)
            for k in self.pump_curve:
                hl_c, hg_c = self.custom_hl_hg(k, q[..., k])
                hl_p[..., k] = hl_c
                hg_p[..., k] = hg_c
        with np.errstate(divide="ignore", invalid="ignore"):
            P = 1.0 / hg_p
            Y = hl_p / hg_p
        P = np.where(self.pump_nocurve, 1.0 / CSMALL, P)
        Y = np.where(self.pump_nocurve, q, Y)
        return P, Y

    def valve_PY_ml(self, q):
        """dgga/solver.py:834-848 (fixed PRVs, setting == MISSING)."""
        km = self.km_valve_ml
        qa = np.abs(q)
        hgrad = 2.0 * km * qa
        lin = hgrad < self.rqtol
        hgrad = np.where(lin, self.rqtol, hgrad)
        hloss = np.where(lin, q * hgrad, q * hgrad / 2.0)
        P = np.where(km > 0.0, 1.0 / np.where(km > 0, hgrad, 1.0), 1.0 / CSMALL)
        Y = np.where(km > 0.0, hloss / hgrad, q)
        return P, Y

    def emitter_hloss(self, e, ke):
        """dgga/solver.py:2368-2377."""
        ke_adj = np.maximum(ke, CSMALL)
        hgrad = self.qexp * ke_adj * self._pow(np.maximum(np.abs(e), 1e-30), self.qexp - 1.0)
        lin = hgrad < self.rqtol
        hgrad = np.where(lin, self.rqtol, hgrad)
        hloss = np.where(lin, hgrad * e, hgrad * e / self.qexp)
        return hloss, hgrad

    def link_PY(self, q, S, r_hw):
        """headlosscoeffs for every link (dgga/solver.py:2924-2956).
        Returns (P, Y) with pcv PRV positions zeroed."""
        if self.headloss == "D-W":
            P, Y = self.dw_PY(q, r_hw)[:2]
        else:
            P, Y = self.pipe_PY(q, r_hw)
        Pt, Yt = self.tcv_PY(q)
        P = np.where(self.is_tcv, Pt, P)
        Y = np.where(self.is_tcv, Yt, Y)
        if self.n_pumps:
            Pp, Yp = self.pump_PY(q)
            P = np.where(self.is_pump, Pp, P)
            Y = np.where(self.is_pump, Yp, Y)
        if self.prv_fixed.any():
            Pv, Yv = self.valve_PY_ml(q)
            P = np.where(self.prv_fixed, Pv, P)
            Y = np.where(self.prv_fixed, Yv, Y)
        closed_now = (S <= ST_CLOSED) | self.sm_zero_speed
        if self.n_pcv:
            closed_now = closed_now & ~self.pcv
        P = np.where(closed_now, 1.0 / CBIG, P)
        Y = np.where(closed_now, q, Y)
        if self.n_pcv:
            P = np.where(self.pcv, 0.0, P)
            Y = np.where(self.pcv, 0.0, Y)
        return P, Y

    # ------------------------------------------------------------- K2
    def assemble(self, P, Y, q, H, e_j, ke_j, d_j, S, csr=False):
        """linkcoeffs + emittercoeffs + nodecoeffs + valvecoeffs in the
        reference's order (dgga/solver.py:2958-3017).  Returns
        (A [B,Nj,Nj] or csr [B,nnz], F [B,Nj], Xflow [B,Nj], P, Y, em,
        hloss_e, hgrad_e)."""
        B = q.shape[0]
        Nj = self.Nj
        pl = (P != 0.0).astype(np.float64)
        Pm, Ym, qm = P * pl, Y * pl, q * pl
        Xflow = np.zeros((B, Nj))
        for b in range(B):
            np.add.at(Xflow[b], self.f_idx1, -qm[b, self.lk_m1])
            np.add.at(Xflow[b], self.f_idx2, qm[b, self.lk_m2])
        vals = np.concatenate([-Pm[:, self.lk_both], -Pm[:, self.lk_both],
                               Pm[:, self.lk_m1], Pm[:, self.lk_m2]], axis=1)
        if csr:
            A = np.zeros((B, self.csr_pos.size))
            for b in range(B):
                np.add.at(A[b], self.csr_scatter, vals[b])
        else:
            A = np.zeros((B, Nj * Nj))
            for b in range(B):
                np.add.at(A[b], self.A_idx, vals[b])
        F = np.zeros((B, Nj))
        for b in range(B):
            np.add.at(F[b], self.f_idx1, Ym[b, self.lk_m1])
            np.add.at(F[b], self.f_idx2, -Ym[b, self.lk_m2])
            if self.lk_g1.size:
                np.add.at(F[b], self.g1_row, Pm[b, self.lk_g1] * H[b, self.g1_src])
            if self.lk_g2.size:
                np.add.at(F[b], self.g2_row, Pm[b, self.lk_g2] * H[b, self.g2_src])
        hloss_e, hgrad_e = self.emitter_hloss(e_j, ke_j)
        em = (ke_j > 0.0).astype(np.float64)
        if csr:
            A[:, self.csr_diag] = A[:, self.csr_diag] + em / hgrad_e
        else:
            A3 = A.reshape(B, Nj, Nj)
            idx = np.arange(Nj)
            A3[:, idx, idx] = A3[:, idx, idx] + em / hgrad_e
        F = F + em * (hloss_e + self.el_junc) / hgrad_e
        Xflow = Xflow - em * e_j
        Xflow = Xflow - d_j
        F = F + Xflow
        if self.n_pcv:
            P, Y, F, A = self.prvcoeffs(P, Y, F, Xflow, q, S, A, csr)
        if not csr:
            A = A.reshape(B, Nj, Nj)
        return A, F, Xflow, P, Y, em, hloss_e, hgrad_e

    def prvcoeffs(self, P, Y, F, Xflow, q, S, A, csr):
        """Control-valve terms in valve order: prvcoeff (dgga/solver.py:850-912,
        batched), psvcoeff and fcvcoeff (the scalar replica's valvecoeffs,
        dgga/solver.py:1616-1688).  An ACTIVE FCV moves its setting out of
        Xflow[j1] into Xflow[j2] for the valves after it."""
        B = P.shape[0]
        a_cols, f_cols, p_cols, y_cols = [], [], [], []
        Xf = Xflow.copy()
        for i in range(self.n_pcv):
            ki = int(self.prv_k[i])
            kind = int(self.prv_kind[i])
            j1 = int(self.prv_j1[i])
            j2 = int(self.prv_j2[i])
            hset = float(self.prv_hset[i])
            qk = q[:, ki]
            sk = S[:, ki]
            m_act = sk == ST_ACTIVE
            m_cl = sk <= ST_CLOSED
            km = float(self.prv_km[i])
            qa = np.abs(qk)
            if km > 0.0:
                hgrad = 2.0 * km * qa
                lin = hgrad < self.rqtol
                hgrad = np.where(lin, self.rqtol, hgrad)
                hloss = np.where(lin, qk * hgrad, qk * hgrad / 2.0)
                pk_o = 1.0 / hgrad
                yk_o = hloss / hgrad
            else:
                pk_o = np.full_like(qk, 1.0 / CSMALL)
                yk_o = qk
            pk_na = np.where(m_cl, 1.0 / CBIG, pk_o)
            yk_na = np.where(m_cl, qk, yk_o)
            mna = (~m_act).astype(np.float64)
            mact = m_act.astype(np.float64)
            xf1 = Xf[:, j1].copy()
            xf2 = Xf[:, j2].copy()
            zero = np.zeros_like(qk)
            dy = (yk_na - qk) * mna
            if kind == 0:       # PRV: pins its downstream row
                pna = pk_na * mna
                a_cols += [-pna, -pna, pna, zero, pna, CBIG * mact]
                f1 = dy + np.where(m_act & (xf2 < 0.0), xf2, 0.0)
                f2 = -dy + hset * CBIG * mact
                y = np.where(m_act, qk + xf2, yk_na)
            elif kind == 1:     # PSV: pins its upstream row
                pna = pk_na * mna
                a_cols += [-pna, -pna, pna, CBIG * mact, pna, zero]
                f1 = dy + hset * CBIG * mact
                f2 = -dy + np.where(m_act & (xf1 > 0.0), xf1, 0.0)
                y = np.where(m_act, qk - xf1, yk_na)
            else:               # FCV: ACTIVE cuts the link, flow = setting
                pna = np.where(m_act, 1.0 / CBIG, pk_na)
                a_cols += [-pna, -pna, pna, zero, pna, zero]
                f1 = np.where(m_act, -hset, yk_na - qk)
                f2 = np.where(m_act, hset, -(yk_na - qk))
                y = np.where(m_act, qk - hset, yk_na)
                Xf[:, j1] = np.where(m_act, xf1 - hset, xf1)
                Xf[:, j2] = np.where(m_act, xf2 + hset, xf2)
            f_cols += [f1, f2]
            p_cols.append(pna)
            y_cols.append(y)
        a_vals = np.stack(a_cols, axis=1)
        f_vals = np.stack(f_cols, axis=1)
        P = P.copy()
        Y = Y.copy()
        P[:, self.prv_k] = np.stack(p_cols, axis=1)
        Y[:, self.prv_k] = np.stack(y_cols, axis=1)
        F = F.copy()
        for b in range(B):
            np.add.at(F[b], self.prv_F_idx, f_vals[b])
            if csr:
                np.add.at(A[b], self.prv_A_csr, a_vals[b])
            else:
                np.add.at(A[b], self.prv_A_idx, a_vals[b])
        return P, Y, F, A

    def csr_to_dense(self, data):
        B = data.shape[0]
        A = np.zeros((B, self.Nj * self.Nj))
        A[:, self.csr_pos] = data
        return A.reshape(B, self.Nj, self.Nj)

    # ------------------------------------------------------------- status
    def cvstatus_closed(self, s_is_open, dh, q):
        big = np.abs(dh) > self.htol
        closed_big = (dh < -self.htol) | (q < -self.qtol)
        rev = q < -self.qtol
        closed_small = rev if s_is_open else np.ones_like(rev)
        return np.where(big, closed_big, closed_small)

    def cvstatus(self, s, dh, q):
        big = np.abs(dh) > self.htol
        rev = q < -self.qtol
        new_big = np.where((dh < -self.htol) | rev, ST_CLOSED, ST_OPEN)
        new_small = np.where(rev, ST_CLOSED, s)
        return np.where(big, new_big, new_small).astype(np.int8)

    def linkstatus(self, S, H, q):
        """dgga/solver.py:697-739."""
        st0 = S
        s = np.where((st0 == ST_XHEAD) | (st0 == ST_TEMPCLOSED), ST_OPEN, st0).astype(np.int8)
        dh = H[:, self.n1] - H[:, self.n2]
        if self.is_cv.any():
            s = np.where(self.is_cv, self.cvstatus(s, dh, q), s).astype(np.int8)
        if self.n_pumps:
            gain = -dh
            xh = np.where(gain > self.sm_pump_hmax + self.htol, ST_XHEAD, ST_OPEN)
            s = np.where(self.sm_pump_live & (s >= ST_OPEN), xh, s).astype(np.int8)
        k = self.tl_k
        if k.size:
            live = s[:, k] > ST_CLOSED
            ht = H[:, self.tl_tank]
            dht = ht - H[:, self.tl_other]
            qk = q[:, k] * self.tl_qsign
            full = (ht >= self.tl_hmax - self.htol) & (~self.tl_ovf)
            empty = ht <= self.tl_hmin + self.htol
            cut_f = np.where(self.tl_is_pump, self.tl_pump_in,
                             self.cvstatus_closed(True, dht, qk))
            cut_e = np.where(self.tl_is_pump, self.tl_pump_out,
                             ~self.cvstatus_closed(False, dht, qk))
            hit = live & ((full & cut_f) | (empty & cut_e))
            s[:, k] = np.where(hit, ST_TEMPCLOSED, s[:, k])
        if self.n_pcv and (self.prv_kind == 2).any():
            # fcvstatus on the status before the reopening (the scalar
            # replica's _linkstatus 1875-1877, _fcvstatus 1819-1831)
            f = np.nonzero(self.prv_kind == 2)[0]
            kf = self.prv_k[f]
            s0 = st0[:, kf]
            h1 = H[:, self.prv_n1[f]]
            h2 = H[:, self.prv_n2[f]]
            qf = q[:, kf]
            new = np.where(h1 - h2 < -self.htol, ST_XFCV,
                           np.where(qf < -self.qtol, ST_XFCV,
                                    np.where((s0 == ST_XFCV) & (qf >= self.prv_hset[f]),
                                             ST_ACTIVE, s0)))
            s[:, kf] = new.astype(np.int8)
        return s, (s != st0).any(axis=1)

    def prvstatus(self, S, H, q):
        """valvestatus: prvstatus (dgga/solver.py:914-947) for PRVs and
        psvstatus (the scalar replica's 1740-1767) for PSVs; FCVs keep their
        status here (fcvstatus runs in linkstatus)."""
        k = self.prv_k
        h1 = H[:, self.prv_n1]
        h2 = H[:, self.prv_n2]
        qk = q[:, k]
        sk = S[:, k]
        hset = self.prv_hset
        hml = self.prv_km * (qk * qk)
        htol, qtol = self.htol, self.qtol
        neg = qk < -qtol
        c_act = np.where(neg, ST_CLOSED, np.where(h1 - hml < hset - htol, ST_OPEN, ST_ACTIVE))
        c_open = np.where(neg, ST_CLOSED, np.where(h2 >= hset + htol, ST_ACTIVE, ST_OPEN))
        c_cl = np.where((h1 >= hset + htol) & (h2 < hset - htol), ST_ACTIVE,
                        np.where((h1 < hset - htol) & (h1 > h2 + htol), ST_OPEN, ST_CLOSED))
        c_xp = np.where(neg, ST_CLOSED, sk)
        s_prv = np.where(sk == ST_ACTIVE, c_act,
                         np.where(sk == ST_OPEN, c_open,
                                  np.where(sk == ST_CLOSED, c_cl,
                                           np.where(sk == ST_XPRESSURE, c_xp, sk))))
        p_act = np.where(neg, ST_CLOSED, np.where(h2 + hml > hset + htol, ST_OPEN, ST_ACTIVE))
        p_open = np.where(neg, ST_CLOSED, np.where(h1 < hset - htol, ST_ACTIVE, ST_OPEN))
        p_cl = np.where((h2 > hset + htol) & (h1 > h2 + htol), ST_OPEN,
                        np.where((h1 >= hset + htol) & (h1 > h2 + htol), ST_ACTIVE, ST_CLOSED))
        s_psv = np.where(sk == ST_ACTIVE, p_act,
                         np.where(sk == ST_OPEN, p_open,
                                  np.where(sk == ST_CLOSED, p_cl,
                                           np.where(sk == ST_XPRESSURE, c_xp, sk))))
        kind = self.prv_kind
        s_new = np.where(kind == 0, s_prv, np.where(kind == 1, s_psv, sk))
        S_out = S.copy()
        S_out[:, k] = s_new
        return S_out.astype(np.int8), (s_new != sk).any(axis=1)

    # ------------------------------------------------------------- helpers
    def full_heads(self, Hj, rh_fix):
        B = Hj.shape[0]
        H = np.empty((B, self.N))
        H[:, self.junc_nodes] = Hj
        H[:, self.fixed_nodes] = rh_fix
        return H


def wave_tree_sum(x):
    """Row sums of x [B, n] in one fixed order: 64 partial sums, partial l
    adding x[:, l], x[:, l + 64], x[:, l + 128], ... in ascending order from
    +0.0, then an xor butterfly p_l <- p_l + p_(l ^ m) for m = 32, 16, ..., 1
    (every lane ends with the same value).  The reference sums with torch's
    pairwise reduction (dgga/solver.py:3104-3110) and its scalar replica
    with a left fold (2058-2062); this build fixes the wavefront tree above
    for both the HIP kernel and this oracle, so their relerr, and hence
    their iteration counts, agree bit for bit."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    B, n = x.shape
    p = np.zeros((B, 64))
    for c0 in range(0, n, 64):
        blk = x[:, c0:c0 + 64]
        p[:, :blk.shape[1]] = p[:, :blk.shape[1]] + blk
    lanes = np.arange(64)
    for m in (32, 16, 8, 4, 2, 1):
        p = p + p[:, lanes ^ m]
    return p[:, 0]


def _two_prod_v(a, b):
    """Exact product a*b = p + e (Veltkamp splitting, elementwise)."""
    p = a * b
    t = 134217729.0 * a
    ah = t - (t - a)
    al = a - ah
    t = 134217729.0 * b
    bh = t - (t - b)
    bl = b - bh
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def csr_residual_dd(m, vals, x, F):
    """F - A x for CSR values [B, nnz] (the model's CSR pattern), every
    product split exactly and every row summed in double-double, rounded
    once: the residual to about one rounding of its own size."""
    if not hasattr(m, "_ell"):
        rows = m.csr_row
        start = np.searchsorted(rows, np.arange(m.Nj))
        cnt = np.bincount(rows, minlength=m.Nj)
        ell = np.full((m.Nj, int(cnt.max())), -1, dtype=np.int64)
        for w in range(ell.shape[1]):
            has = cnt > w
            ell[has, w] = start[has] + w
        m._ell = ell
    ell = m._ell
    hi = F.copy()
    lo = np.zeros_like(F)
    xc = x[:, m.csr_col]
    p, e = _two_prod_v(vals, xc)
    for w in range(ell.shape[1]):
        idx = ell[:, w]
        ok = idx >= 0
        pw = np.where(ok, p[:, np.maximum(idx, 0)], 0.0)
        ew = np.where(ok, e[:, np.maximum(idx, 0)], 0.0)
        s = hi - pw
        bb = s - hi
        err = (hi - (s - bb)) + (-pw - bb)
        hi = s
        lo = lo + (err - ew)
    return hi + lo


def sparse_solve_refined(m, vals, F, active=None, n_refine=4):
    """A H = F per scenario with SuperLU on the CSR values [B, nnz], then
    n_refine steps of refinement with the double-double residual: the
    correctly rounded solution while kappa * eps < 1 (the same target as
    chol_solve_refined(..., extended=True)).  Raises RuntimeError naming
    singular samples.  Rows of inactive scenarios are returned as zeros."""
    import scipy.sparse as sps
    from scipy.sparse.linalg import splu
    B, Nj = F.shape
    H = np.zeros_like(F)
    todo = np.arange(B) if active is None else np.nonzero(active)[0]
    lus = {}
    bad = []
    for b in todo:
        A = sps.csc_matrix((vals[b], (m.csr_row, m.csr_col)), shape=(Nj, Nj))
        try:
            lus[b] = splu(A, permc_spec="MMD_AT_PLUS_A")
        except RuntimeError:
            bad.append(int(b))
            continue
        H[b] = lus[b].solve(F[b])
    if bad:
        raise RuntimeError("sparse factorisation failed: samples %s are singular" % bad)
    idx = np.asarray(sorted(lus))
    if idx.size == 0:
        return H
    for _ in range(n_refine):
        r = csr_residual_dd(m, vals[idx], H[idx], F[idx])
        for i, b in enumerate(idx):
            H[b] = H[b] + lus[b].solve(r[i])
    return H


def chol_solve_refined(A, F, n_refine=2, extended=False):
    """Dense Cholesky + cholesky_solve + n_refine refinement steps with the
    residual A*H summed along the last axis (dgga/solver.py:3037-3064).
    Raises RuntimeError naming non-SPD samples (3043-3051).

    extended=True evaluates the refinement residual in x87 extended precision
    (64-bit significand) and refines 4 times: the result is then the linear
    solution to about one rounding, which separates the reference's own
    rounding noise from a real discrepancy in a comparison."""
    B, Nj = F.shape
    H = np.empty_like(F)
    bad = []
    for b in range(B):
        try:
            c = sla.cho_factor(A[b], lower=True, check_finite=False)
        except np.linalg.LinAlgError:
            bad.append(b)
            continue
        x = sla.cho_solve(c, F[b], check_finite=False)
        if extended:
            Al = A[b].astype(np.longdouble)
            Fl = F[b].astype(np.longdouble)
            for _ in range(max(n_refine, 4)):
                r = (Fl - Al @ x.astype(np.longdouble)).astype(np.float64)
                x = x + sla.cho_solve(c, r, check_finite=False)
        else:
            for _ in range(n_refine):
                r = F[b] - (A[b] * x[None, :]).sum(-1)
                x = x + sla.cho_solve(c, r, check_finite=False)
        H[b] = x
    if bad:
        raise RuntimeError("dense Cholesky failed: samples %s are not SPD" % bad)
    return H


def solve(model, demand, res_head, ke=None, r_hw=None, q0=None, e0=None,
          max_iter=None, accuracy=None, status_machine=True, assemble="dense",
          n_refine=2, record=False, extended=False, status0=None,
          record_schedule=False, linsolve="dense"):
    """Batched dense GGA (dgga/solver.py:2703-3212, mode='dense',
    dense_status_machine=True).  demand/res_head/ke: [B,N] or [N];
    r_hw: [L] or [B,L].  Returns the reference's output dict (head_ft,
    flow_cfs, emitter_cfs, iters, relerr, converged, status, diag_ratio)
    plus, when ``record``, per-iteration traces.  ``extended`` replaces each
    linear solve by its correctly rounded solution (sparse_solve_refined),
    which separates the reference's own rounding noise from a real
    discrepancy in a comparison.
    ``status0`` ([L] or [B,L] internal codes) replaces the initial statuses
    (the warm start of an extended-period frame, as run_gga's status0 in the
    reference's EPANET-mode solve, dgga/solver.py:2855).  ``record_schedule``
    adds out["schedule"]: the status machine's events in the reference's
    layout (dgga/solver.py:3136-3160).  ``linsolve="sparse"`` solves each
    scenario's CSR system with SuperLU refined to the correctly rounded
    solution (sparse_solve_refined), the result the dense path gives with
    ``extended``, at a fraction of its cost on large networks; scenarios
    already converged are skipped (their heads are frozen anyway)."""
    m = model
    d = np.atleast_2d(np.asarray(demand, dtype=np.float64))
    rh = np.atleast_2d(np.asarray(res_head, dtype=np.float64))
    single = np.asarray(demand).ndim == 1
    B = d.shape[0]
    rh = np.broadcast_to(rh, (B, m.N))
    ke = m.node_ke if ke is None else np.asarray(ke, dtype=np.float64)
    ke = np.broadcast_to(np.atleast_2d(ke), (B, m.N))
    ke_j = ke[:, m.junc_nodes]
    has_em = ke_j > 0.0
    r = m.r_hw if r_hw is None else np.asarray(r_hw, dtype=np.float64)
    r = np.broadcast_to(np.atleast_2d(r), (B, m.L))
    max_iter = m.max_iter_default if max_iter is None else int(max_iter)
    hacc = m.hacc_default if accuracy is None else float(accuracy)
    if m.n_pcv and not status_machine:
        raise NotImplementedError("PRV networks need status_machine=True")
    if m.is_cv.any() and not status_machine:
        raise NotImplementedError("CV pipes need status_machine=True")
    q = np.broadcast_to(m.init_flow(), (B, m.L)).copy() if q0 is None else \
        np.broadcast_to(np.atleast_2d(q0), (B, m.L)).copy()
    if e0 is None:
        e_j = np.where(has_em, 1.0, 0.0)
    else:
        e_j = np.broadcast_to(np.atleast_2d(e0), (B, m.N))[:, m.junc_nodes].copy()
    rh_fix = rh[:, m.fixed_nodes]
    if np.isnan(rh_fix).any():
        raise ValueError("res_head_ft has nan at a fixed-head node")
    H_junc = np.zeros((B, m.Nj))
    H = m.full_heads(H_junc, rh_fix)
    d_j = d[:, m.junc_nodes]
    active = np.ones(B, dtype=bool)
    iters = np.zeros(B, dtype=np.int64)
    relerr_out = np.zeros(B)
    sm = bool(status_machine)
    maxtrials = max_iter
    S = np.broadcast_to(m.init_status_int if status0 is None else np.asarray(status0, np.int8),
                        (B, m.L)).copy()
    nextcheck = None
    diag_ratio = np.zeros(B) if m.n_pcv else None
    if sm:
        maxtrials = max_iter + (m.extra_iter if m.extra_iter > 0 else 0)
        nextcheck = np.full(B, m.checkfreq, dtype=np.int64)
    trace = dict(status=[], active=[], relerr=[]) if record else None
    sched = [] if (record_schedule and status_machine) else None
    if extended:
        # the correctly rounded solve: double-double residuals (an x87
        # extended residual is one ulp off on ~1 % of the heads here)
        linsolve = "sparse"
    csr = assemble == "csr" or linsolve == "sparse"
    for it in range(1, maxtrials + 1):
        P, Y = m.link_PY(q, S if sm else np.broadcast_to(
            np.where(m.closed_dense, ST_CLOSED, ST_OPEN), (B, m.L)), r)
        A, F, Xflow, P, Y, em, hloss_e, hgrad_e = m.assemble(
            P, Y, q, H, e_j, ke_j, d_j, S, csr=csr)
        if linsolve == "sparse":
            if m.n_pcv:
                dn = np.abs(A[:, m.csr_diag])
                dr_now = dn.max(1) / np.maximum(dn.min(1), 1e-300)
                diag_ratio = np.where(active, dr_now, diag_ratio)
            Hj = sparse_solve_refined(m, A, F, active)
        else:
            if csr:
                A = m.csr_to_dense(A)
            if m.n_pcv:
                dn = np.abs(np.diagonal(A, axis1=1, axis2=2))
                dr_now = dn.max(1) / np.maximum(dn.min(1), 1e-300)
                diag_ratio = np.where(active, dr_now, diag_ratio)
            Hj = chol_solve_refined(A, F, n_refine, extended)
        if m.solve_jitter is not None:
            u = m.solve_jitter.integers(-1, 2, size=Hj.shape)
            Hj = Hj * (1.0 + u * 2.0 ** -52)
        H_junc = np.where(active[:, None], Hj, H_junc)
        H = m.full_heads(H_junc, rh_fix)
        dh = H[:, m.n1] - H[:, m.n2]
        dq = Y - P * dh
        if m.n_pumps:
            dq = np.where(m.pump_chp & (dq > q), q / 2.0, dq)
        q_new = q - dq
        dh_e = H[:, m.junc_nodes] - m.el_junc
        dq_e = (hloss_e - dh_e) / hgrad_e
        e_new = e_j - em * dq_e
        dqsum = wave_tree_sum(np.abs(dq)) + wave_tree_sum(em * np.abs(dq_e))
        qsum = wave_tree_sum(np.abs(q_new)) + wave_tree_sum(em * np.abs(e_new))
        relerr = np.where(qsum > hacc, dqsum / qsum, dqsum)
        am = active[:, None]
        q = np.where(am, q_new, q)
        e_j = np.where(am, e_new, e_j)
        iters = np.where(active, it, iters)
        relerr_out = np.where(active, relerr, relerr_out)
        if not sm:
            active = active & (relerr > hacc)
            if record:
                trace["active"].append(active.copy())
                trace["relerr"].append(relerr.copy())
            if not active.any():
                break
            continue
        if m.n_pcv:
            S_v, vchg = m.prvstatus(S, H, q)
            if sched is not None:
                sched.append(dict(it=it, kind="valve", act=active.copy(),
                                  chg=(S_v != S) & active[:, None]))
            S = np.where(active[:, None], S_v, S).astype(np.int8)
            valve_change = vchg & active
        else:
            valve_change = None
        conv = relerr <= hacc
        br_conv = active & conv & (it <= max_iter)
        br_per = active & (~conv) & (it <= m.maxcheck) & (nextcheck == it)
        run_ls = br_conv | br_per
        if run_ls.any():
            S_new, chg = m.linkstatus(S, H, q)
            if sched is not None:
                sched.append(dict(it=it, conv=br_conv.copy(), per=br_per.copy(),
                                  chg=(S_new != S) & run_ls[:, None]))
            S = np.where(run_ls[:, None], S_new, S).astype(np.int8)
            stat_change = chg & br_conv
        else:
            stat_change = np.zeros(B, dtype=bool)
        if valve_change is not None:
            stat_change = stat_change | (valve_change & br_conv)
        if run_ls.any():
            nextcheck = np.where(br_conv & stat_change, it + m.checkfreq, nextcheck)
            nextcheck = np.where(br_per, nextcheck + m.checkfreq, nextcheck)
        done_now = active & conv & ((it > max_iter) | (~stat_change))
        active = active & (~done_now)
        if record:
            trace["status"].append(S.copy())
            trace["active"].append(active.copy())
            trace["relerr"].append(relerr.copy())
        if not active.any():
            break
    iters = np.where(active, maxtrials + 1, iters)
    emitter = np.zeros((B, m.N))
    emitter[:, m.junc_nodes] = e_j * has_em
    out = dict(head_ft=H, flow_cfs=q, emitter_cfs=emitter, iters=iters,
               relerr=relerr_out, converged=~active, status=S)
    if diag_ratio is not None:
        out["diag_ratio"] = diag_ratio
    if single:
        out = {k: v[0] for k, v in out.items()}
    if record:
        out["trace"] = trace
    if sched is not None:
        out["schedule"] = sched
    return out
