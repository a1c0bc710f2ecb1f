# -*- coding: utf-8 -*-
"""GGASolver: the batched GGA forward solve on MI355X.

Mirrors dgga's ``GGASolver`` constructor and ``solve`` keyword API
(dgga/solver.py:160-162, 2703-2706) so that code written against the
reference's dense batched path can switch to this class.  All numerics run
in libwds.so (hand-written HIP kernels, one wavefront per scenario); this
module only prepares the static network description, allocates torch
tensors and maps status codes to the reference's exceptions.

Differences from the reference, all refusals rather than approximations:
  * mode must be "dense" (the HIP path is the batched path; the scalar EPANET
    replica mode="epanet" is not part of it);
  * linear_solver "dense"/"hip" and assemble "dense"/"csr" all run the HIP
    sparse LDL^T path; "cudss" is CUDA-only and refused;
  * float64 only, CUDA/HIP device only.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import time
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from .network import (CSMALL, CBIG, BIG, TINY, MISSING, QZERO, HEXP_HW, PI_EPANET,
                      CVPIPE, PIPE, PUMP, PRV, PSV, PBV, FCV, TCV, JUNCTION,
                      ST_CLOSED, ST_OPEN, ST_ACTIVE, PUMP_CONST_HP, PUMP_CUSTOM, M_PER_FT,
                      SI_FLOW_UNITS)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class GGASolver:
    """Batched GGA forward solver (dgga/solver.py:140-495 interface).

    Beyond the reference's arguments: ``ordering`` ("auto": the cheaper by
    scheduled program cost of minimum degree and depth-limited
    independent-set rounds; "mmd" minimum degree; "rounds3".."rounds6"
    independent-set rounds with that degree cap; "natural") and ``arena``
    (run the scenario kernels in arena mode, with
    the per-scenario factor and vectors in global memory, even for a network
    that fits a CU's LDS; used to test that mode)."""

    def _check_pswitch(self, nt):
        """Refuse simple controls that watch a junction's head (the guard of
        dgga/solver.py:617-644, run by its __init__ for every dense solver,
        :478-484).  EPANET's pswitch (hydsolver.c:268-355) re-checks them after
        convergence and changes the link's status or setting; the batched
        status machine has no counterpart, so solving would be silently wrong.
        Controls on tank levels and timers belong to the extended-period
        driver and pass."""
        net = self.net
        if net.ctl_node is None or len(net.ctl_node) == 0:
            return
        cn = np.asarray(net.ctl_node, dtype=np.int64)
        ct = np.asarray(net.ctl_type, dtype=np.int64)
        hit = (cn >= 0) & (ct < 2) & (nt[np.where(cn >= 0, cn, 0)] == JUNCTION)
        if hit.any():
            i = int(np.nonzero(hit)[0][0])
            raise NotImplementedError(
                "mode=dense: control #%d watches the head of junction %s (link %s); "
                "pswitch (EPANET hydsolver.c:268-355) would change that link's status "
                "or setting after convergence, and the batched status machine has no "
                "counterpart: use the scalar EPANET path"
                % (i, net.node_id[int(cn[i])], net.link_id[int(net.ctl_link[i])]))

    # record_schedule: status changes logged per scenario before a re-run
    # with room for all of them (None: 4 L + 64)
    sched_cap = None

    def __init__(self, net, device="cuda", dtype=torch.float64, mode="dense",
                 inp_path=None, dense_tank_bound_check=True,
                 dense_status_machine=False, ordering="auto", arena=False):
        if mode != "dense":
            raise NotImplementedError(
                "mode=%r: the MI355X path is the batched dense-mode solver "
                "(mode='dense'); the scalar EPANET replica is not on it" % (mode,))
        if dtype != torch.float64:
            raise NotImplementedError("only float64 is supported (all arithmetic is FP64)")
        self.net = net
        self.mode = mode
        self.device = torch.device(device)
        self.dtype = dtype
        self.dense_status_machine = bool(dense_status_machine)
        self.dense_refine = 2
        # the reference's linear-solver cache knobs (dgga/solver.py:378-423):
        # workspaces are cached per batch size in an LRU of cudss_cache_max;
        # the adjoint's refinement steps are cudss_grad_refine; factors live
        # in each autograd context, so every forward keeps its own slot and
        # cudss_grad_slots is accepted without a limit to enforce
        self.cudss_cache_max = 8
        self.cudss_grad_refine = 2
        self.cudss_grad_slots = 1
        self._ws_cache = OrderedDict()
        lt = np.asarray(net.link_type, dtype=np.int64)
        nt = np.asarray(net.node_type, dtype=np.int64)
        self.N, self.L = net.N, net.L
        self.lt_np = lt
        allowed = {CVPIPE, PIPE, TCV, PUMP, PRV, PSV, FCV} if self.dense_status_machine \
            else {PIPE, TCV, PUMP}
        bad = set(lt.tolist()) - allowed
        if bad:
            raise NotImplementedError(
                f"mode=dense supports link_type in {sorted(allowed)}, got {sorted(bad)}")
        self.headloss_form = str(net.meta.get("headloss", "H-W")).upper()
        if self.headloss_form not in ("H-W", "D-W"):
            raise NotImplementedError(f"headloss {self.headloss_form}")
        if str(net.meta.get("demand_model", "DDA")).upper() != "DDA":
            raise NotImplementedError("only demand-driven analysis (DDA)")
        if float(net.meta.get("damp_limit", 0.0)) > 0.0:
            raise NotImplementedError("DAMPLIMIT > 0 is not supported")
        self._check_pswitch(nt)
        self.junc_nodes = np.where(nt == JUNCTION)[0]
        self.fixed_nodes = np.where(nt != JUNCTION)[0]
        self.Nj = int(self.junc_nodes.size)
        self.hexp = HEXP_HW if self.headloss_form == "H-W" else 2.0
        self.qexp = float(net.meta["qexp"])
        self.rqtol = float(net.meta.get("rqtol", 1e-7))
        self.max_iter_default = int(net.meta["trials"])
        self.hacc_default = min(max(float(net.meta["accuracy"]), 1e-5), 1e-1)
        self.htol = float(net.meta.get("htol", 0.0005))
        self.qtol = float(net.meta.get("qtol", 0.0001))
        self.checkfreq = int(net.meta.get("checkfreq", 2))
        self.maxcheck = int(net.meta.get("maxcheck", 10))
        self.extra_iter = int(net.meta.get("extra_iter", -1))
        # ---- internal status codes and settings (dgga/solver.py:1277-1319) ----
        st = np.asarray(net.init_status)
        map_init = {0: ST_CLOSED, 1: ST_OPEN, 2: ST_ACTIVE}
        self.init_status_int = np.asarray([map_init[int(s)] for s in st], dtype=np.int8)
        self.init_status_int_np = self.init_status_int
        setting = _f64(net.valve_setting_user).copy()
        ucf_p, ucf_q = net.ucf_pressure(), net.ucf_flow()
        init_setting = setting.copy()
        for k in np.where(lt >= PRV)[0]:
            if self.init_status_int[k] != ST_ACTIVE:
                init_setting[k] = MISSING
            elif lt[k] in (PRV, PSV, PBV):
                init_setting[k] = init_setting[k] / ucf_p
            elif lt[k] == FCV:
                init_setting[k] = init_setting[k] / ucf_q
        self.init_setting = init_setting
        diam = _f64(net.diam_ft)
        d4 = diam ** 2 * diam ** 2
        km = _f64(net.km_int)
        km_tcv = np.where(st == 2, np.divide(0.02517 * setting, d4, out=np.zeros_like(setting),
                                             where=d4 > 0), km)
        # ---- pumps (dgga/solver.py:1327-1348, 950-1011) ----
        pump_kind = np.full(self.L, -1, dtype=np.int32)
        ph0, pr, pn, phm, pq0 = (np.zeros(self.L) for _ in range(5))
        pn[:] = 1.0
        self.n_pumps = int(len(net.pump_link))
        if self.n_pumps:
            pl = np.asarray(net.pump_link, dtype=np.int64)
            kinds = np.asarray(net.pump_ptype, dtype=np.int64)
            pump_kind[pl] = kinds
            ph0[pl] = net.pump_h0
            pr[pl] = net.pump_r
            nn = np.asarray(net.pump_n, dtype=np.float64).copy()
            nn[np.abs(nn - 1.0) < TINY] = 1.0
            pn[pl] = nn
            sp = init_setting[pl]
            phm[pl] = np.where(kinds == PUMP_CONST_HP, BIG, sp * sp * np.asarray(net.pump_hmax))
            pq0[pl] = net.pump_q0
        # CUSTOM curves (dgga/solver.py:1346-1370, 985-1011): per-link point
        # ranges in user units, converted at solve time (curvecoeff)
        cptr = np.zeros(self.L + 1, dtype=np.int32)
        cx, cy = [], []
        if self.n_pumps:
            pptr = np.asarray(net.pump_curve_ptr, dtype=np.int64)
            cu = {int(pl[j]): j for j in range(self.n_pumps) if kinds[j] == PUMP_CUSTOM}
            for kk in range(self.L):
                j = cu.get(kk)
                if j is not None:
                    a, b = int(pptr[j]), int(pptr[j + 1])
                    cx.extend(np.asarray(net.pump_curve_q[a:b], dtype=np.float64).tolist())
                    cy.extend(np.asarray(net.pump_curve_h[a:b], dtype=np.float64).tolist())
                cptr[kk + 1] = len(cx)
        # pump parameters per link (batch-shared; set_pumps replaces them for
        # a call, as the reference's speed / pump_h0 / pump_r arguments do)
        self.pump_links = np.asarray(net.pump_link, dtype=np.int64)
        self.pl_h0, self.pl_r = ph0.copy(), pr.copy()
        self._pl_Hmax = np.zeros(self.L)
        self._pl_chp = np.zeros(self.L, dtype=bool)
        self._pl_q0 = pq0.copy()
        if self.n_pumps:
            self._pl_Hmax[pl] = np.asarray(net.pump_hmax, dtype=np.float64)
            self._pl_chp[pl] = kinds == PUMP_CONST_HP
        self._pumps_now = None   # (speed, h0, r) on the device when not the defaults
        self.ucf_flow = ucf_q
        self.ucf_head = M_PER_FT if str(net.meta["flow_units"]).upper() in SI_FLOW_UNITS else 1.0
        closed = st == 0
        self._closed0 = closed
        q0 = PI_EPANET * diam ** 2 / 4.0
        if self.n_pumps:
            q0 = np.where(lt == PUMP, init_setting * pq0, q0)
        q0 = np.where(closed, QZERO, q0)
        self.init_flow_np = np.asarray(q0, dtype=np.float64).copy()  # initlinkflow
        self.n_tanks = int(len(net.tank_node))
        self._keep = dict(
            node_type=_i32(nt), elev=_f64(net.elev_ft), n1=_i32(net.link_n1),
            n2=_i32(net.link_n2), lt=_i32(lt), st=_i32(self.init_status_int), diam=diam,
            km=km, km_tcv=_f64(km_tcv), km_valve=km.copy(), kc=_f64(net.roughness),
            q0=_f64(q0), setting=_f64(init_setting), pump_kind=pump_kind, ph0=_f64(ph0),
            pr=_f64(pr), pn=_f64(pn), phm=_f64(phm),
            tank_node=_i32(net.tank_node), tank_hmin=_f64(net.tank_hmin),
            tank_hmax=_f64(net.tank_hmax), tank_ovf=_i32(net.tank_overflow),
            cptr=cptr, cx=_f64(cx), cy=_f64(cy))
        k = self._keep
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        desc = _lib.NetworkDesc(
            N=self.N, L=self.L, node_type=P(k["node_type"], C.c_int32),
            elev=P(k["elev"], C.c_double), link_n1=P(k["n1"], C.c_int32),
            link_n2=P(k["n2"], C.c_int32), link_type=P(k["lt"], C.c_int32),
            init_status=P(k["st"], C.c_int32), diam=P(k["diam"], C.c_double),
            km=P(k["km"], C.c_double), km_tcv=P(k["km_tcv"], C.c_double),
            km_valve=P(k["km_valve"], C.c_double), kc=P(k["kc"], C.c_double),
            q0=P(k["q0"], C.c_double), setting=P(k["setting"], C.c_double),
            pump_kind=P(k["pump_kind"], C.c_int32), pump_h0=P(k["ph0"], C.c_double),
            pump_r=P(k["pr"], C.c_double), pump_n=P(k["pn"], C.c_double),
            pump_hmax=P(k["phm"], C.c_double), ntank=self.n_tanks,
            tank_node=P(k["tank_node"], C.c_int32), tank_hmin=P(k["tank_hmin"], C.c_double),
            tank_hmax=P(k["tank_hmax"], C.c_double),
            tank_overflow=P(k["tank_ovf"], C.c_int32),
            headloss=0 if self.headloss_form == "H-W" else 1, hexp=self.hexp,
            qexp=self.qexp, rqtol=self.rqtol,
            viscos=float(net.meta.get("viscosity", 1.1e-5)), htol=self.htol,
            qtol=self.qtol, checkfreq=self.checkfreq, maxcheck=self.maxcheck,
            extra_iter=self.extra_iter, pump_curve_ptr=P(k["cptr"], C.c_int32),
            pump_curve_x=P(k["cx"], C.c_double), pump_curve_y=P(k["cy"], C.c_double),
            ucf_flow=self.ucf_flow, ucf_head=self.ucf_head)
        self._lib = _lib.load()
        h = C.c_void_p()
        # arena=True runs the kernels with each scenario's factor and vectors in
        # a global-memory arena even when they would fit in LDS (networks
        # larger than a CU's LDS run that way regardless)
        codes = {"mmd": 0, "natural": 1, "auto": 2}
        codes.update({"rounds%d" % d: 3 + d for d in range(9)})
        if ordering not in codes:
            raise ValueError("ordering must be one of %s" % sorted(codes))
        order = codes[ordering] | (0x100 if arena else 0)
        _lib.check(self._lib.wds_plan_create(C.byref(desc), order, C.byref(h)),
                   "wds_plan_create")
        self._plan = h
        st_ = _lib.PlanStats()
        _lib.check(self._lib.wds_plan_stats_get(self._plan, C.byref(st_)), "wds_plan_stats_get")
        self.stats = {f: getattr(st_, f) for f, _ in _lib.PlanStats._fields_}
        self.n_pcv = int(self.stats["npcv"])
        self.has_cv = bool((lt == CVPIPE).any())
        # tensors used by the autograd wrapper
        self.r_hw = torch.as_tensor(_f64(net.r_hw), dtype=dtype, device=self.device)
        self.node_ke_default = torch.as_tensor(_f64(net.node_ke), dtype=dtype,
                                               device=self.device)
        self.junc_nodes_t = torch.as_tensor(self.junc_nodes, dtype=torch.int64,
                                            device=self.device)
        self.prv_links = np.where((lt == PRV) & (init_setting != MISSING))[0]
        n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
        self.prv_n1, self.prv_n2 = n1[self.prv_links], n2[self.prv_links]
        self._uploaded = False
        # optional kernel timing: when a list, every library launch appends
        # (name, start_event, end_event, batch, iterations tensor) recorded on
        # the launching stream
        self.event_log = None
        self._pinned = {}
        self._pinned_adj = {}
        # lazy_checks: the implicit solve's status words (forward and
        # backward) go to pinned host memory behind events and are checked at
        # the next call into this solver or at check(), instead of at a
        # synchronisation per backward; consecutive training steps then queue
        # without a host round trip.  The reference's raise points (forward /
        # backward) are the default.
        self.lazy_checks = False
        self._pending = []

    def _defer(self, check, buf, *rest):
        """Queue a status check: check(buf, *rest) once the stream has
        produced buf (an event recorded now)."""
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self._pending.append((ev, check, buf, rest))

    def poll(self):
        """Run the queued status checks whose data has arrived (no wait)."""
        while self._pending and self._pending[0][0].query():
            ev, check, buf, rest = self._pending.pop(0)
            check(buf, *rest)

    def check(self):
        """Wait for and run every queued status check (lazy_checks)."""
        while self._pending:
            ev, check, buf, rest = self._pending.pop(0)
            ev.synchronize()
            check(buf, *rest)

    def _timed(self, name, fn, B, iters=None):
        if self.event_log is None:
            return fn()
        st = torch.cuda.current_stream(self.device)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        rc = fn()
        e1.record(st)
        self.event_log.append((name, e0, e1, int(B), iters))
        return rc

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and getattr(self, "_lib", None) is not None:
            self._lib.wds_plan_destroy(h)
            self._plan = None

    # ------------------------------------------------------------------
    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def pump_theta(self, speed=None, pump_h0=None, pump_r=None):
        """(speed, pump_h0, pump_r) as float64 [L] numpy arrays, defaults
        (the network's speeds and curve coefficients) where None."""
        f = lambda x, d: d.copy() if x is None else np.array(
            x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x,
            dtype=np.float64).reshape(-1)
        th = (f(speed, self.init_setting), f(pump_h0, self.pl_h0), f(pump_r, self.pl_r))
        for a, what in zip(th, ("speed", "pump_h0", "pump_r")):
            if a.shape != (self.L,):
                raise ValueError("%s must be [L] indexed by link (L=%d)" % (what, self.L))
        if self.n_pumps and (not np.isfinite(th[0][self.pump_links]).all() or
                             (th[0][self.pump_links] < 0.0).any()):
            raise ValueError("pump speeds must be finite and >= 0")
        return th

    def set_pumps(self, theta=None):
        """Queue the pump parameters theta = (speed, pump_h0, pump_r) [L]
        (pump_theta) on the solver's stream for the calls that follow; None
        restores the network's own.  Like the reference's per-call
        speed / pump_h0 / pump_r (dgga/autodiff.py:539-603): the cut-off
        head of pumpstatus is speed^2 Hmax and the initial flow speed Q0."""
        if not self.n_pumps:
            return
        if theta is None and self._pumps_now is None:
            return
        if theta is not None and self._pumps_now is not None and all(
                np.array_equal(a, b) for a, b in zip(theta, self._pumps_now)):
            return
        sp, h0, r = theta if theta is not None else self.pump_theta()
        hmax = np.where(self._pl_chp, BIG, sp * sp * self._pl_Hmax)
        q0 = np.where(self.lt_np == PUMP, np.where(self._closed0, QZERO, sp * self._pl_q0), 0.0)
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (sp, h0, r, hmax, q0)]
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_plan_set_pumps(self._plan, self._stream(),
                                                    *[P(a) for a in keep]),
                       "wds_plan_set_pumps")
        self._pumps_now = None if theta is None else tuple(a.copy() for a in theta)

    def _workspace(self, B):
        key = (int(B), self.dtype, self.device, 0)
        ent = self._ws_cache.get(key)
        if ent is None:
            t0 = time.perf_counter()
            nbytes = int(_lib.check(self._lib.wds_workspace_bytes(self._plan, int(B)),
                                    "wds_workspace_bytes"))
            ws = torch.empty(nbytes // 8, dtype=torch.float64, device=self.device)
            ent = dict(ws=ws, plan_ms=1e3 * (time.perf_counter() - t0),
                       dev_mib=nbytes / 2.0 ** 20)
            self._ws_cache[key] = ent
            self._trim_cache()
        else:
            self._ws_cache.move_to_end(key)
        return ent["ws"]

    def _trim_cache(self):
        cap = self.cudss_cache_max
        if cap is not None and (not isinstance(cap, int) or cap < 1):
            raise ValueError("cudss_cache_max must be >= 1 or None (no limit), got %r" % (cap,))
        n = 0
        while cap is not None and len(self._ws_cache) > cap:
            self._ws_cache.popitem(last=False)
            n += 1
        return n

    @property
    def cudss_refine(self):
        """Refinement steps of every forward linear solve (the reference's
        cudss_refine, dgga/solver.py:384)."""
        return self.dense_refine

    @cudss_refine.setter
    def cudss_refine(self, v):
        self.dense_refine = int(v)

    def cudss_free(self, empty_cache=False):
        """Drop every cached workspace; returns how many
        (dgga/solver.py:2529-2562).  Memory still held by tensors a caller
        keeps (outputs, autograd factors) is theirs."""
        n = len(self._ws_cache)
        self._ws_cache.clear()
        if empty_cache and torch.cuda.is_available():
            torch.cuda.empty_cache()
        return n

    @contextlib.contextmanager
    def cudss_session(self, empty_cache=False):
        """Context manager that frees the cache on exit, exceptions included
        (dgga/solver.py:2564-2572)."""
        try:
            yield self
        finally:
            self.cudss_free(empty_cache=empty_cache)

    def cudss_cache_info(self):
        """Cached workspaces, least recently used first: (B, dtype, device,
        matrix type, allocation ms, MiB, slot) (dgga/solver.py:2574-2580)."""
        return [(k[0], k[1], k[2], "SPD", v["plan_ms"], v["dev_mib"], k[3])
                for k, v in self._ws_cache.items()]

    def cudss_plan(self, B, dtype=None, device=None):
        """Warm the plan upload and the workspace of batch size B; returns the
        milliseconds spent (dgga/solver.py:2595-2603)."""
        if dtype not in (None, self.dtype) or device not in (None, self.device):
            raise NotImplementedError("the plan serves the solver's own dtype and device")
        t0 = time.perf_counter()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_plan_upload(self._plan), "wds_plan_upload")
            self._workspace(int(B))
        return 1e3 * (time.perf_counter() - t0)

    def counters(self, reset=False):
        """Linear-algebra call counters (dgga/solver.py:2580-2593): factorize,
        solve, bwd_solve, bwd_reuse, bwd_refactorize, launches."""
        c = _lib.Counters()
        _lib.check(self._lib.wds_counters_get(self._plan, C.byref(c), int(bool(reset))),
                   "wds_counters_get")
        return {f: int(getattr(c, f)) for f, _ in _lib.Counters._fields_}

    def cudss_counters(self, reset=False):
        """The reference's name for ``counters``."""
        return self.counters(reset)

    def plan_export(self, what):
        n = int(self._lib.wds_plan_export(self._plan, int(what), None, 0))
        buf = np.zeros(max(n, 1), dtype=np.int32)
        self._lib.wds_plan_export(self._plan, int(what),
                                  buf.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return buf[:n]

    def _prep(self, x, B=None, width=None):
        t = x.to(dtype=self.dtype, device=self.device) if isinstance(x, torch.Tensor) \
            else torch.as_tensor(np.asarray(x), dtype=self.dtype, device=self.device)
        return t

    def maxtrials(self, max_iter, status_machine):
        m = int(max_iter)
        if status_machine and self.extra_iter > 0:
            m += self.extra_iter
        return m

    # ------------------------------------------------------------------
    def factor_solve(self, csr_data, F, n_refine=2, want_factor=False):
        """Batched factor + refined solve on the reference's CSR values
        ([B,nnz] in the CSR pattern of dgga/solver.py:336-369) and right-hand
        sides F [B,Nj]; returns H [B,Nj] (and the factor when asked)."""
        csr = self._prep(csr_data).contiguous()
        Fv = self._prep(F).contiguous()
        B = int(Fv.shape[0])
        if csr.shape != (B, int(self.stats["nnz_csr"])) or Fv.shape != (B, self.Nj):
            raise ValueError("factor_solve: shapes %s %s" % (tuple(csr.shape), tuple(Fv.shape)))
        H = torch.empty_like(Fv)
        info = torch.zeros(B, dtype=torch.int32, device=self.device)
        fac = torch.empty(B, int(self.stats["nnzL"]), dtype=self.dtype, device=self.device) \
            if want_factor else None
        ws = self._workspace(B)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_factor_solve(
                self._plan, self._stream(), B, _lib.ptr(csr), _lib.ptr(Fv), _lib.ptr(H),
                int(n_refine), _lib.ptr(fac), _lib.ptr(info), _lib.ptr(ws)), "wds_factor_solve")
        bad = torch.nonzero(info != 0).flatten().tolist()
        if bad:
            raise RuntimeError("batched LDL^T failed: samples %s are not positive definite" % bad)
        return (H, fac) if want_factor else H

    def link_coeffs(self, q, r_hw=None, status=None):
        """The kernels' link coefficients at flows q [B,L] (or [L]) with
        resistances r_hw ([L] or [B,L]; None: the solver's) and internal
        statuses status ([L] or [B,L]; None: the initial ones), as
        _link_coeffs_np returns them (dgga/autodiff.py:239-362): (P, hloss,
        dr) [B,L] tensors with P = 1/hgrad as assembled (1/CBIG closed, 0
        for an ACTIVE PRV) and dr = d hloss / d r of open pipes."""
        qt = self._prep(q)
        if qt.dim() == 1:
            qt = qt.unsqueeze(0)
        B = int(qt.shape[0])
        if qt.dim() != 2 or qt.shape[1] != self.L:
            raise ValueError("q must be [L] or [B,L] with L=%d" % self.L)
        qt = qt.contiguous()
        r = self.r_hw if r_hw is None else self._prep(r_hw)
        self._check_shape(r, self.L, B, "r_hw")
        if r.dim() == 2 and r.shape[0] == 1:
            r = r[0]
        r = r.contiguous()
        st = torch.as_tensor(self.init_status_int if status is None else status,
                             dtype=torch.int8, device=self.device)
        if st.dim() == 1:
            st = st.unsqueeze(0)
        if st.shape[-1] != self.L or st.shape[0] not in (1, B):
            raise ValueError("status must be [L] or [B,L]")
        st = st.expand(B, -1).contiguous()
        P = torch.empty(B, self.L, dtype=self.dtype, device=self.device)
        hl = torch.empty_like(P)
        dr = torch.empty_like(P)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_link_coeffs(
                self._plan, self._stream(), B, _lib.ptr(qt), _lib.ptr(r),
                self.L if r.dim() == 2 else 0, _lib.ptr(st), _lib.ptr(P), _lib.ptr(hl),
                _lib.ptr(dr)), "wds_link_coeffs")
        return P, hl, dr

    def factor(self, csr_data):
        """Numeric LDL^T of every scenario's CSR matrix [B,nnz] (the cuDSS
        factorize phase, dgga/solver.py:2630-2640); returns the factor
        [B,nnzL] in the plan's internal layout, for ``trisolve``."""
        csr = self._prep(csr_data).contiguous()
        B = int(csr.shape[0])
        if csr.dim() != 2 or csr.shape[1] != int(self.stats["nnz_csr"]):
            raise ValueError("factor: csr must be [B,%d]" % int(self.stats["nnz_csr"]))
        fac = torch.empty(B, int(self.stats["nnzL"]), dtype=self.dtype, device=self.device)
        info = torch.zeros(B, dtype=torch.int32, device=self.device)
        ws = self._workspace(B)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_factor(self._plan, self._stream(), B, _lib.ptr(csr),
                                            _lib.ptr(fac), _lib.ptr(info), _lib.ptr(ws)),
                       "wds_factor")
        bad = torch.nonzero(info != 0).flatten().tolist()
        if bad:
            raise RuntimeError("batched LDL^T failed: samples %s are not positive definite" % bad)
        return fac

    def trisolve(self, fac, csr_data, F, n_refine=2):
        """H [B,Nj] with A H = F from a ``factor`` result plus n_refine
        refinement steps on the CSR values (dgga/solver.py:2641-2651,
        3052-3064)."""
        csr = self._prep(csr_data).contiguous()
        Fv = self._prep(F).contiguous()
        B = int(Fv.shape[0])
        if csr.shape != (B, int(self.stats["nnz_csr"])) or Fv.shape != (B, self.Nj) or \
                tuple(fac.shape) != (B, int(self.stats["nnzL"])):
            raise ValueError("trisolve: shapes %s %s %s" % (tuple(fac.shape), tuple(csr.shape),
                                                           tuple(Fv.shape)))
        fac = fac.to(device=self.device, dtype=self.dtype).contiguous()
        H = torch.empty_like(Fv)
        ws = self._workspace(B)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_trisolve(self._plan, self._stream(), B, _lib.ptr(fac),
                                              _lib.ptr(csr), _lib.ptr(Fv), _lib.ptr(H),
                                              int(n_refine), _lib.ptr(ws)), "wds_trisolve")
        return H

    # ------------------------------------------------------------------
    def _host_status(self, out, B):
        """info and iters of a solve on the host: one asynchronous copy into a
        cached pinned buffer and one stream synchronisation."""
        buf = self._pinned.get(B)
        if buf is None:
            buf = torch.empty(2 * B, dtype=torch.int32, pin_memory=True)
            self._pinned = {B: buf}
        buf.copy_(out["st"], non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        return buf[:B].numpy(), buf[B:].numpy()

    def _pinned_info(self, B):
        buf = self._pinned_adj.get(B)
        if buf is None:
            buf = torch.empty(B, dtype=torch.int32, pin_memory=True)
            self._pinned_adj = {B: buf}
        return buf

    @staticmethod
    def _check_shape(t, n, B, what):
        """t must be [n] or [1|B, n] (ValueError otherwise)."""
        ok = (t.dim() == 1 and t.shape[0] == n) or \
            (t.dim() == 2 and t.shape[1] == n and t.shape[0] in (1, B))
        if not ok:
            raise ValueError("%s must be [%d] or [1|%d, %d], got %s"
                             % (what, n, B, n, tuple(t.shape)))

    def _args(self, demand_cfs, res_head_ft, ke_int, q0, e0, max_iter, accuracy,
              status_machine, r_hw=None, polish_steps=-1, want_factor=False, fused=True,
              status0=None, ws=None, status0_trusted=False):
        d = self._prep(demand_cfs)
        rh = self._prep(res_head_ft)
        single = d.dim() == 1
        if single:
            d = d.unsqueeze(0)
        if rh.dim() == 1:
            rh = rh.unsqueeze(0)
        B = int(d.shape[0])
        if d.dim() != 2 or rh.dim() != 2 or d.shape[1] != self.N or \
                rh.shape[-1] != self.N or rh.shape[0] not in (1, B) or B < 1:
            raise ValueError("demand/res_head must be [N] or [B,N] with N=%d" % self.N)
        # one shared vector is passed as is (no [B,N] copy); NaN heads
        # (junction entries) read as 0 in the kernel
        rh_shared = rh.shape[0] == 1 and B > 1
        rh = rh.contiguous()
        # every per-scenario array the kernels read is checked against B, N
        # and L here: the kernels index them without bounds checks
        ke = self.node_ke_default if ke_int is None else self._prep(ke_int)
        self._check_shape(ke, self.N, B, "ke")
        if ke.dim() == 2 and ke.shape[0] == 1:
            ke = ke[0]
        ke = ke.contiguous()
        ke_stride = self.N if ke.dim() == 2 else 0
        r = self.r_hw if r_hw is None else self._prep(r_hw)
        self._check_shape(r, self.L, B, "r_hw")
        if r.dim() == 2 and r.shape[0] == 1:
            r = r[0]
        r = r.contiguous()
        r_stride = self.L if r.dim() == 2 else 0
        dev = self.device
        # the fused kernel writes every per-scenario scalar; the per-iteration
        # path leaves some untouched on early exits
        scalar = torch.empty if fused else torch.zeros
        out = dict(
            head=torch.empty(B, self.N, dtype=self.dtype, device=dev),
            flow=torch.empty(B, self.L, dtype=self.dtype, device=dev),
            emitter=torch.empty(B, self.N, dtype=self.dtype, device=dev),
            st=scalar(2 * B, dtype=torch.int32, device=dev),
            relerr=scalar(B, dtype=self.dtype, device=dev),
            status=torch.empty(B, self.L, dtype=torch.int8, device=dev),
            diag_ratio=scalar(B, dtype=self.dtype, device=dev),
            resid=scalar(B, dtype=self.dtype, device=dev),
            factor=torch.empty(B, int(self.stats["nnzL"]), dtype=self.dtype, device=dev)
            if want_factor else None)
        # info and iters share one buffer: one copy to the host reads both
        out["info"], out["iters"] = out["st"][:B], out["st"][B:]
        q0t = e0t = None
        if q0 is not None:
            q0t = self._prep(q0)
            self._check_shape(q0t, self.L, B, "q0")
            q0t = q0t.expand(B, -1).contiguous()
        if e0 is not None:
            e0t = self._prep(e0)
            self._check_shape(e0t, self.N, B, "e0")
            e0t = e0t.expand(B, -1).contiguous()
        s0t = None
        if status0 is not None:
            s0t = torch.as_tensor(status0, dtype=torch.int8, device=dev)
            if s0t.dim() == 1:
                s0t = s0t.unsqueeze(0)
            if s0t.shape[-1] != self.L or s0t.shape[0] not in (1, B):
                raise ValueError("link_status must be [L] or [B,L] internal status codes")
            # statuses produced by a previous solve (status0_trusted) skip the
            # range check and its host synchronisation
            if not status0_trusted and bool(((s0t < 0) | (s0t > 7)).any()):
                raise ValueError("link_status holds internal status codes 0..7")
            s0t = s0t.expand(B, -1).contiguous()
        if ws is None:
            ws = self._workspace(B)
        keep = (d.contiguous(), rh, ke, r, q0t, e0t, ws, s0t)
        a = _lib.SolveArgs(
            B=B, demand=_lib.ptr(keep[0]), res_head=_lib.ptr(rh), ke=_lib.ptr(ke),
            res_head_shared=int(rh_shared),
            ke_stride=ke_stride, r_hw=_lib.ptr(r), r_stride=r_stride, q0=_lib.ptr(q0t),
            e0=_lib.ptr(e0t), max_iter=int(max_iter), accuracy=float(accuracy),
            status_machine=int(bool(status_machine)), n_refine=int(self.dense_refine),
            polish_steps=int(polish_steps), head=_lib.ptr(out["head"]),
            flow=_lib.ptr(out["flow"]), emitter=_lib.ptr(out["emitter"]),
            iters=_lib.ptr(out["iters"]), relerr=_lib.ptr(out["relerr"]),
            status=_lib.ptr(out["status"]), diag_ratio=_lib.ptr(out["diag_ratio"]),
            info=_lib.ptr(out["info"]), resid=_lib.ptr(out["resid"]),
            factor=_lib.ptr(out["factor"]), workspace=_lib.ptr(ws), status0=_lib.ptr(s0t))
        return a, out, keep, single, B

    def _validate(self, assemble, linear_solver, status_machine, link_status, link_setting):
        if assemble not in ("dense", "csr"):
            raise ValueError(f"assemble must be 'dense' or 'csr', got {assemble!r}")
        if linear_solver == "cudss":
            raise NotImplementedError("linear_solver='cudss' is CUDA-only; use 'hip'")
        if linear_solver not in ("dense", "hip"):
            raise ValueError(f"linear_solver must be 'dense' or 'hip', got {linear_solver!r}")
        if not self.dense_status_machine and (link_status is not None or
                                              link_setting is not None or status_machine):
            raise NotImplementedError(
                "link_status/link_setting/status_machine need dense_status_machine=True")
        if link_setting is not None:
            raise NotImplementedError("link_setting warm starts are not supported "
                                      "(per-scenario pump speeds are not on this path)")
        if link_status is not None and not status_machine:
            raise NotImplementedError("link_status needs status_machine=True")
        if self.has_cv and not status_machine:
            raise NotImplementedError("networks with CV pipes need status_machine=True")
        if self.n_pcv and not status_machine:
            raise NotImplementedError("networks with PRVs need status_machine=True")
        if (self.n_pumps or self.n_tanks) and not status_machine:
            raise NotImplementedError(
                "networks with pumps or tanks need status_machine=True on this path "
                "(pumpstatus/tankstatus run inside the batched status machine)")
        if not self.device.type == "cuda":
            raise NotImplementedError("the solver runs on a HIP device only")

    def solve(self, demand_cfs, res_head_ft, ke_int=None, q0=None, e0=None,
              max_iter=None, accuracy=None, link_status=None, link_setting=None,
              status_machine=False, assemble="dense", linear_solver="dense",
              record_schedule=False, r_hw=None):
        """Batched steady-state solve (dgga/solver.py:2703-3212), with the
        reference's keyword defaults; every assemble / linear_solver value
        this path accepts ("dense", "csr"; "dense", "hip") runs the same HIP
        sparse LDL^T solve, whose results match the reference's dense
        Cholesky path (the keyword only names the route).  Returns the
        reference's dict: head_ft [B,N], flow_cfs [B,L], emitter_cfs [B,N],
        iters [B], relerr [B], converged [B], status [B,L] (status machine),
        diag_ratio [B] (networks with PRVs).  1-D inputs give 1-D outputs."""
        self._validate(assemble, linear_solver, status_machine, link_status, link_setting)
        if record_schedule and not status_machine:
            raise NotImplementedError(
                "record_schedule=True covers the batched status machine only "
                "(status_machine=True; dgga/solver.py:2802-2806)")
        max_iter = self.max_iter_default if max_iter is None else int(max_iter)
        hacc = self.hacc_default if accuracy is None else float(accuracy)
        a, out, keep, single, B = self._args(demand_cfs, res_head_ft, ke_int, q0, e0,
                                             max_iter, hacc, status_machine, r_hw=r_hw,
                                             status0=link_status)
        cap = self.sched_cap if self.sched_cap else 4 * self.L + 64
        for _ in range(2):
            sched = self._sched_buffers(a, B, self.maxtrials(max_iter, True), cap) \
                if record_schedule else None
            with torch.cuda.device(self.device):
                _lib.check(self._lib.wds_solve(self._plan, self._stream(), C.byref(a)),
                           "wds_solve")
            if sched is None:
                break
            need = int(sched[2].max().item())
            if need <= cap:
                break
            cap = need   # deterministic re-run with room for every change
        res = self._finish(out, single, max_iter, status_machine)
        if sched is not None:
            res["schedule"] = self._schedule(sched, B)
        return res

    def _sched_buffers(self, a, B, maxtrials, cap):
        dev = self.device
        flags = torch.zeros(B, maxtrials, dtype=torch.uint8, device=dev)
        events = torch.empty(B, max(cap, 1), dtype=torch.int32, device=dev)
        count = torch.zeros(B, dtype=torch.int32, device=dev)
        a.sched_flags, a.sched_events = _lib.ptr(flags), _lib.ptr(events)
        a.sched_count, a.sched_cap = _lib.ptr(count), int(cap)
        return flags, events, count

    def _schedule(self, sched, B):
        """record_schedule output in the reference's layout (dgga/solver.py:
        3136-3160): per executed iteration, a prvstatus event {it, kind:
        "valve", act[B], chg[B,L]} when the network has PRVs, and a
        linkstatus event {it, conv[B], per[B], chg[B,L]} when any scenario
        ran linkstatus; chg marks the links whose status that step changed."""
        flags, events, count = (t.cpu().numpy() for t in sched)
        ran = (flags & 1) != 0
        n_it = int(np.nonzero(ran.any(0))[0].max()) + 1 if ran.any() else 0
        chg = {}
        for b in range(B):
            ev = events[b, :int(count[b])].astype(np.int64)
            for it, kind, k in zip(ev >> 20, (ev >> 19) & 1, ev & ((1 << 19) - 1)):
                m = chg.setdefault((int(it), int(kind)), np.zeros((B, self.L), dtype=bool))
                m[b, int(k)] = True
        none = np.zeros((B, self.L), dtype=bool)
        out = []
        for it in range(1, n_it + 1):
            f = flags[:, it - 1]
            if self.n_pcv:
                out.append(dict(it=it, kind="valve", act=(f & 1) != 0,
                                chg=chg.get((it, 0), none).copy()))
            conv, per = (f & 2) != 0, (f & 4) != 0
            if (conv | per).any():
                out.append(dict(it=it, conv=conv, per=per, chg=chg.get((it, 1), none).copy()))
        return out

    def _finish(self, out, single, max_iter, status_machine):
        info, _ = self._host_status(out, int(out["info"].shape[0]))
        bad = np.nonzero(info > 0)[0].tolist()
        if bad:
            its = info[bad].tolist()
            raise RuntimeError(
                "batched LDL^T failed: samples %s have a matrix that is not positive "
                "definite (iterations %s); EPANET would apply badvalve, the batched "
                "path has no counterpart (dgga/solver.py:3043-3051)" % (bad, its))
        maxtrials = self.maxtrials(max_iter, status_machine)
        iters = out["iters"].to(torch.int64)
        res = dict(head_ft=out["head"], flow_cfs=out["flow"], emitter_cfs=out["emitter"],
                   iters=iters, relerr=out["relerr"], converged=iters <= maxtrials)
        if status_machine:
            res["status"] = out["status"]
        if self.n_pcv:
            res["diag_ratio"] = out["diag_ratio"]
        if single:
            res = {k: v[0] for k, v in res.items()}
        return res

    # ------------------------------------------------------------------
    def solve_unfused(self, demand_cfs, res_head_ft, ke_int=None, max_iter=None,
                      accuracy=None, status_machine=False, sync_every=1):
        """Plain per-iteration path: one kernel launch per Newton iteration
        with the state in global memory and a host-side convergence check
        after every ``sync_every`` launches (the reference's host-driven loop,
        dgga/solver.py:2922-3179)."""
        self._validate("csr", "hip", status_machine, None, None)
        max_iter = self.max_iter_default if max_iter is None else int(max_iter)
        hacc = self.hacc_default if accuracy is None else float(accuracy)
        a, out, keep, single, B = self._args(demand_cfs, res_head_ft, ke_int, None, None,
                                             max_iter, hacc, status_machine, fused=False)
        ctrl = torch.zeros(B, 4, dtype=torch.int32, device=self.device)
        maxtrials = self.maxtrials(max_iter, status_machine)
        with torch.cuda.device(self.device):
            for it in range(1, maxtrials + 1):
                _lib.check(self._lib.wds_newton_iter(self._plan, self._stream(), C.byref(a),
                                                     it, _lib.ptr(ctrl)), "wds_newton_iter")
                if it % sync_every == 0 and not bool(ctrl[:, 0].any()):
                    break
        return self._finish(out, single, max_iter, status_machine)
