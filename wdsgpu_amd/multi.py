# -*- coding: utf-8 -*-
"""Ragged batches: many networks solved together (forward and implicit
adjoint) in one persistent launch per kernel mode.

Each network keeps its own GGASolver (plan, keyword semantics, error
classes); ``MultiSolver`` hands all of them to the library's multi-network
entry points (include/wds.h wds_multi_*), whose persistent kernels take
(network, scenario) work items from a queue, costliest networks first.
Networks whose per-scenario layout fits the LDS budget run in the LDS-mode
kernel, the others in the arena-mode kernel; budget 0 runs everything in one
arena-mode launch.  ``autotune`` times the candidate budgets on the actual
batch and keeps the fastest; every budget gives the same bits.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib


class MultiSolver:
    def __init__(self, solvers, B, lds_budget=0):
        if len(solvers) != len(B) or not solvers:
            raise ValueError("one batch size per solver")
        self.solvers = list(solvers)
        self.B = [int(b) for b in B]
        self.device = self.solvers[0].device
        if any(s.device != self.device for s in self.solvers):
            raise ValueError("all solvers must share one device")
        self._lib = _lib.load()
        self._h = None
        self.timings = {}
        # a list to record (name, start event, end event) of every library
        # call on the launching stream, or None
        self.event_log = None
        self._build(lds_budget)

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            self._lib.wds_multi_destroy(self._h)
            self._h = None

    def _build(self, budget):
        if self._h is not None:
            self._lib.wds_multi_destroy(self._h)
            self._h = None
        n = len(self.solvers)
        plans = (C.c_void_p * n)(*[s._plan.value for s in self.solvers])
        Bs = (C.c_int32 * n)(*self.B)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.wds_multi_create(plans, n, Bs, int(budget), C.byref(h)),
                       "wds_multi_create")
        self._h = h
        self.lds_budget = int(budget)
        self._ws = [torch.empty(int(self._lib.wds_multi_workspace_bytes(h, i)) // 8,
                                dtype=torch.float64, device=self.device) for i in range(n)]

    def info(self):
        """(scenarios in LDS mode, scenarios in arena mode, LDS bytes)."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int64()
        _lib.check(self._lib.wds_multi_info(self._h, C.byref(a), C.byref(b), C.byref(c)),
                   "wds_multi_info")
        return a.value, b.value, c.value

    def forward(self, inputs, accuracy=1e-6, max_iter=200, polish_steps=3,
                status_machine=True, want_factor=True, per_network=False):
        """inputs: per network (demand [B,N], res_head [N] or [B,N], ke [N] or
        [B,N] or None, r_hw [L] or [B,L] or None).  Returns per network the
        output dict of the solve (head, flow, emitter, iters, info, relerr,
        status, factor, ...) and the kept inputs, for ``adjoint``; the
        status words are checked by ``check_forward``.  per_network=True
        makes the same work one wds_solve call per network instead (the
        comparison point for the single launch)."""
        if len(inputs) != len(self.solvers):
            raise ValueError("one input tuple per network")
        args, outs, keeps = [], [], []
        for i, (s, (d, rh, ke, r)) in enumerate(zip(self.solvers, inputs)):
            a, out, keep, single, B = s._args(d, rh, ke, None, None, max_iter, accuracy,
                                              status_machine, r_hw=r,
                                              polish_steps=int(polish_steps),
                                              want_factor=want_factor, ws=self._ws[i])
            if B != self.B[i]:
                raise ValueError("network %d: %d scenarios, the batch was built for %d"
                                 % (i, B, self.B[i]))
            args.append(a)
            outs.append(out)
            keeps.append(keep)
        if per_network:
            with torch.cuda.device(self.device):
                for s, a in zip(self.solvers, args):
                    self._timed("wds_solve", lambda: s._lib.wds_solve(s._plan, s._stream(),
                                                                      C.byref(a)))
            return outs, keeps
        arr = (_lib.SolveArgs * len(args))(*args)
        with torch.cuda.device(self.device):
            self._timed("wds_multi_solve",
                        lambda: self._lib.wds_multi_solve(self._h, self.solvers[0]._stream(),
                                                          arr))
        return outs, keeps

    def _timed(self, name, fn):
        if self.event_log is None:
            _lib.check(fn(), name)
            return
        st = torch.cuda.current_stream(self.device)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        rc = fn()
        e1.record(st)
        _lib.check(rc, name)
        self.event_log.append((name, e0, e1))

    def check_forward(self, outs, max_iter=200):
        """One host copy of every network's status words; RuntimeError
        naming the network and samples that failed or did not converge."""
        st = torch.cat([o["st"] for o in outs]).cpu().numpy()
        pos = 0
        for i, (s, o) in enumerate(zip(self.solvers, outs)):
            B = self.B[i]
            info, iters = st[pos:pos + B], st[pos + B:pos + 2 * B]
            pos += 2 * B
            bad = np.nonzero((info != 0) | (iters > s.maxtrials(max_iter, True)))[0]
            if bad.size:
                raise RuntimeError("network %d: samples %s failed or did not converge "
                                   "(info %s, iters %s)" % (i, bad[:8].tolist(),
                                                           info[bad[:8]].tolist(),
                                                           iters[bad[:8]].tolist()))

    def adjoint(self, outs, keeps, cotangents, n_refine=2, per_network=False):
        """Implicit adjoint of every network: cotangents per network (gH
        [B,N], gQ [B,L], gE [B,N], each may be None).  Returns per network
        (g_demand, g_res_head, g_ke, g_r) [B,*] and the info tensors.
        per_network=True: one wds_adjoint call per network instead."""
        args, grads, infos = [], [], []
        for i, (s, o, k, (gH, gQ, gE)) in enumerate(zip(self.solvers, outs, keeps, cotangents)):
            B = self.B[i]
            dev = s.device
            gd = torch.empty(B, s.N, dtype=s.dtype, device=dev)
            grh, gke = torch.empty_like(gd), torch.empty_like(gd)
            gr = torch.empty(B, s.L, dtype=s.dtype, device=dev)
            info = torch.empty(B, dtype=torch.int32, device=dev)
            c = lambda g: None if g is None else g.to(dtype=s.dtype, device=dev).contiguous()
            gHc, gQc, gEc = c(gH), c(gQ), c(gE)
            ke, r = k[2], k[3]
            a = _lib.AdjointArgs(
                B=B, demand=_lib.ptr(k[0]), ke=_lib.ptr(ke),
                ke_stride=s.N if ke.dim() == 2 else 0, r_hw=_lib.ptr(r),
                r_stride=s.L if r.dim() == 2 else 0, head=_lib.ptr(o["head"]),
                flow=_lib.ptr(o["flow"]), emitter=_lib.ptr(o["emitter"]),
                status=_lib.ptr(o["status"]), factor=_lib.ptr(o["factor"]),
                gH=_lib.ptr(gHc), gQ=_lib.ptr(gQc), gE=_lib.ptr(gEc), n_refine=int(n_refine),
                g_demand=_lib.ptr(gd), g_res_head=_lib.ptr(grh), g_ke=_lib.ptr(gke),
                g_r=_lib.ptr(gr), info=_lib.ptr(info), workspace=_lib.ptr(self._ws[i]))
            args.append(a)
            grads.append((gd, grh, gke, gr, gHc, gQc, gEc))
            infos.append(info)
        if per_network:
            with torch.cuda.device(self.device):
                for s, a in zip(self.solvers, args):
                    self._timed("wds_adjoint", lambda: s._lib.wds_adjoint(s._plan, s._stream(),
                                                                          C.byref(a)))
            return [g[:4] for g in grads], infos
        arr = (_lib.AdjointArgs * len(args))(*args)
        with torch.cuda.device(self.device):
            self._timed("wds_multi_adjoint",
                        lambda: self._lib.wds_multi_adjoint(self._h,
                                                            self.solvers[0]._stream(), arr))
        return [g[:4] for g in grads], infos

    def autotune(self, inputs, cotangents=None, budgets=None, reps=2, **kw):
        """Time forward (+ adjoint when cotangents are given) for each LDS
        budget on this batch, keep the fastest; returns {budget: ms}."""
        if budgets is None:
            lds = sorted({int(s.stats["lds_bytes"]) for s in self.solvers
                          if s.stats["arena_bytes"] == 0})
            # 0 (one arena launch), and budgets that admit ever larger
            # networks into the LDS launch
            budgets = [0] + [b for b in (24 << 10, 48 << 10, 96 << 10, 160 << 10)
                             if any(x <= b for x in lds)]
        res = {}
        for b in budgets:
            self._build(b)
            best = None
            for _ in range(reps + 1):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                outs, keeps = self.forward(inputs, **kw)
                if cotangents is not None:
                    self.adjoint(outs, keeps, cotangents)
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) * 1e3
                best = t if best is None else min(best, t)
            res[b] = best
        pick = min(res, key=res.get)
        self._build(pick)
        self.timings = res
        return res
