# -*- coding: utf-8 -*-
"""Benchmark and test workloads on the seeded networks.

``Calibration`` is the forward + implicit-adjoint workload: B scenarios on
the 782-junction planar network, loss = sum over scenarios of the squared
head misfit at 30 sensor junctions against heads produced by a 'true'
Hazen-Williams roughness field (C perturbed by N(0, 5)); gradients with
respect to every demand, every pipe C (through r = 4.727 L / C^1.852 /
D^4.871), both reservoir heads and the 10 leak emitter coefficients
(through Ke = (Ucf_Q^2 / Ucf_P) / C^2).  This mirrors the reference's
training use of implicit_solve (scripts/mainline_v2/mv2_time_gpu.py:119-140:
forward + implicit adjoint, accuracy 1e-6, polish 3, max_iter 200).
"""
from __future__ import annotations

import numpy as np
import torch

from .autodiff import emitter_ke, hw_resistance, implicit_solve
from .generators import planar_network, roughness_truth, scenario_batch, sensor_nodes
from .network import PIPE, RESERVOIR
from .solver import GGASolver


class Calibration:
    def __init__(self, B=1024, seed=1, device="cuda", accuracy=1e-6, polish_steps=3,
                 n_sensors=30, net=None, lazy_checks=False):
        self.net = planar_network(seed) if net is None else net
        net = self.net
        self.B = B
        self.device = torch.device(device)
        self.accuracy = accuracy
        self.polish_steps = polish_steps
        self.solver = GGASolver(net, device=device, dense_status_machine=True)
        s = self.solver
        dev = self.device
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        d, rh = scenario_batch(net, B, seed=seed + 3)
        self.d0 = f64(d)
        self.rh0 = f64(np.nan_to_num(net.reservoir_head_ft()))
        self.pipe = torch.as_tensor(np.asarray(net.link_type) == PIPE, device=dev)
        self.len_ft = f64(np.where(np.asarray(net.link_type) == PIPE, net.len_ft, 1.0))
        self.diam_ft = f64(np.where(np.asarray(net.link_type) == PIPE, net.diam_ft, 1.0))
        self.C0 = f64(np.where(np.asarray(net.link_type) == PIPE, net.roughness, 1.0))
        self.Ce0 = f64(net.meta.get("emitter_c", np.zeros(net.N)))
        self.sensors = torch.as_tensor(sensor_nodes(net, n_sensors, seed + 4), device=dev)
        C_true = f64(np.where(np.asarray(net.link_type) == PIPE,
                              roughness_truth(net, seed + 2), 1.0))
        with torch.no_grad():
            r_true = self.r_of(C_true)
            out = s.solve(self.d0, self.rh0, ke_int=self.ke_of(self.Ce0), accuracy=1e-12,
                          status_machine=True, r_hw=r_true)
            self.H_true = out["head_ft"].index_select(1, self.sensors).clone()
        # status checks at the next step (or check()) instead of a
        # synchronisation per step
        s.lazy_checks = bool(lazy_checks)

    def r_of(self, C):
        return hw_resistance(C, self.len_ft, self.diam_ft, mask=self.pipe)

    def ke_of(self, Ce):
        return emitter_ke(Ce)

    def leaves(self):
        """Fresh leaf tensors (demand [B,N], reservoir head [N], C [L], emitter C [N])."""
        return (self.d0.clone().requires_grad_(True), self.rh0.clone().requires_grad_(True),
                self.C0.clone().requires_grad_(True), self.Ce0.clone().requires_grad_(True))

    def loss(self, d, rh, C, Ce):
        H, Q, E = implicit_solve(self.solver, d, rh, ke=self.ke_of(Ce), r_hw=self.r_of(C),
                                 accuracy=self.accuracy, max_iter=200,
                                 polish_steps=self.polish_steps, adjoint="gpu",
                                 status_machine=True)
        Hs = H.index_select(1, self.sensors)
        return ((Hs - self.H_true) ** 2).sum()

    def check(self):
        """Wait for the queued steps and raise any solver error they hit."""
        self.solver.check()

    def step(self):
        """One forward + adjoint pass at the workload's parameters; returns
        the loss and the four gradients (fresh tensors each step: the
        persistent leaves' gradients are reset, not accumulated)."""
        if getattr(self, "_params", None) is None:
            self._params = self.leaves()
        d, rh, C, Ce = self._params
        for t in self._params:
            t.grad = None
        L = self.loss(d, rh, C, Ce)
        L.backward()
        return L.detach(), d.grad, rh.grad, C.grad, Ce.grad


class Ragged:
    """The ragged multi-network workload: the 64 seeded networks of
    ``ragged_batch(seed)`` (50-5000 junctions, Hazen-Williams and
    Darcy-Weisbach, 0-4 pumps with 3-point curves, 0-3 PRVs/TCVs, 16-256
    lognormal demand scenarios each, about 4100 solves).  A step is the
    forward (accuracy 1e-6, polish 3, terminal factor kept) of every solve
    plus the implicit adjoint of a fixed seeded linear functional of the
    heads at each network's sensor junctions (about one in ten junctions, at
    most 30), all networks in one MultiSolver launch per kernel mode; the
    LDS budget is chosen by the autotuner on the batch itself."""

    def __init__(self, seed=0, device="cuda", accuracy=1e-6, polish_steps=3, autotune=True,
                 batch=None):
        from .generators import ragged_batch
        from .multi import MultiSolver
        self.batch = ragged_batch(seed) if batch is None else batch
        self.device = torch.device(device)
        self.accuracy = accuracy
        self.polish_steps = polish_steps
        dev = self.device
        self.solvers = [GGASolver(net, device=dev, dense_status_machine=True)
                        for net, _ in self.batch]
        self.inputs, self.cotangents, self.sensors = [], [], []
        g = np.random.default_rng(seed + 7)
        for i, (net, B) in enumerate(self.batch):
            d, _ = scenario_batch(net, B, seed=100 + i)
            rh = np.nan_to_num(net.reservoir_head_ft())
            self.inputs.append((torch.tensor(d, device=dev), torch.tensor(rh, device=dev),
                                torch.tensor(net.node_ke, device=dev), None))
            sens = sensor_nodes(net, min(30, max(1, net.Nj // 10)), seed + 1000 + i)
            W = np.zeros((B, net.N))
            W[:, sens] = g.normal(size=(B, sens.size))
            self.sensors.append(sens)
            self.cotangents.append((torch.tensor(W, device=dev), None, None))
        self.multi = MultiSolver(self.solvers, [b for _, b in self.batch])
        self.tuning = None
        if autotune:
            self.tuning = self.multi.autotune(self.inputs, self.cotangents, accuracy=accuracy,
                                              polish_steps=polish_steps)
        self._pending = []

    @property
    def n_solves(self):
        return sum(b for _, b in self.batch)

    def step(self, per_network=False):
        """Forward + adjoint of every solve, queued without a host
        synchronisation; the status words are checked by check().  Returns
        the forward outputs and the gradients per network.  per_network:
        the same work as one wds_solve and one wds_adjoint call per network
        (same kernels, no shared launch)."""
        ms = self.multi
        outs, keeps = ms.forward(self.inputs, accuracy=self.accuracy,
                                 polish_steps=self.polish_steps, per_network=per_network)
        grads, infos = ms.adjoint(outs, keeps, self.cotangents, per_network=per_network)
        self._pending.append(([{"st": o["st"]} for o in outs], infos))
        return outs, grads

    def check(self):
        """Raise if any queued step's forward failed or did not converge, or
        its adjoint failed."""
        pending, self._pending = self._pending, []
        for outs, infos in pending:
            self.multi.check_forward(outs)
            info = torch.cat(infos)
            if bool((info != 0).any()):
                raise RuntimeError("ragged adjoint failed on %d solves"
                                   % int((info != 0).sum().item()))


class EpsCalibration:
    """The extended-period calibration workload: B scenarios (lognormal
    demand factors, sigma 0.3) of the seeded 20000-junction geometric
    network (4 reservoirs, 6 tanks, 8 pumps with tank-level controls,
    diurnal demand pattern) over 24 hourly steps.  Loss: the squared misfit
    of 30 sensor heads and the 6 tank levels at every step against the run
    with a 'true' roughness field (C + N(0, 5)); gradients with respect to
    every demand, pipe C, reservoir head, initial tank level and leak C."""

    def __init__(self, B=256, seed=0, device="cuda", net=None, steps=24, accuracy=1e-6,
                 polish_steps=3, checkpoint_every=4, n_sensors=30):
        from .eps_adjoint import EpsModel
        from .generators import geometric_network
        self.net = geometric_network(seed) if net is None else net
        net = self.net
        self.B = B
        self.device = torch.device(device)
        self.polish_steps = polish_steps
        self.model = EpsModel(net, device=device, steps=steps, accuracy=accuracy,
                              polish_steps=polish_steps, checkpoint_every=checkpoint_every)
        dev = self.device
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        d, _ = scenario_batch(net, B, seed=seed + 3)
        self.d0 = f64(d)
        self.rh0 = f64(np.nan_to_num(net.reservoir_head_ft()))
        self.h00 = f64(net.tank_h0)
        lt = np.asarray(net.link_type)
        self.pipe = torch.as_tensor(lt == PIPE, device=dev)
        self.len_ft = f64(np.where(lt == PIPE, net.len_ft, 1.0))
        self.diam_ft = f64(np.where(lt == PIPE, net.diam_ft, 1.0))
        self.C0 = f64(np.where(lt == PIPE, net.roughness, 1.0))
        self.Ce0 = f64(net.meta.get("emitter_c", np.zeros(net.N)))
        sens = sensor_nodes(net, n_sensors, seed + 4)
        self.obs = torch.as_tensor(np.concatenate([sens, np.asarray(net.tank_node)]))
        C_true = f64(np.where(lt == PIPE, roughness_truth(net, seed + 2), 1.0))
        with torch.no_grad():
            self.H_true = self.model.simulate(self.d0, self.rh0, emitter_ke(self.Ce0),
                                              hw_resistance(C_true, self.len_ft, self.diam_ft,
                                                            mask=self.pipe),
                                              self.h00, observe=self.obs).clone()

    def step(self):
        """One forward + adjoint of the whole simulation; returns the loss
        and the gradients (demand, reservoir head, C, leak C, tank h0)."""
        leaves = [x.clone().requires_grad_(True)
                  for x in (self.d0, self.rh0, self.C0, self.Ce0, self.h00)]
        d, rh, C, Ce, h0 = leaves
        H = self.model.simulate(d, rh, emitter_ke(Ce),
                                hw_resistance(C, self.len_ft, self.diam_ft, mask=self.pipe),
                                h0, observe=self.obs)
        loss = ((H - self.H_true) ** 2).sum()
        loss.backward()
        return loss.detach(), [x.grad for x in leaves]


class LeakInversion:
    """The reference's training-loop workload (data/mainline_evidence.md
    §5.4, scripts/demo_leak_inversion.py): leak localisation by Adam on the
    emitter coefficients of candidate junctions.  B scenarios of the
    782-junction planar network (lognormal demand factors, sigma 0.3), no
    emitters but n_leaks true leaks of user coefficient leak_C among
    n_candidates random junctions; observations: the heads at n_sensors
    sensor junctions of the same batch solved with the true leaks.
    Parameters: C = softplus(theta) at the candidates (Ke = (Ucf_Q^Qexp /
    Ucf_P) / C^Qexp, the HIP emitter map); loss = mean squared head misfit
    (ft^2) + reg * sum C; Adam(lr).  A step is one forward + implicit
    adjoint of the batch and one Adam update."""

    def __init__(self, B=1024, seed=1, device="cuda", n_candidates=60, n_leaks=3, leak_C=1.5,
                 n_sensors=33, lr=0.1, reg=1e-4, accuracy=1e-6, polish_steps=3,
                 theta0=-3.0):
        net = planar_network(seed)
        net.node_ke = np.zeros(net.N)
        self.net = net
        self.B = B
        self.device = torch.device(device)
        self.accuracy, self.polish_steps, self.reg = accuracy, polish_steps, reg
        self.solver = GGASolver(net, device=device, dense_status_machine=True)
        s = self.solver
        dev = self.device
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        rng = np.random.default_rng(seed + 11)
        junc = np.nonzero(np.asarray(net.node_type) == 0)[0]
        self.candidates = np.sort(rng.choice(junc, n_candidates, replace=False))
        self.true_leaks = np.sort(rng.choice(self.candidates, n_leaks, replace=False))
        self.cand_t = torch.as_tensor(self.candidates, device=dev)
        self.sensors = torch.as_tensor(sensor_nodes(net, n_sensors, seed + 12), device=dev)
        d, _ = scenario_batch(net, B, seed=seed + 13)
        self.d0 = f64(d)
        self.rh0 = f64(np.nan_to_num(net.reservoir_head_ft()))
        C_true = np.zeros(net.N)
        C_true[self.true_leaks] = leak_C
        with torch.no_grad():
            out = s.solve(self.d0, self.rh0, ke_int=emitter_ke(f64(C_true)), accuracy=1e-12,
                          status_machine=True)
            self.H_obs = out["head_ft"].index_select(1, self.sensors).clone()
        self.theta = torch.full((n_candidates,), float(theta0), dtype=torch.float64,
                                device=dev, requires_grad=True)
        self.opt = torch.optim.Adam([self.theta], lr=lr)

    def C(self):
        return torch.nn.functional.softplus(self.theta)

    def loss(self):
        C = self.C()
        ke = torch.zeros(self.net.N, dtype=torch.float64, device=self.device).index_copy(
            0, self.cand_t, emitter_ke(C))
        H, _, _ = implicit_solve(self.solver, self.d0, self.rh0, ke=ke, accuracy=self.accuracy,
                                 max_iter=200, polish_steps=self.polish_steps, adjoint="gpu",
                                 status_machine=True)
        mis = H.index_select(1, self.sensors) - self.H_obs
        return (mis * mis).mean() + self.reg * C.sum()

    def step(self):
        """One Adam step; returns the loss before the update (a tensor)."""
        self.opt.zero_grad(set_to_none=True)
        L = self.loss()
        L.backward()
        self.opt.step()
        return L.detach()

    def ranking(self):
        """Candidate junctions by current C, largest first."""
        C = self.C().detach().cpu().numpy()
        return self.candidates[np.argsort(-C)]
