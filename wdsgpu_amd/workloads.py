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
        """One forward + adjoint pass; returns the four gradients."""
        d, rh, C, Ce = self.leaves()
        L = self.loss(d, rh, C, Ce)
        L.backward()
        return L.detach(), d.grad, rh.grad, C.grad, Ce.grad
