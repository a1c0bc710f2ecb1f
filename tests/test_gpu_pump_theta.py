# -*- coding: utf-8 -*-
"""Pump parameters on the GPU: implicit_solve(adjoint="cpu", speed=,
pump_h0=, pump_r=) runs the HIP forward with the given pump parameters and
returns d loss / d (speed, pump_h0, pump_r) from the HIP adjoint; both are
compared with the oracle (its pump derivatives are pinned by central
differences in tests/test_pump_theta.py), for CUSTOM, power-function,
linear and constant-horsepower pumps."""
import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.autodiff import implicit_solve
from wdsgpu_amd.generators import feature_network
from wdsgpu_amd.solver import GGASolver

import scalar_gga as sg

pytestmark = pytest.mark.gpu

TOL_H_FT = 1e-10 / 0.3048


def _pump_loop(ptype):
    r_pipe = sg.hw_r(800.0, 1.0, 110.0)
    nodes = [dict(kind="r", head=90.0), dict(kind="j", elev=0.0, demand=1.2),
             dict(kind="j", elev=0.0, demand=0.8), dict(kind="j", elev=0.0, demand=1.0),
             dict(kind="r", head=150.0)]
    pump = {"pf": dict(h0=-120.0, r=6.0, n=1.8), "lin": dict(h0=-120.0, r=9.0, n=1.0),
            "chp": dict(r=-8.814 * 20.0)}[ptype]
    links = [dict(kind="pump", n1=0, n2=1, ptype=ptype, q0=1.0, hmax=sg.BIG, **pump),
             dict(kind="pipe", n1=1, n2=2, r=r_pipe, diam=1.0),
             dict(kind="pipe", n1=2, n2=3, r=r_pipe, diam=1.0),
             dict(kind="pipe", n1=1, n2=3, r=1.5 * r_pipe, diam=1.0),
             dict(kind="pipe", n1=3, n2=4, r=2.0 * r_pipe, diam=1.0)]
    return sg.to_net(nodes, links, accuracy=1e-10)


def _net(which):
    if which == "feature":
        return feature_network(0, n_pumps=2, n_custom=1, n_prv=1, n_tcv=1, n_cv=0, tank=False)
    return _pump_loop(which)


@pytest.mark.parametrize("which", ["feature", "pf", "lin", "chp"])
@pytest.mark.parametrize("perturb", [False, True])
def test_pump_parameters_and_gradients_match_oracle(cuda, which, perturb):
    net = _net(which)
    B = 4
    g = np.random.default_rng(21)
    d = net.demand_cfs * np.exp(0.1 * g.standard_normal((B, net.N)))
    rh = np.nan_to_num(net.reservoir_head_ft())
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    theta = [m.pump_speed.copy(), m.pl_h0.copy(), m.pl_r.copy()]
    if perturb:
        pl = m.pump_links
        theta[0][pl] *= 1.05
        theta[1][pl] *= 0.97
        theta[2][pl] *= 1.02
    adj.set_pump_theta(m, *theta)
    W = g.normal(size=(B, net.N))
    WQ = g.normal(size=(B, net.L))
    leaves = [torch.tensor(x, device=cuda, requires_grad=True) for x in theta]
    H, Q, E = implicit_solve(s, torch.tensor(d, device=cuda), torch.tensor(rh, device=cuda),
                             accuracy=1e-10, polish_steps=3, speed=leaves[0],
                             pump_h0=leaves[1], pump_r=leaves[2])
    loss = (torch.tensor(W, device=cuda) * H).sum() + (torch.tensor(WQ, device=cuda) * Q).sum()
    loss.backward()
    sol = adj.solve_polished(m, d, np.broadcast_to(rh, (B, net.N)),
                             np.broadcast_to(net.node_ke, (B, net.N)), m.r_hw, accuracy=1e-10,
                             linsolve="sparse")
    dh = np.abs(H.detach().cpu().numpy() - sol["head"]).max()
    ref = adj.cpu_adjoint(m, sol, W, WQ, np.zeros((B, net.N)), pumps=True)[4:]
    pl = m.pump_links
    rel = []
    for name, x, r in zip(("speed", "pump_h0", "pump_r"), leaves, ref):
        a = x.grad.cpu().numpy()
        b = r.sum(0)
        assert np.all(a[np.setdiff1d(np.arange(net.L), pl)] == 0.0)
        scale = max(np.abs(b[pl]).max(), 1e-300)
        rel.append((name, float(np.abs(a[pl] - b[pl]).max() / scale)))
    print("%s perturb=%s: max|dH| %.2e ft, pump statuses %s, gradient rel %s"
          % (which, perturb, dh, sol["status"][:, pl].tolist()[0],
             ["%s %.1e" % x for x in rel]), flush=True)
    assert dh <= TOL_H_FT
    for name, r in rel:
        assert r < 1e-9, (name, r)
    # the solver's own pump parameters are back for the next call
    out = s.solve(torch.tensor(d, device=cuda), torch.tensor(rh, device=cuda), accuracy=1e-10,
                  status_machine=True)
    m0 = gga.GGAModel(net)
    ref0 = gga.solve(m0, d, np.broadcast_to(rh, (B, net.N)), accuracy=1e-10,
                     status_machine=True, extended=True)
    assert np.abs(out["head_ft"].cpu().numpy() - ref0["head_ft"]).max() <= TOL_H_FT


def test_gpu_route_refuses_pump_parameters(cuda):
    net = _net("pf")
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d = torch.tensor(net.demand_cfs, device=cuda)
    rh = torch.tensor(np.nan_to_num(net.reservoir_head_ft()), device=cuda)
    with pytest.raises(NotImplementedError):
        implicit_solve(s, d, rh, adjoint="gpu", speed=np.ones(net.L))
    with pytest.raises(ValueError):
        implicit_solve(s, d, rh, adjoint="cpu", speed=np.ones(net.L + 1))
