# -*- coding: utf-8 -*-
"""Pump-parameter gradients of the oracle's implicit adjoint
(oracle/adjoint.py pump_dtheta, cpu_adjoint(pumps=True)): d loss / d
(speed, pump_h0, pump_r) against Richardson-extrapolated central
differences of the polished solve, for power-function, linear,
constant-horsepower and CUSTOM pumps.  CPU only."""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.generators import feature_network

import scalar_gga as sg


def _pump_loop(ptype):
    # R0 -> pump -> J1, a loop J1-J2-J3, and a second reservoir R4 at 150 ft
    # feeding J3: the pump works against the loop's heads
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


def _case(which):
    if which == "custom":
        # both pumps open (seed 0): one CUSTOM, one power-function
        net = feature_network(0, n_pumps=2, n_custom=1, n_prv=0, n_tcv=0, n_cv=0, tank=False)
    else:
        net = _pump_loop(which)
    return net


def _loss_and_grads(net, theta, W, WQ):
    m = gga.GGAModel(net)
    adj.set_pump_theta(m, *theta)
    d = net.demand_cfs[None, :]
    rh = np.nan_to_num(net.reservoir_head_ft())[None, :]
    ke = net.node_ke[None, :]
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=1e-10, polish_steps=4,
                             linsolve="sparse")
    loss = float((W * sol["head"]).sum() + (WQ * sol["q"]).sum())
    return loss, sol, m


@pytest.mark.parametrize("which", ["pf", "lin", "chp", "custom"])
def test_pump_parameter_gradients_match_central_differences(which):
    net = _case(which)
    g = np.random.default_rng(11)
    W = g.normal(size=(1, net.N))
    WQ = g.normal(size=(1, net.L))
    m0 = gga.GGAModel(net)
    theta = [m0.pump_speed.copy(), m0.pl_h0.copy(), m0.pl_r.copy()]
    loss, sol, m = _loss_and_grads(net, theta, W, WQ)
    grads = adj.cpu_adjoint(m, sol, W, WQ, np.zeros((1, net.N)), pumps=True)[4:]
    for k in m0.pump_links:
        k = int(k)
        assert sol["status"][0, k] == gga.ST_OPEN
        for i, name in enumerate(("speed", "pump_h0", "pump_r")):
            ga = float(grads[i][0, k])
            base = theta[i][k]
            if which == "custom" and m0.pl_ptype[k] == 2 and i > 0:
                # a CUSTOM pump's curve coefficients come from its points
                assert ga == 0.0
                continue
            if which == "chp" and i == 1:
                assert ga == 0.0       # h0 does not enter the constant-HP branch
                continue

            def f(h):
                th = [a.copy() for a in theta]
                th[i][k] = base + h
                L_, s_, _ = _loss_and_grads(net, th, W, WQ)
                assert np.array_equal(s_["status"], sol["status"])
                return L_

            est = []
            h = 1e-3 * abs(base)
            for _ in range(4):
                d1 = (f(h) - f(-h)) / (2 * h)
                d2 = (f(h / 2) - f(-h / 2)) / h
                est.append((4 * d2 - d1) / 3)
                h /= 4
            j = min(range(len(est) - 1), key=lambda a: abs(est[a] - est[a + 1]))
            fd = est[j]
            rel = abs(ga - fd) / max(abs(fd), 1e-300)
            print("%s pump %d d/d%s: adjoint %.10e fd %.10e rel %.1e"
                  % (which, k, name, ga, fd, rel))
            assert rel < 1e-7, (which, k, name, ga, fd)


def test_closed_pump_has_no_parameter_gradient():
    # a pump that cannot lift into the loop (XHEAD) passes no derivative
    net = _pump_loop("pf")
    net.pump_h0 = np.array([-30.0])
    net.pump_hmax = np.array([30.0])
    m = gga.GGAModel(net)
    sol = adj.solve_polished(m, net.demand_cfs[None, :],
                             np.nan_to_num(net.reservoir_head_ft())[None, :],
                             net.node_ke[None, :], m.r_hw, accuracy=1e-10, linsolve="sparse")
    assert sol["status"][0, 0] <= gga.ST_CLOSED
    g = adj.cpu_adjoint(m, sol, np.ones((1, net.N)), np.ones((1, net.L)),
                        np.zeros((1, net.N)), pumps=True)[4:]
    assert all(np.all(x[0, 0] == 0.0) for x in g)
