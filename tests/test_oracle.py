# -*- coding: utf-8 -*-
"""Pins the CPU oracle (oracle/) with closed forms, invariants, brute force
and independent library routines.  CPU only, seconds."""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.generators import grid_network, planar_network, scenario_batch
from wdsgpu_amd.network import (CBIG, JUNCTION, PIPE, PRV, RESERVOIR, ST_ACTIVE,
                                ST_CLOSED, ST_OPEN, hw_resistance, read_inp)

from netutil import make_net, res_heads


def test_hw_resistance_matches_textbook_si_formula():
    # SI Hazen-Williams: hf[m] = 10.67 L Q^1.852 / (C^1.852 D^4.87); EPANET's
    # 4.727 is its US-unit form, so the two agree to the rounding of the
    # constants (~0.3%), while a wrong exponent or unit factor is off by far more.
    L_m, D_m, C, Q = 500.0, 0.3, 120.0, 0.05
    r = hw_resistance(C, L_m / 0.3048, D_m / 0.3048)
    hf_ft = r * (Q / 0.0283168466) ** 1.852
    hf_si = 10.67 * L_m * Q ** 1.852 / (C ** 1.852 * D_m ** 4.87)
    assert abs(hf_ft * 0.3048 / hf_si - 1.0) < 5e-3


def _tiny_series(n_pipes=3, demand=0.2, H0=300.0):
    nodes = [(RESERVOIR, H0, 0.0)] + [(JUNCTION, 10.0 * i, 0.0) for i in range(n_pipes)]
    nodes[-1] = (JUNCTION, 10.0 * (n_pipes - 1), demand)
    links = [dict(n1=i, n2=i + 1, len_ft=1000.0 + 100 * i, diam_ft=0.5 + 0.1 * i,
                  C=100.0 + 5 * i) for i in range(n_pipes)]
    return make_net(nodes, links)


def test_series_pipes_closed_form():
    net = _tiny_series()
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-12, status_machine=True)
    d = net.demand_cfs[-1]
    h = 300.0 - np.cumsum(net.r_hw * d ** 1.852)
    assert np.allclose(out["head_ft"][1:], h, rtol=0, atol=1e-9)
    assert np.allclose(out["flow_cfs"], d, rtol=0, atol=1e-12)


def test_parallel_pipes_split_closed_form():
    nodes = [(RESERVOIR, 200.0, 0.0), (JUNCTION, 0.0, 1.5)]
    links = [dict(n1=0, n2=1, len_ft=800.0, diam_ft=0.6, C=110.0),
             dict(n1=0, n2=1, len_ft=1200.0, diam_ft=0.8, C=130.0)]
    net = make_net(nodes, links)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-12, status_machine=True)
    r1, r2 = net.r_hw
    # equal head loss: r1 q1^1.852 = r2 q2^1.852, q1 + q2 = d
    a = (r2 / r1) ** (1 / 1.852)
    q2 = 1.5 / (1 + a)
    q1 = 1.5 - q2
    assert np.allclose(out["flow_cfs"], [q1, q2], atol=1e-10)
    assert abs(out["head_ft"][1] - (200.0 - r1 * q1 ** 1.852)) < 1e-9


def test_emitter_balance():
    # reservoir -> pipe -> junction with only an emitter: H0 - r e^1.852 = El + Ke e^2
    Ke = 50.0
    nodes = [(RESERVOIR, 150.0, 0.0), (JUNCTION, 20.0, 0.0)]
    links = [dict(n1=0, n2=1, len_ft=900.0, diam_ft=0.4, C=120.0)]
    net = make_net(nodes, links, ke=[0.0, Ke])
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), ke=net.node_ke, accuracy=1e-12,
                    status_machine=True)
    e = out["emitter_cfs"][1]
    H = out["head_ft"][1]
    assert abs(150.0 - net.r_hw[0] * e ** 1.852 - H) < 1e-9
    assert abs(H - 20.0 - Ke * e * e) < 1e-9
    assert abs(out["flow_cfs"][0] - e) < 1e-12


def _prv_net(setting_ft, H0=250.0):
    nodes = [(RESERVOIR, H0, 0.0), (JUNCTION, 0.0, 0.0), (JUNCTION, 0.0, 0.0),
             (JUNCTION, 0.0, 0.8)]
    links = [dict(n1=0, n2=1, len_ft=500.0, diam_ft=0.8, C=120.0),
             dict(n1=1, n2=2, type=PRV, diam_ft=0.8, setting_ft=setting_ft),
             dict(n1=2, n2=3, len_ft=700.0, diam_ft=0.6, C=110.0)]
    return make_net(nodes, links)


def test_prv_active_holds_setting():
    net = _prv_net(120.0)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    assert out["status"][1] == ST_ACTIVE
    assert abs(out["head_ft"][2] - 120.0) < 1e-6   # big-M penalty: O(1/CBIG)
    assert abs(out["flow_cfs"][1] - 0.8) < 1e-10


def test_prv_opens_when_upstream_below_setting():
    net = _prv_net(400.0)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    assert out["status"][1] == ST_OPEN
    assert out["head_ft"][2] < 400.0


def test_converged_planar_solution_satisfies_equations():
    net = planar_network(1)
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, 2, seed=3)
    sol = adj.solve_polished(m, d, rh, net.node_ke[None], m.r_hw, accuracy=1e-12)
    assert (sol["resid_inf"] < 1e-9).all()
    # mass balance at every junction from the raw flows (independent of the
    # residual code path)
    for b in range(2):
        q = sol["q"][b]
        inflow = np.zeros(m.N)
        np.add.at(inflow, m.n2, q)
        np.add.at(inflow, m.n1, -q)
        net_out = inflow[m.junc_nodes] - sol["emitter"][b, m.junc_nodes] - d[b, m.junc_nodes]
        assert np.abs(net_out).max() < 1e-9


def test_batch_equals_single_bitwise():
    net = planar_network(1)
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, 3, seed=4)
    outb = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, status_machine=True)
    for b in range(3):
        o1 = gga.solve(m, d[b:b + 1], rh[b:b + 1], ke=net.node_ke, accuracy=1e-6,
                       status_machine=True)
        assert np.array_equal(o1["head_ft"][0], outb["head_ft"][b])
        assert o1["iters"][0] == outb["iters"][b]


def test_csr_equals_dense_bitwise():
    net = planar_network(1)
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, 2, seed=4)
    a = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, status_machine=True,
                  assemble="dense")
    b = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, status_machine=True,
                  assemble="csr")
    assert np.array_equal(a["head_ft"], b["head_ft"])
    assert np.array_equal(a["iters"], b["iters"])


def test_assembled_matrix_is_bitwise_symmetric():
    net = planar_network(1)
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, 1, seed=4)
    q = m.init_flow()[None]
    S = m.init_status_int[None]
    P, Y = m.link_PY(q, S, m.r_hw[None])
    H = m.full_heads(np.zeros((1, m.Nj)), rh[:, m.fixed_nodes])
    A, *_ = m.assemble(P, Y, q, H, np.zeros((1, m.Nj)), net.node_ke[m.junc_nodes][None],
                       d[:, m.junc_nodes], S)
    assert np.array_equal(A[0], A[0].T)


def test_shipped_inp_fixture_converges_with_equal_iterations_across_batch():
    # networks/random_main/rand_0009.inp (the reference's smoke-test network),
    # copied as tests/golden/rand_0009.inp
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "rand_0009.inp")
    net = read_inp(path)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), status_machine=True)
    assert out["converged"]
    assert out["relerr"] <= m.hacc_default
    # energy equation per pipe at the converged state: hloss(q) ~= H1 - H2
    sol = adj.solve_polished(m, net.demand_cfs, res_heads(net), net.node_ke, m.r_hw,
                             accuracy=1e-12)
    assert sol["resid_inf"][0] < 1e-10


def _fd_grad(m, d, rh, ke, r, W, kind, j, h):
    def loss(dd, rr, kk, hh):
        s = adj.solve_polished(m, dd, hh, kk, rr, accuracy=1e-12)
        return float((s["head"] * W).sum())
    if kind == "d":
        p, q_ = d.copy(), d.copy()
        p[0, j] += h
        q_[0, j] -= h
        return (loss(p, r, ke, rh) - loss(q_, r, ke, rh)) / (2 * h)
    if kind == "r":
        p, q_ = r.copy(), r.copy()
        p[j] += h
        q_[j] -= h
        return (loss(d, p, ke, rh) - loss(d, q_, ke, rh)) / (2 * h)
    if kind == "rh":
        p, q_ = rh.copy(), rh.copy()
        p[0, j] += h
        q_[0, j] -= h
        return (loss(d, r, ke, p) - loss(d, r, ke, q_)) / (2 * h)
    p, q_ = ke.copy(), ke.copy()
    p[0, j] += h
    q_[0, j] -= h
    return (loss(d, r, p, rh) - loss(d, r, q_, rh)) / (2 * h)


def test_cpu_adjoint_matches_central_differences_grid():
    net = grid_network(0)
    m = gga.GGAModel(net)
    d = net.demand_cfs[None].copy()
    rh = res_heads(net)[None]
    ke = net.node_ke[None]
    r = m.r_hw.copy()
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=1e-12)
    W = np.zeros((1, m.N))
    W[0, m.junc_nodes[[5, 37, 88]]] = [1.0, -2.0, 0.5]
    gd, grh, gke, gr = adj.cpu_adjoint(m, sol, W, np.zeros((1, m.L)), np.zeros((1, m.N)))
    for j in (3, 41, 97):
        fd = _fd_grad(m, d, rh, ke, r, W, "d", j, 1e-3 * d[0, j])
        assert abs(gd[0, j] - fd) <= 1e-6 * max(abs(fd), 1e-3)
    for k in (0, 17, 150):
        fd = _fd_grad(m, d, rh, ke, r, W, "r", k, 1e-3 * r[k])
        assert abs(gr[0, k] - fd) <= 1e-6 * max(abs(fd), 1e-6)
    fd = _fd_grad(m, d, rh, ke, r, W, "rh", 100, 1e-2)
    assert abs(grh[0, 100] - fd) <= 1e-6 * abs(fd)
    # translation identity: raising every fixed head by c raises every head by c
    assert abs(grh[0].sum() - W[0].sum()) < 1e-9


def test_cpu_adjoint_transposed_solve_matches_dense():
    # splu(J).solve(v, trans='T') against a dense solve of J^T
    net = _prv_net(120.0)
    m = gga.GGAModel(net)
    sol = adj.solve_polished(m, net.demand_cfs[None], res_heads(net)[None],
                             net.node_ke[None], m.r_hw, accuracy=1e-12)
    hl, hg, dr, act = adj.link_coeffs(m, sol["q"][0], m.r_hw, sol["status"][0])
    _, hge, _ = adj.emitter_coeffs(m, sol["e_j"][0], net.node_ke[m.junc_nodes])
    J, _ = adj.build_J(m, hg, hge, net.node_ke[m.junc_nodes] > 0, act)
    v = np.arange(1.0, J.shape[0] + 1)
    from scipy.sparse.linalg import splu
    x1 = splu(J).solve(v, trans="T")
    x2 = np.linalg.solve(J.toarray().T, v)
    assert np.allclose(x1, x2, rtol=1e-12, atol=1e-12)
