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


def _one_link_net(link, H0=200.0, el=0.0, demand=0.5, headloss="H-W", pump=None):
    """reservoir -> one link -> junction with a demand (a tree: the link flow
    is the demand and the junction head follows from one energy equation)."""
    nodes = [(RESERVOIR, H0, 0.0), (JUNCTION, el, demand)]
    net = make_net(nodes, [dict(n1=0, n2=1, **link)])
    net.meta["headloss"] = headloss
    if pump is not None:
        net.pump_link = np.array([0], dtype=np.int32)
        net.pump_ptype = np.array([1], dtype=np.int8)
        for k in ("h0", "r", "n", "q0", "qmax", "hmax"):
            setattr(net, "pump_" + k, np.array([pump[k]]))
        net.valve_setting_user[0] = 1.0
    return net


def test_pump_power_curve_fit_passes_through_its_points():
    from wdsgpu_amd.network import pump_power_curve
    h0, h1, q1, h2, q2 = 60.0, 45.0, 80.0, 0.0, 160.0
    c = pump_power_curve(h0, h1, q1, h2, q2, flow_units="LPS", si=True)
    # back in user units: h(q) = -h0_int*0.3048 - r_int*0.3048/28.317^n * q^n
    for q, h in ((0.0, h0), (q1, h1), (q2, h2)):
        hq = (-c["h0"] - c["r"] * (q / 28.317) ** c["n"]) * 0.3048
        assert abs(hq - h) < 1e-9


def test_pump_lift_closed_form():
    from wdsgpu_amd.network import PUMP, pump_power_curve
    c = pump_power_curve(50.0, 37.5, 60.0, 0.0, 120.0, flow_units="CFS", si=False)
    net = _one_link_net(dict(type=PUMP, diam_ft=0.0), H0=100.0, demand=30.0, pump=c)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    q = 30.0
    assert abs(out["flow_cfs"][0] - q) < 1e-9
    lift = -c["h0"] - c["r"] * q ** c["n"]
    assert abs(out["head_ft"][1] - (100.0 + lift)) < 1e-8


CUSTOM_POINTS = [(2.0, 60.0), (6.0, 50.0), (10.0, 32.0), (14.0, 5.0)]


@pytest.mark.parametrize("q", [1.0, 2.0, 4.0, 6.0, 8.0, 12.0, 15.0])
def test_custom_pump_lift_closed_form(q):
    # on the first point, inside each segment, on an inner point, beyond the
    # last; below the first point the extended segment lifts above the
    # shutoff head Hmax = first head, so pumpstatus closes it (XHEAD)
    from netutil import curve_lift, custom_pump_tree
    net = custom_pump_tree(CUSTOM_POINTS, q)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    if q < CUSTOM_POINTS[0][0]:
        assert out["status"][0] == 0
        return
    assert out["status"][0] == 3
    assert abs(out["flow_cfs"][0] - q) < 1e-9
    assert abs(out["head_ft"][1] - (100.0 + curve_lift(CUSTOM_POINTS, q))) < 1e-8


def test_custom_pump_affinity_law_and_si_units():
    # speed w: lift(q) = w^2 curve(q / w); SI points (L/s, m) are read in
    # user units and converted with 28.317 L/s per cfs and 0.3048 m per ft
    from netutil import curve_lift, custom_pump_tree
    w, q = 0.8, 6.0
    net = custom_pump_tree(CUSTOM_POINTS, q, speed=w)
    out = gga.solve(gga.GGAModel(net), net.demand_cfs, res_heads(net), accuracy=1e-10,
                    status_machine=True)
    assert abs(out["head_ft"][1] - (100.0 + w * w * curve_lift(CUSTOM_POINTS, q / w))) < 1e-8
    pts = [(40.0, 30.0), (120.0, 24.0), (200.0, 10.0)]
    q_cfs = 100.0 / 28.317
    net = custom_pump_tree(pts, q_cfs, flow_units="LPS")
    out = gga.solve(gga.GGAModel(net), net.demand_cfs, res_heads(net), accuracy=1e-10,
                    status_machine=True)
    assert abs(out["head_ft"][1] - (100.0 + curve_lift(pts, 100.0) / 0.3048)) < 1e-8


def test_custom_pump_curve_checks():
    from wdsgpu_amd.network import pump_custom_curve
    with pytest.raises(ValueError, match="227"):
        pump_custom_curve([(1.0, 50.0), (2.0, 50.0)])
    c = pump_custom_curve([(10.0, 50.0), (30.0, 40.0), (50.0, 20.0)], flow_units="LPS")
    # EPANET input2.c:441-443: Q0 mid-range, Qmax last flow, Hmax first head
    assert abs(c["q0"] * 28.317 - 30.0) < 1e-12
    assert abs(c["qmax"] * 28.317 - 50.0) < 1e-12
    assert abs(c["hmax"] * 0.3048 - 50.0) < 1e-12


def test_darcy_weisbach_turbulent_closed_form():
    # Swamee-Jain friction factor computed independently of the oracle's
    # constants: f = 0.25 / log10(e/(3.7 D) + 5.74 / Re^0.9)^2, h = f L/D v^2/2g
    from wdsgpu_amd.network import dw_resistance
    L_ft, D_ft, e_ft, Q = 1500.0, 0.5, 0.0005, 0.6
    net = _one_link_net(dict(len_ft=L_ft, diam_ft=D_ft, C=100.0), headloss="D-W", demand=Q)
    net.roughness[:] = e_ft
    net.r_hw[:] = dw_resistance(L_ft, D_ft)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    nu = 1.1e-5
    A = np.pi * D_ft ** 2 / 4.0
    v = Q / A
    Re = v * D_ft / nu
    assert Re > 4000
    f = 0.25 / np.log10(e_ft / (3.7 * D_ft) + 5.74 / Re ** 0.9) ** 2
    hL = f * L_ft / D_ft * v * v / (2.0 * 32.2)
    assert abs((200.0 - out["head_ft"][1]) / hL - 1.0) < 1e-6


def test_darcy_weisbach_laminar_closed_form():
    # Hagen-Poiseuille: h = 32 nu L v / (g D^2)
    from wdsgpu_amd.network import dw_resistance
    L_ft, D_ft, Q = 300.0, 0.3, 1e-4
    net = _one_link_net(dict(len_ft=L_ft, diam_ft=D_ft, C=100.0), headloss="D-W", demand=Q)
    net.roughness[:] = 0.0005
    net.r_hw[:] = dw_resistance(L_ft, D_ft)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    v = Q / (np.pi * D_ft ** 2 / 4.0)
    assert v * D_ft / 1.1e-5 < 2000
    hL = 32.0 * 1.1e-5 * L_ft * v / (32.2 * D_ft ** 2)
    assert abs((200.0 - out["head_ft"][1]) / hL - 1.0) < 1e-3  # pi = 3.141592654


def test_tcv_closed_form():
    # minor loss K v^2 / 2g; EPANET's 0.02517 is 8 / (pi^2 g) with g = 32.2
    # ft/s^2 rounded to four digits
    from wdsgpu_amd.network import TCV
    K, D_ft, Q = 6.0, 0.5, 0.4
    net = _one_link_net(dict(type=TCV, diam_ft=D_ft, setting=K), demand=Q)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    v = Q / (np.pi * D_ft ** 2 / 4.0)
    hL = K * v * v / (2.0 * 32.2)
    assert abs((200.0 - out["head_ft"][1]) / hL - 1.0) < 2e-4


def test_check_valve_closes_against_reverse_head():
    # two reservoirs joined through a junction; the CV pipe points from the
    # low reservoir towards the junction, so it must close
    from wdsgpu_amd.network import CVPIPE, ST_CLOSED
    nodes = [(RESERVOIR, 100.0, 0.0), (RESERVOIR, 150.0, 0.0), (JUNCTION, 0.0, 0.3)]
    links = [dict(n1=0, n2=2, type=CVPIPE, len_ft=500.0, diam_ft=0.5, C=120.0),
             dict(n1=1, n2=2, len_ft=500.0, diam_ft=0.5, C=120.0)]
    net = make_net(nodes, links)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=1e-10, status_machine=True)
    assert out["status"][0] == ST_CLOSED
    assert abs(out["flow_cfs"][0]) < 1e-5
    assert abs(out["flow_cfs"][1] - 0.3) < 1e-5


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
    # copied as tests/golden/networks/random_main/rand_0009.inp
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "networks", "random_main",
                        "rand_0009.inp")
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


@pytest.mark.parametrize("headloss,seed,n_custom,pump_st,r_links", [
    ("H-W", 0, 0, 3, (3, 40)), ("H-W", 2, 0, 0, (3, 40)), ("H-W", 0, 1, 3, (3, 40)),
    # D-W: turbulent (3), laminar (20, 43; 4, 23) and Dunlop-transition (37;
    # 5) pipes
    ("D-W", 0, 0, 3, (3, 20, 37, 43)), ("D-W", 2, 0, 0, (3, 4, 5, 23))])
def test_cpu_adjoint_with_pumps_matches_central_differences(headloss, seed, n_custom,
                                                            pump_st, r_links):
    # feature network: a power-function pump (OPEN for seed 0, XHEAD for
    # seed 2) or a CUSTOM one (OPEN, inside a segment), an ACTIVE PRV
    # branch, a TCV, a check valve, a tank and leaks
    from wdsgpu_amd.generators import feature_network
    net = feature_network(seed, headloss=headloss, n_prv=1, n_custom=n_custom)
    m = gga.GGAModel(net)
    d = net.demand_cfs[None].copy()
    rh = res_heads(net)[None]
    ke = net.node_ke[None]
    r = m.r_hw.copy()
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=1e-12)
    pump = int(m.pump_links[0])
    assert sol["status"][0, pump] == pump_st
    W = np.zeros((1, m.N))
    jn = m.junc_nodes
    W[0, jn[[1, 7, 20, len(jn) - 1]]] = [1.0, -2.0, 0.5, 1.5]
    gd, grh, gke, gr = adj.cpu_adjoint(m, sol, W, np.zeros((1, m.L)), np.zeros((1, m.N)))
    for j in jn[[0, 9, 33, len(jn) - 2]]:
        fd = _fd_grad(m, d, rh, ke, r, W, "d", j, 1e-3 * d[0, j])
        assert abs(gd[0, j] - fd) <= 1e-6 * max(abs(fd), 1e-3)
    for k in r_links:
        # Richardson-extrapolated central differences: laminar D-W pipes
        # carry gradients near 1e-7, where a single small step is noise
        f1 = _fd_grad(m, d, rh, ke, r, W, "r", k, 2e-2 * r[k])
        f2 = _fd_grad(m, d, rh, ke, r, W, "r", k, 1e-2 * r[k])
        fd = (4.0 * f2 - f1) / 3.0
        assert abs(gr[0, k] - fd) <= 1e-6 * max(abs(fd), 1e-6)
    # fixed heads (the pump's suction reservoir, the other reservoir, the
    # tank); no translation identity here: the ACTIVE PRV pins an absolute head
    for f in m.fixed_nodes:
        fd = _fd_grad(m, d, rh, ke, r, W, "rh", int(f), 1e-2)
        assert abs(grh[0, f] - fd) <= 1e-6 * max(abs(fd), 1e-6)


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


def test_feature_cases_cover_the_status_machine():
    seen = set()
    from wdsgpu_amd.generators import feature_network
    from wdsgpu_amd.network import CVPIPE, PUMP, ST_XHEAD
    from netutil import FEATURE_CASES
    for headloss, seed, level in FEATURE_CASES:
        net = feature_network(seed, headloss=headloss, n_prv=2, tank_level_m=level)
        m = gga.GGAModel(net)
        d, rh = scenario_batch(net, 2, seed=11)
        ref = gga.solve(m, d, np.nan_to_num(rh), ke=net.node_ke, accuracy=1e-3,
                        status_machine=True)
        lt = np.asarray(net.link_type)
        st = ref["status"]
        for kind, code in ((PUMP, ST_OPEN), (PUMP, ST_XHEAD), (CVPIPE, ST_OPEN),
                           (CVPIPE, ST_CLOSED), (PRV, ST_ACTIVE)):
            if (st[:, lt == kind] == code).any():
                seen.add((kind, code))
        if (st[:, m.tl_k] == 1).any():
            seen.add(("tank", 1))
    assert seen == {(PUMP, ST_OPEN), (PUMP, ST_XHEAD), (CVPIPE, ST_OPEN),
                    (CVPIPE, ST_CLOSED), (PRV, ST_ACTIVE), ("tank", 1)}


@pytest.mark.parametrize("seed", [0, 2])
def test_schedule_accounts_for_every_status_change(seed):
    # record_schedule: at every executed iteration the links whose status
    # moved are exactly those the prvstatus and linkstatus events mark, one
    # prvstatus event runs per iteration, and every scenario ends on a
    # convergence-branch linkstatus event at its final iteration
    from wdsgpu_amd.generators import feature_network
    net = feature_network(seed, headloss="H-W", n_prv=2)
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, 6, seed=11)
    rh = np.nan_to_num(rh)
    out = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-3, status_machine=True,
                    record=True, record_schedule=True)
    sched, tr = out["schedule"], out["trace"]["status"]
    n_it = len(tr)
    valve = [e for e in sched if e.get("kind") == "valve"]
    assert [e["it"] for e in valve] == list(range(1, n_it + 1))
    prev = np.broadcast_to(m.init_status_int, tr[0].shape)
    for it in range(1, n_it + 1):
        moved = tr[it - 1] != prev
        marked = np.zeros_like(moved)
        for e in sched:
            if e["it"] == it:
                marked |= e["chg"]
        assert np.array_equal(moved, marked)
        prev = tr[it - 1]
    for b, last in enumerate(out["iters"]):
        ls = [e for e in sched if "conv" in e and e["conv"][b]]
        assert ls and ls[-1]["it"] == last
        assert not any(e["it"] > last and e.get("act", e.get("conv"))[b] for e in sched)
