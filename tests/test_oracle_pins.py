# -*- coding: utf-8 -*-
"""Pins of the oracle's branches that no closed form elsewhere fixes:

- Newton iteration counts and the status-machine cadence (CheckFreq /
  MaxCheck, convergence freeze) against a brute-force scalar GGA written
  from the EPANET equations (tests/scalar_gga.py, no oracle code);
- tank TEMPCLOSED exactly when the tank is full with inflow or empty with
  outflow;
- CONST_HP pumps (head gain 8.814 P / Q, with the half-step of newflows)
  and NOCURVE pumps against closed forms, pump XHEAD;
- the Darcy-Weisbach transition cubic against the laminar and Swamee-Jain
  laws it joins (value and slope at Re = 2000 and 4000);
- the emitter-coefficient gradient against central differences;
- the fixed relerr reduction tree against a lane-by-lane simulation.
CPU only."""
import math

import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.network import CSMALL, DW_A1, DW_A2, PI_EPANET

import scalar_gga as sg


def _oracle(nodes, links, acc, checkfreq=2, maxcheck=10):
    net = sg.to_net(nodes, links, accuracy=acc, checkfreq=checkfreq, maxcheck=maxcheck)
    m = gga.GGAModel(net)
    rh = sg.fixed_heads(nodes)[None]
    out = gga.solve(m, net.demand_cfs[None], rh, ke=net.node_ke, accuracy=acc,
                    status_machine=True, record_schedule=True, extended=True)
    ev = []
    for e in out["schedule"]:
        if "conv" in e:
            br = "conv" if e["conv"][0] else "per"
            ev += [(e["it"], int(k), br) for k in np.nonzero(e["chg"][0])[0]]
    return net, m, out, ev


def _compare(nodes, links, acc, checkfreq=2, maxcheck=10, tol_h=1e-8):
    bf = sg.solve(nodes, links, acc, checkfreq=checkfreq, maxcheck=maxcheck)
    net, m, out, ev = _oracle(nodes, links, acc, checkfreq, maxcheck)
    assert int(out["iters"][0]) == bf["iters"]
    assert out["status"][0].tolist() == list(bf["status"])
    assert ev == [(it, k, br) for it, k, _, _, br in bf["events"]]
    assert np.abs(out["head_ft"][0] - np.array(bf["head"])).max() < tol_h
    return bf, out


def _res(h):
    return dict(kind="r", head=h)


def _junc(elev, dem, ke=0.0):
    return dict(kind="j", elev=elev, demand=dem, ke=ke)


def _pipe(a, b, length, diam, C, kind="pipe"):
    return dict(kind=kind, n1=a, n2=b, r=sg.hw_r(length, diam, C), diam=diam)


@pytest.mark.parametrize("acc", [1e-3, 1e-6, 1e-10])
def test_one_pipe_iterations(acc):
    nodes = [_res(100.0), _junc(0.0, 2.0)]
    links = [_pipe(0, 1, 1000.0, 1.0, 100.0)]
    bf, out = _compare(nodes, links, acc)
    # converged head: energy equation of the one pipe
    if acc <= 1e-6:
        q = 2.0
        assert abs(out["head_ft"][0, 1] - (100.0 - links[0]["r"] * q ** 1.852)) < 1e-6


@pytest.mark.parametrize("acc", [1e-3, 1e-8])
def test_pipe_and_emitter_iterations(acc):
    nodes = [_res(100.0), _junc(10.0, 0.5, ke=4.0)]
    links = [_pipe(0, 1, 2000.0, 0.8, 110.0)]
    bf, out = _compare(nodes, links, acc)
    # emitter balance at the converged state: e = sqrt((H - El) / Ke)
    if acc <= 1e-8:
        H = out["head_ft"][0, 1]
        e = out["emitter_cfs"][0, 1]
        assert abs(e - math.sqrt((H - 10.0) / 4.0)) < 1e-6


def _cv_case():
    # reservoir R1 (100 ft) feeds J1; a check valve J1 -> R2 (120 ft) would
    # carry reverse flow, so it must close
    nodes = [_res(100.0), _junc(0.0, 1.0), _res(120.0)]
    links = [_pipe(0, 1, 1000.0, 1.0, 100.0), _pipe(1, 2, 500.0, 0.5, 100.0, kind="cv")]
    return nodes, links


def test_check_valve_closes_on_the_checkfreq_cadence():
    nodes, links = _cv_case()
    bf, out = _compare(nodes, links, 1e-6, checkfreq=2)
    # the valve closes at the first periodic check (iteration CheckFreq = 2)
    assert bf["events"][0][:2] == (2, 1) and bf["events"][0][4] == "per"
    assert bf["status"][1] == sg.CLOSED


def test_check_valve_cadence_follows_checkfreq():
    nodes, links = _cv_case()
    bf, _ = _compare(nodes, links, 1e-6, checkfreq=3)
    assert bf["events"][0][0] == 3


def test_check_valve_closes_at_convergence_past_maxcheck():
    # MaxCheck = 1: no periodic check runs, the valve closes on the
    # convergence branch and the solve goes on with it closed
    nodes, links = _cv_case()
    bf, _ = _compare(nodes, links, 1e-6, maxcheck=1)
    assert bf["events"][0][4] == "conv"
    assert bf["iters"] > bf["events"][0][0]


def _tank_case(h_res, level, full):
    hmin, hmax = 100.0, 130.0
    head = hmax if full else hmin
    nodes = [_res(h_res), _junc(0.0, 0.5),
             dict(kind="t", head=head, hmin=hmin, hmax=hmax)]
    links = [_pipe(0, 1, 1000.0, 1.0, 100.0), _pipe(1, 2, 800.0, 0.6, 100.0)]
    return nodes, links


@pytest.mark.parametrize("h_res,full,closed", [(160.0, True, True), (80.0, True, False),
                                               (80.0, False, True), (160.0, False, False)])
def test_tank_tempclosed_when_full_with_inflow_or_empty_with_outflow(h_res, full, closed):
    nodes, links = _tank_case(h_res, 0.0, full)
    bf, out = _compare(nodes, links, 1e-6)
    st = out["status"][0, 1]
    assert (st == sg.TEMPCLOSED) == closed
    q = out["flow_cfs"][0, 1]                        # J -> tank
    if closed:
        # no flow through the closed tank link, the junction hangs off R only
        assert abs(q) < 1e-6
        assert abs(out["head_ft"][0, 1] - (h_res - links[0]["r"] * 0.5 ** 1.852)) < 1e-5
    else:
        # the open link carries the flow the heads drive: into a full tank
        # never, out of an empty one never
        assert (q < 0.0) if full else (q > 0.0)


@pytest.mark.parametrize("power,qd", [(10.0, 1.0), (25.0, 3.0), (5.0, 0.2)])
def test_const_hp_pump_closed_form(power, qd):
    # reservoir -> constant-horsepower pump -> junction with demand qd: the
    # pump delivers qd against a head gain of 8.814 P / qd (hp, cfs -> ft)
    nodes = [_res(100.0), _junc(0.0, qd)]
    links = [dict(kind="pump", n1=0, n2=1, ptype="chp", r=-8.814 * power, q0=1.0,
                  hmax=sg.BIG)]
    bf, out = _compare(nodes, links, 1e-10)
    assert abs(out["head_ft"][0, 1] - (100.0 + 8.814 * power / qd)) < 1e-8
    assert abs(out["flow_cfs"][0, 0] - qd) < 1e-10


def test_const_hp_half_step_is_taken():
    # a 1 hp pump from a 100 ft reservoir into a junction that a 150 ft
    # reservoir also feeds: from q0 = 1 cfs the first Newton step asks for
    # dq > q (a reversal) and newflows halves the pump flow instead
    r_pipe = sg.hw_r(1000.0, 1.0, 100.0)
    nodes = [_res(100.0), _junc(0.0, 0.1), _res(150.0)]
    links = [dict(kind="pump", n1=0, n2=1, ptype="chp", r=-8.814, q0=1.0, hmax=sg.BIG),
             dict(kind="pipe", n1=1, n2=2, r=r_pipe, diam=1.0)]
    net = sg.to_net(nodes, links, accuracy=1e-10)
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs[None], sg.fixed_heads(nodes)[None], accuracy=1e-10,
                    max_iter=1, status_machine=True)
    assert out["flow_cfs"][0, 0] == 0.5
    bf, out = _compare(nodes, links, 1e-10)
    # converged: pump gain 8.814 / q and the pipe's energy equation
    qp, qj = out["flow_cfs"][0]
    hj = out["head_ft"][0, 1]
    assert abs(hj - (100.0 + 8.814 / qp)) < 1e-8
    assert abs((hj - 150.0) - r_pipe * abs(qj) ** 1.852 * np.sign(qj)) < 1e-8
    assert abs(qp - qj - 0.1) < 1e-10


def test_nocurve_pump_closed_form():
    # a pump without a curve is an open link of head loss CSMALL q
    nodes = [_res(100.0), _junc(0.0, 2.0)]
    links = [dict(kind="pump", n1=0, n2=1, ptype="nocurve", q0=1.0, hmax=sg.BIG)]
    bf, out = _compare(nodes, links, 1e-10)
    assert abs(out["head_ft"][0, 1] - (100.0 - CSMALL * 2.0)) < 1e-10


def test_power_pump_xhead_and_linear_pump():
    # a pump of shutoff head 60 ft cannot lift from 0 ft into a 200 ft zone:
    # pumpstatus closes it (XHEAD) and the junction is fed from above
    nodes = [_res(0.0), _junc(0.0, 1.0), _res(200.0)]
    h0, r, n = -60.0, 5.0, 2.0
    links = [dict(kind="pump", n1=0, n2=1, ptype="pf", h0=h0, r=r, n=n, q0=1.0, hmax=60.0),
             _pipe(1, 2, 1000.0, 1.0, 100.0)]
    bf, out = _compare(nodes, links, 1e-8)
    assert out["status"][0, 0] == sg.XHEAD
    # linear curve (N = 1): head gain 60 - 4 q at the delivered flow
    nodes = [_res(0.0), _junc(0.0, 3.0)]
    links = [dict(kind="pump", n1=0, n2=1, ptype="lin", h0=-60.0, r=4.0, n=1.0, q0=1.0,
                  hmax=60.0)]
    bf, out = _compare(nodes, links, 1e-10)
    assert abs(out["head_ft"][0, 1] - (60.0 - 4.0 * 3.0)) < 1e-9


# ---------------------------------------------------------------- D-W
def _dw_f(m, k, w):
    """Friction factor of pipe k at w = |Q| / (nu D): hloss / (r Q |Q|) on a
    pipe without minor loss."""
    s = m.viscos * m.diam[k]
    q = np.zeros(m.L)
    q[k] = w * s
    _, _, hl, _, _ = m.dw_PY(q[None], m.r_hw[None])
    r = m.r_hw[k]
    if w <= DW_A2:
        # laminar branch: hloss = 16 pi s r Q, i.e. f = 64 / Re with Re = 4 w / pi
        return hl[0, k] / (r * q[k] * abs(q[k]))
    return hl[0, k] / (r * q[k] * abs(q[k]))


def test_darcy_weisbach_transition_joins_laminar_and_swamee_jain():
    # The Dunlop cubic between Re = 2000 (w = A2) and Re = 4000 (w = A1)
    # meets 64/Re and Swamee-Jain with equal value and slope at both ends:
    # four conditions that fix a cubic, so a wrong coefficient fails
    from wdsgpu_amd.generators import feature_network
    net = feature_network(0, headloss="D-W", n_prv=0)
    m = gga.GGAModel(net)
    k = int(np.nonzero(np.asarray(net.link_type) == 1)[0][3])
    e = m.kc[k] / m.diam[k]

    def lam(w):
        return 64.0 / (4.0 * w / PI_EPANET)

    def sj(w):
        Re = 4.0 * w / PI_EPANET
        return 0.25 / math.log10(e / 3.7 + 5.74 / Re ** 0.9) ** 2

    eps = 1e-7
    for w0, left, right in ((DW_A2, lam, None), (DW_A1, None, sj)):
        f_in = _dw_f(m, k, w0 * (1 + eps)) if left else _dw_f(m, k, w0 * (1 - eps))
        ref = left(w0) if left else right(w0)
        assert abs(f_in - ref) <= 1e-6 * ref
        # slopes: one-sided differences inside the cubic vs the law's own
        h = 1e-4 * w0
        if left:
            s_in = (_dw_f(m, k, w0 + 2 * h) - _dw_f(m, k, w0 + h)) / h
            s_ref = (left(w0 + 2 * h) - left(w0 + h)) / h
        else:
            s_in = (_dw_f(m, k, w0 - h) - _dw_f(m, k, w0 - 2 * h)) / h
            s_ref = (right(w0 - h) - right(w0 - 2 * h)) / h
        # the cubic meets 64/Re with its slope exactly; at Re = 4000 its slope
        # uses EPANET's rounded constant AC and matches Swamee-Jain's to 0.75 %
        assert abs(s_in - s_ref) <= (2e-3 if left else 1.5e-2) * abs(s_ref)
    # the laminar branch itself is 64 / Re and the turbulent one Swamee-Jain
    assert abs(_dw_f(m, k, 0.5 * DW_A2) - lam(0.5 * DW_A2)) <= 1e-12 * lam(0.5 * DW_A2)
    assert abs(_dw_f(m, k, 20 * DW_A1) - sj(20 * DW_A1)) <= 1e-9 * sj(20 * DW_A1)


# ---------------------------------------------------------------- emitters
def _ke_fd(m, d, rh, ke, r, W, j, rel=1e-3):
    def loss(kk):
        s = adj.solve_polished(m, d, rh, kk, r, accuracy=1e-9)
        return float((s["head"] * W).sum())

    def cd(h):
        p, q_ = ke.copy(), ke.copy()
        p[0, j] += h
        q_[0, j] -= h
        return (loss(p) - loss(q_)) / (2 * h)
    h = rel * ke[0, j]
    return (4.0 * cd(h / 2) - cd(h)) / 3.0


@pytest.mark.parametrize("case", ["feature-H-W", "feature-D-W", "planar"])
def test_emitter_coefficient_gradient_matches_central_differences(case):
    from wdsgpu_amd.generators import feature_network, planar_network, scenario_batch
    if case == "planar":
        net = planar_network(1)
        d, rh = scenario_batch(net, 1, seed=5)
        leaks = np.nonzero(net.node_ke)[0][[0, 4, 9]]
    else:
        net = feature_network(0, headloss=case.split("-", 1)[1], n_prv=1, n_leaks=3)
        d = net.demand_cfs[None].copy()
        rh = np.nan_to_num(net.reservoir_head_ft())[None]
        leaks = np.nonzero(net.node_ke)[0]
    m = gga.GGAModel(net)
    ke = net.node_ke[None].copy()
    g = np.random.default_rng(3)
    W = np.zeros((1, m.N))
    W[0, m.junc_nodes] = g.normal(size=m.Nj)
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=1e-9)
    _, _, gke, _ = adj.cpu_adjoint(m, sol, W, np.zeros((1, m.L)), np.zeros((1, m.N)))
    for j in leaks:
        fd = _ke_fd(m, d, rh, ke, m.r_hw, W, int(j))
        assert abs(gke[0, j] - fd) <= 1e-6 * max(abs(fd), 1e-12), (j, gke[0, j], fd)
    # junctions without an emitter have no Ke gradient
    none = np.setdiff1d(m.junc_nodes, np.nonzero(net.node_ke)[0])
    assert np.all(gke[0, none] == 0.0)


# ---------------------------------------------------------------- relerr tree
def test_relerr_tree_is_the_lane_butterfly():
    g = np.random.default_rng(0)
    for n in (1, 3, 64, 65, 905, 1000):
        x = np.abs(g.normal(size=(2, n))) * 10.0 ** g.integers(-8, 3, size=(2, n))
        got = gga.wave_tree_sum(x)
        for b in range(2):
            lanes = [0.0] * 64
            for i in range(n):
                lanes[i % 64] = lanes[i % 64] + float(x[b, i])
            for mm in (32, 16, 8, 4, 2, 1):
                lanes = [lanes[l] + lanes[l ^ mm] for l in range(64)]
            assert got[b] == lanes[0]
            assert abs(got[b] - math.fsum(x[b])) <= 2 * n * 2.0 ** -53 * math.fsum(x[b])


# ---------------------------------------------------------------- linear solve
def test_sparse_refined_solve_is_correctly_rounded():
    # the oracle's reference solve (extended=True): every entry within half
    # an ulp of the exact solution, which an exact rational residual and
    # one more correction reveal
    from fractions import Fraction
    import scipy.sparse as sps
    from scipy.sparse.linalg import splu
    from wdsgpu_amd.generators import grid_network, scenario_batch
    net = grid_network(0)
    m = gga.GGAModel(net)
    B = 2
    d, rh = scenario_batch(net, B, seed=9)
    rh = np.nan_to_num(rh)
    q = np.broadcast_to(m.init_flow(), (B, m.L)).copy()
    H = m.full_heads(np.zeros((B, m.Nj)), rh[:, m.fixed_nodes])
    S = np.broadcast_to(m.init_status_int, (B, m.L)).copy()
    P, Y = m.link_PY(q, S, m.r_hw[None])
    csr, F, *_ = m.assemble(P, Y, q, H, np.zeros((B, m.Nj)), np.zeros((B, m.Nj)),
                            d[:, m.junc_nodes], S, csr=True)
    x = gga.sparse_solve_refined(m, csr, F)
    for b in range(B):
        A = sps.csc_matrix((csr[b], (m.csr_row, m.csr_col)), shape=(m.Nj, m.Nj))
        r = [Fraction(v) for v in F[b]]
        for k in range(csr.shape[1]):
            r[m.csr_row[k]] -= Fraction(csr[b, k]) * Fraction(x[b, m.csr_col[k]])
        delta = splu(A).solve(np.array([float(v) for v in r]))
        assert np.all(np.abs(delta) <= 0.5000001 * np.spacing(np.abs(x[b])))
    # the dense Cholesky path with plain refinement lands close by
    xd = gga.chol_solve_refined(m.csr_to_dense(csr), F)
    assert np.all(np.abs(xd - x) <= 1e-12 * np.abs(x))


def test_solve_jitter_separates_ill_conditioned_trajectories():
    # The oracle's one-ulp solve jitter (accuracy 1e-6, so the final heads
    # carry the trajectory's rounding): on a network without PRVs the final
    # heads move by about 1e-10 ft; on the ragged-batch member whose third
    # Newton system has kappa ~ 2e15 (a PRV closed, a junction held by
    # 1/CBIG links) by several 1e-9 ft: no float64 solver determines that
    # trajectory to 1e-10 m, and the ragged GPU test holds it to 3x this
    # motion instead
    from wdsgpu_amd.generators import grid_network, ragged_batch, scenario_batch
    cases = []
    member, B = ragged_batch(0)[58]
    for net, B, seed in ((grid_network(0), 2, 5), (member, B, 158)):
        m = gga.GGAModel(net)
        d, rh = scenario_batch(net, B, seed=seed)
        d, rh = d[[0, B - 1]], np.nan_to_num(rh[[0, B - 1]])
        base = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, extended=True)
        m.solve_jitter = np.random.default_rng(7)
        jit = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, extended=True)
        m.solve_jitter = None
        cases.append(np.abs(jit["head_ft"] - base["head_ft"]).max())
    print("1-ulp solve jitter of the final heads (ft): grid %.2e, ragged member 58 %.2e"
          % tuple(cases))
    assert cases[0] < 1e-9
    assert cases[1] > 2e-9
