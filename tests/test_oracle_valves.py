# -*- coding: utf-8 -*-
"""Pins of the oracle's pressure-sustaining (PSV) and flow-control (FCV)
valves: closed forms of small series networks (an ACTIVE PSV holds its
upstream head, an OPEN PSV does not throttle, an ACTIVE FCV passes exactly
its setting, an FCV whose head reverses goes XFCV and passes less), and on
the generated networks the converged states: junction mass balance, the
valves' own equations and a status machine at its fixed point.  CPU only."""
import numpy as np
import pytest

from oracle import gga
from wdsgpu_amd.generators import feature_network
from wdsgpu_amd.network import (FCV, HEXP_HW, JUNCTION, PIPE, PSV, RESERVOIR, ST_ACTIVE,
                                ST_CLOSED, ST_OPEN, ST_XFCV)

from netutil import make_net, res_heads

PSI_PER_FT = 0.4333   # Ucf[PRESSURE] of make_net's US units


def _series(valve, H0=200.0, H3=50.0, d1=0.2, d2=0.3, Ca=100.0, Cb=100.0):
    """reservoir (H0) -> pipe a -> J1 -> valve -> J2 -> pipe b -> reservoir (H3)."""
    nodes = [(RESERVOIR, H0, 0.0), (JUNCTION, 0.0, d1), (JUNCTION, 0.0, d2),
             (RESERVOIR, H3, 0.0)]
    links = [dict(n1=0, n2=1, len_ft=3000.0, diam_ft=0.5, C=Ca),
             dict(n1=1, n2=2, diam_ft=0.5, status=2, **valve),
             dict(n1=2, n2=3, len_ft=3000.0, diam_ft=0.5, C=Cb)]
    return make_net(nodes, links, accuracy=1e-10)


def _solve(net, acc=1e-10):
    m = gga.GGAModel(net)
    out = gga.solve(m, net.demand_cfs, res_heads(net), accuracy=acc, status_machine=True)
    return m, out


def _hw(r, q):
    return r * abs(q) ** HEXP_HW * np.sign(q)


def test_active_psv_holds_its_upstream_head():
    # unthrottled, J1 would sit near the middle of 200 and 50 ft: a 150 ft
    # setting makes the PSV throttle
    net = _series(dict(type=PSV, setting=150.0 * PSI_PER_FT))
    m, out = _solve(net)
    H, Q = out["head_ft"], out["flow_cfs"]
    assert out["status"][1] == ST_ACTIVE
    assert abs(H[1] - 150.0) < 1e-6                # big-M pin of the upstream row
    ra, rb = net.r_hw[0], net.r_hw[2]
    qa = ((200.0 - 150.0) / ra) ** (1.0 / HEXP_HW)  # pipe a carries what 50 ft drive
    assert abs(Q[0] - qa) < 1e-7
    assert abs(Q[1] - (qa - 0.2)) < 1e-7          # the valve takes the rest of J1's inflow
    assert abs(Q[2] - (qa - 0.5)) < 1e-7
    assert abs(H[2] - (50.0 + _hw(rb, Q[2]))) < 1e-6


def test_psv_below_its_setting_stays_open():
    net = _series(dict(type=PSV, setting=20.0 * PSI_PER_FT))
    m, out = _solve(net)
    H, Q = out["head_ft"], out["flow_cfs"]
    assert out["status"][1] == ST_OPEN
    assert H[1] > 20.0
    # open with no minor loss: the valve is a 1/CSMALL conductance, the two
    # pipes share the 150 ft drop
    ra, rb = net.r_hw[0], net.r_hw[2]
    assert abs(_hw(ra, Q[0]) + _hw(rb, Q[2]) + (H[1] - H[2]) - 150.0) < 1e-8
    assert abs(H[1] - H[2]) < 1e-5


def test_psv_against_reversed_flow_closes():
    # the downstream reservoir is the higher one: flow would run backwards
    # through the valve, so it closes and each side feeds from its own end
    net = _series(dict(type=PSV, setting=20.0 * PSI_PER_FT), H0=50.0, H3=200.0)
    m, out = _solve(net)
    H, Q = out["head_ft"], out["flow_cfs"]
    assert out["status"][1] == ST_CLOSED
    # a closed link keeps the 1/CBIG conductance: a leak of dh / CBIG
    assert abs(Q[1]) <= abs(H[1] - H[2]) / 1e8 + 1e-12
    assert abs(Q[0] - (0.2 + Q[1])) < 1e-12 and abs(Q[2] - (Q[1] - 0.3)) < 1e-12
    ra, rb = net.r_hw[0], net.r_hw[2]
    assert abs(H[1] - (50.0 - _hw(ra, Q[0]))) < 1e-8
    assert abs(H[2] - (200.0 + _hw(rb, Q[2]))) < 1e-8


def test_active_fcv_passes_its_setting():
    net = _series(dict(type=FCV, setting=0.5), d1=0.0, d2=0.2)
    m, out = _solve(net)
    H, Q = out["head_ft"], out["flow_cfs"]
    assert out["status"][1] == ST_ACTIVE
    dh = H[1] - H[2]
    assert abs(Q[1] - 0.5) <= abs(dh) / 1e8 + 1e-12   # q = setting + dh / CBIG
    ra, rb = net.r_hw[0], net.r_hw[2]
    assert abs(H[1] - (200.0 - _hw(ra, Q[0]))) < 1e-8
    assert abs(Q[2] - (Q[1] - 0.2)) < 1e-10
    assert abs(H[2] - (50.0 + _hw(rb, Q[2]))) < 1e-8


def test_fcv_with_a_reversed_head_goes_xfcv():
    # asking for more flow than 150 ft can drive reverses the valve's head
    net = _series(dict(type=FCV, setting=4.0), d1=0.0, d2=0.2)
    m, out = _solve(net)
    assert out["status"][1] == ST_XFCV
    assert out["flow_cfs"][1] < 4.0
    # XFCV acts as an open valve: the heads either side nearly meet
    assert abs(out["head_ft"][1] - out["head_ft"][2]) < 1e-4


# (the relerr of these networks levels off above 1e-10 because an ACTIVE
# FCV's flow carries dh / CBIG: they are solved to 1e-6)
VALVE_SEEDS = tuple(range(8))


@pytest.mark.parametrize("seed", VALVE_SEEDS)
def test_generated_valve_networks_converge_to_consistent_states(seed):
    net = feature_network(seed, n_psv=2, n_fcv=2)
    m = gga.GGAModel(net)
    d = np.stack([net.demand_cfs * f for f in (0.8, 1.0, 1.3)])
    rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), d.shape).copy()
    ke = np.broadcast_to(net.node_ke, d.shape)
    out = gga.solve(m, d, rh, ke=ke, accuracy=1e-6, status_machine=True)
    assert bool(out["converged"].all())
    H, Q, S, E = out["head_ft"], out["flow_cfs"], out["status"], out["emitter_cfs"]
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
    for b in range(3):
        # junction mass balance: inflow - outflow = demand + emitter flow, to
        # the dh / CBIG by which an ACTIVE FCV's flow exceeds the setting the
        # node equations carry (well inside the 1e-5 cfs of dgga/CONTRACT.md)
        net_in = np.zeros(net.N)
        np.add.at(net_in, n2, Q[b])
        np.add.at(net_in, n1, -Q[b])
        j = m.junc_nodes
        assert np.abs(net_in[j] - d[b, j] - E[b, j]).max() < 1e-6
        for i, k in enumerate(m.prv_k):
            kind, s = int(m.prv_kind[i]), int(S[b, k])
            if kind == 1 and s == ST_ACTIVE:
                assert abs(H[b, n1[k]] - m.prv_hset[i]) < 1e-5
            if kind == 2 and s == ST_ACTIVE:
                assert abs(Q[b, k] - m.prv_hset[i]) <= abs(H[b, n1[k]] - H[b, n2[k]]) / 1e8 + 1e-12
            if s == ST_CLOSED:
                assert abs(Q[b, k]) <= abs(H[b, n1[k]] - H[b, n2[k]]) / 1e8 + 1e-12
    # the status machine rests at the converged state
    S2, chg = m.prvstatus(S, H, Q)
    assert not chg.any()
    S3, chg3 = m.linkstatus(S, H, Q)
    assert np.array_equal(S3[:, m.prv_k], S[:, m.prv_k])


def test_valve_statuses_cover_the_machine():
    seen = set()
    for seed in VALVE_SEEDS:
        net = feature_network(seed, n_psv=2, n_fcv=2)
        m = gga.GGAModel(net)
        d = net.demand_cfs[None, :]
        rh = np.nan_to_num(net.reservoir_head_ft())[None, :]
        out = gga.solve(m, d, rh, ke=net.node_ke[None, :], accuracy=1e-6, status_machine=True)
        for i, k in enumerate(m.prv_k):
            seen.add((int(m.prv_kind[i]), int(out["status"][0, k])))
    # PRV ACTIVE and OPEN, PSV ACTIVE, OPEN and CLOSED, FCV ACTIVE and XFCV
    assert {(0, ST_ACTIVE), (0, ST_OPEN), (1, ST_ACTIVE), (1, ST_OPEN), (1, ST_CLOSED),
            (2, ST_ACTIVE), (2, ST_XFCV)} <= seen
