# -*- coding: utf-8 -*-
"""GPU forward parity on networks with every link and node kind of the
path: power-function pumps (open and XHEAD), a tank (open and full), TCVs,
check-valve pipes (open and closed), PRVs, leak emitters, Hazen-Williams and
Darcy-Weisbach friction."""
import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd.generators import feature_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

from netutil import FEATURE_CASES as CASES, check_forward_against_oracle

pytestmark = pytest.mark.gpu



@pytest.mark.parametrize("headloss,seed,level", CASES)
def test_feature_network_matches_oracle(cuda, headloss, seed, level):
    net = feature_network(seed, headloss=headloss, n_prv=2, tank_level_m=level)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    B = 6
    d, rh = scenario_batch(net, B, seed=11)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-3, status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, 1e-3)



# CUSTOM multi-point pump curves: both pumps OPEN (H-W 0, D-W 5) or one OPEN
# and one XHEAD (H-W 6, D-W 7)
@pytest.mark.parametrize("headloss,seed", [("H-W", 0), ("H-W", 6), ("D-W", 5), ("D-W", 7)])
def test_custom_pump_network_matches_oracle(cuda, headloss, seed):
    net = feature_network(seed, headloss=headloss, n_prv=2, n_pumps=2, n_custom=2)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 6, seed=11)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-3, status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, 1e-3)


def test_custom_pump_segments_closed_form(cuda):
    # one scenario per operating point: on the first point, inside every
    # segment, on an inner point and beyond the last point
    from netutil import curve_lift, custom_pump_tree
    pts = [(2.0, 60.0), (6.0, 50.0), (10.0, 32.0), (14.0, 5.0)]
    qs = np.array([2.0, 4.0, 6.0, 8.0, 12.0, 15.0])
    net = custom_pump_tree(pts, 1.0)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d = np.zeros((len(qs), 2))
    d[:, 1] = qs
    rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), d.shape).copy()
    out = s.solve(torch.tensor(d), torch.tensor(rh), accuracy=1e-10, status_machine=True)
    H = out["head_ft"].cpu().numpy()
    for b, q in enumerate(qs):
        assert abs(H[b, 1] - (100.0 + curve_lift(pts, q))) < 1e-8
    check_forward_against_oracle(net, m, out, d, rh, 1e-10)


def _same_schedule(got, ref):
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g.keys() == r.keys()
        for k in r:
            if isinstance(r[k], np.ndarray):
                assert np.array_equal(g[k], r[k]), (r["it"], k)
            else:
                assert g[k] == r[k]


@pytest.mark.parametrize("headloss,seed,cap", [("H-W", 0, None), ("H-W", 2, None),
                                               ("D-W", 3, None), ("H-W", 1, 1)])
def test_record_schedule_matches_oracle(cuda, headloss, seed, cap):
    # prvstatus and linkstatus events of the batched status machine; cap 1
    # overflows the first log and takes the re-run with a larger one
    net = feature_network(seed, headloss=headloss, n_prv=2)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    s.sched_cap = cap
    d, rh = scenario_batch(net, 6, seed=11)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-3, status_machine=True, record_schedule=True)
    ref = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-3, status_machine=True,
                    record_schedule=True)
    assert out["iters"].cpu().numpy().tolist() == ref["iters"].tolist()
    assert any(e["chg"].any() for e in ref["schedule"])
    _same_schedule(out["schedule"], ref["schedule"])


def test_record_schedule_planar_prvs(cuda):
    from wdsgpu_amd.generators import planar_network
    net = planar_network(1)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 4, seed=4)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-3, status_machine=True, record_schedule=True)
    ref = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-3, status_machine=True,
                    record_schedule=True)
    _same_schedule(out["schedule"], ref["schedule"])
