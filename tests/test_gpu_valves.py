# -*- coding: utf-8 -*-
"""GPU forward parity on networks with pressure-sustaining (PSV) and
flow-control (FCV) valves beside PRVs, pumps, a tank, TCVs and check valves:
identical iteration counts, statuses (PSV ACTIVE / OPEN / CLOSED, FCV ACTIVE
/ XFCV among them) and status-machine schedules, heads within 1e-10 m of the
oracle; and the polish and implicit adjoint, which are built for PRVs only,
refuse these valves."""
import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd.autodiff import implicit_solve
from wdsgpu_amd.generators import feature_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

from netutil import check_forward_against_oracle
from test_gpu_features import _same_schedule

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("headloss,seed,acc", [("H-W", s, 1e-3) for s in range(8)] +
                         [("H-W", 4, 1e-6), ("H-W", 6, 1e-6), ("D-W", 1, 1e-3),
                          ("D-W", 5, 1e-6)])
def test_valve_network_matches_oracle(cuda, headloss, seed, acc):
    print("valves: feature seed %d %s acc %g" % (seed, headloss, acc), flush=True)
    net = feature_network(seed, headloss=headloss, n_psv=2, n_fcv=2)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 6, seed=11)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=acc, status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, acc)


@pytest.mark.parametrize("seed", [0, 5, 6])
def test_valve_schedule_matches_oracle(cuda, seed):
    net = feature_network(seed, n_psv=2, n_fcv=2)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 6, seed=11)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-3, status_machine=True, record_schedule=True)
    ref = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-3, status_machine=True,
                    record_schedule=True)
    assert out["iters"].cpu().numpy().tolist() == ref["iters"].tolist()
    _same_schedule(out["schedule"], ref["schedule"])


def test_polish_and_adjoint_refuse_psv_and_fcv(cuda):
    net = feature_network(0, n_psv=1, n_fcv=1)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 2, seed=11)
    dt = torch.tensor(d, device=cuda, requires_grad=True)
    with pytest.raises(NotImplementedError):
        implicit_solve(s, dt, torch.tensor(np.nan_to_num(rh), device=cuda), accuracy=1e-6)
