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

