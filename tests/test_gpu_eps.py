# -*- coding: utf-8 -*-
"""Batched extended-period simulation on the GPU against the oracle's
scalar driver with correctly rounded linear solves: identical frame times,
statuses and Newton iteration counts per scenario, heads within 1e-10 m and
flows within 1e-9 cfs at every frame (both paths share the power laws, so
the tank levels they integrate agree)."""
import numpy as np
import pytest

from oracle.eps import run_eps
from wdsgpu_amd.eps import EpsDriver
from wdsgpu_amd.generators import eps_network

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 3])
def test_batched_eps_matches_oracle(cuda, seed):
    net = eps_network(seed)
    scales = np.array([0.9, 1.0, 1.12])
    frames = EpsDriver(net, device=cuda).run_batch(len(scales), demand_scale=scales)
    for b, sc in enumerate(scales):
        ref = run_eps(net, demand_scale=np.full(net.N, sc), extended=True)
        f = frames[b]
        assert f["t_sec"].tolist() == ref["t_sec"].tolist()
        assert np.array_equal(f["status"], ref["status"])
        assert f["iterations"].tolist() == ref["iterations"].tolist()
        dh = np.abs(f["head_ft"] - ref["head_ft"]).max()
        dq = np.abs(f["flow_cfs"] - ref["flow_cfs"]).max()
        print("seed %d scale %.2f: %d frames, max|dH| %.3e ft, max|dQ| %.3e cfs"
              % (seed, sc, len(f["t_sec"]), dh, dq), flush=True)
        assert dh <= 1e-10 / 0.3048
        assert dq <= 1e-9


def test_single_scenario_run_matches_batch_member(cuda):
    net = eps_network(1)
    drv = EpsDriver(net, device=cuda)
    one = drv.run()
    many = drv.run_batch(2)
    assert one["t_sec"].tolist() == many[1]["t_sec"].tolist()
    assert np.array_equal(one["head_ft"], many[1]["head_ft"])
