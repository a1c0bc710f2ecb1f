# -*- coding: utf-8 -*-
"""Batched extended-period simulation on the GPU against the oracle's
scalar driver: identical frame times, statuses and Newton iteration counts
per scenario; heads and flows within 3x the oracle trajectory's own change
under a one-ulp perturbation of its power laws, or 1e-10 m.

Tank levels integrate every frame's flows, so the rounding of each solve
accumulates along the trajectory; the perturbed oracle run measures how far
(about 1e-10 ft on seed 0, about 1e-6 ft on seed 3, whose pump cycles on and
off)."""
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
        ref = run_eps(net, demand_scale=np.full(net.N, sc))
        f = frames[b]
        assert f["t_sec"].tolist() == ref["t_sec"].tolist()
        assert np.array_equal(f["status"], ref["status"])
        assert f["iterations"].tolist() == ref["iterations"].tolist()
        jit = run_eps(net, demand_scale=np.full(net.N, sc),
                      pow_jitter=np.random.default_rng(123))
        assert jit["t_sec"].tolist() == ref["t_sec"].tolist()
        assert jit["iterations"].tolist() == ref["iterations"].tolist()
        hn = np.abs(jit["head_ft"] - ref["head_ft"]).max()
        qn = np.abs(jit["flow_cfs"] - ref["flow_cfs"]).max()
        assert np.abs(f["head_ft"] - ref["head_ft"]).max() <= max(3.0 * hn, 1e-10 / 0.3048)
        assert np.abs(f["flow_cfs"] - ref["flow_cfs"]).max() <= max(3.0 * qn, 1e-12)


def test_single_scenario_run_matches_batch_member(cuda):
    net = eps_network(1)
    drv = EpsDriver(net, device=cuda)
    one = drv.run()
    many = drv.run_batch(2)
    assert one["t_sec"].tolist() == many[1]["t_sec"].tolist()
    assert np.array_equal(one["head_ft"], many[1]["head_ft"])
