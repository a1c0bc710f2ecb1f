# -*- coding: utf-8 -*-
"""GPU runs of the cases that pin the oracle (tests/test_oracle_pins.py)
checked directly against the brute-force scalar GGA: Newton iteration
counts, the iteration and branch of every status change, final statuses
and heads; and the 23 networks the reference ships, solved on the GPU
against the oracle."""
import os

import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd.network import read_inp
from wdsgpu_amd.solver import GGASolver

import scalar_gga as sg
from netutil import check_forward_against_oracle, res_heads
from test_oracle_pins import _cv_case, _junc, _pipe, _res, _tank_case

pytestmark = pytest.mark.gpu


def _cases():
    out = []
    nodes, links = [_res(100.0), _junc(0.0, 2.0)], [_pipe(0, 1, 1000.0, 1.0, 100.0)]
    out += [("one-pipe-%g" % a, nodes, links, a, 2, 10) for a in (1e-3, 1e-6, 1e-10)]
    out.append(("emitter", [_res(100.0), _junc(10.0, 0.5, ke=4.0)],
                [_pipe(0, 1, 2000.0, 0.8, 110.0)], 1e-8, 2, 10))
    nodes, links = _cv_case()
    out += [("cv-cf2", nodes, links, 1e-6, 2, 10), ("cv-cf3", nodes, links, 1e-6, 3, 10),
            ("cv-maxcheck1", nodes, links, 1e-6, 2, 1)]
    for h, full in ((160.0, True), (80.0, True), (80.0, False), (160.0, False)):
        n, lk = _tank_case(h, 0.0, full)
        out.append(("tank-%g-%s" % (h, full), n, lk, 1e-6, 2, 10))
    for P, qd in ((10.0, 1.0), (5.0, 0.2)):
        out.append(("chp-%g" % P, [_res(100.0), _junc(0.0, qd)],
                    [dict(kind="pump", n1=0, n2=1, ptype="chp", r=-8.814 * P, q0=1.0,
                          hmax=sg.BIG)], 1e-10, 2, 10))
    out.append(("chp-half-step", [_res(100.0), _junc(0.0, 0.1), _res(150.0)],
                [dict(kind="pump", n1=0, n2=1, ptype="chp", r=-8.814, q0=1.0, hmax=sg.BIG),
                 dict(kind="pipe", n1=1, n2=2, r=sg.hw_r(1000.0, 1.0, 100.0), diam=1.0)],
                1e-10, 2, 10))
    out.append(("nocurve", [_res(100.0), _junc(0.0, 2.0)],
                [dict(kind="pump", n1=0, n2=1, ptype="nocurve", q0=1.0, hmax=sg.BIG)],
                1e-10, 2, 10))
    out.append(("pump-xhead", [_res(0.0), _junc(0.0, 1.0), _res(200.0)],
                [dict(kind="pump", n1=0, n2=1, ptype="pf", h0=-60.0, r=5.0, n=2.0, q0=1.0,
                      hmax=60.0), _pipe(1, 2, 1000.0, 1.0, 100.0)], 1e-8, 2, 10))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,nodes,links,acc,cf,mc", CASES, ids=[c[0] for c in CASES])
def test_tiny_networks_match_brute_force(cuda, name, nodes, links, acc, cf, mc):
    bf = sg.solve(nodes, links, acc, checkfreq=cf, maxcheck=mc)
    net = sg.to_net(nodes, links, accuracy=acc, checkfreq=cf, maxcheck=mc)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    rh = sg.fixed_heads(nodes)
    out = s.solve(torch.tensor(net.demand_cfs[None]), torch.tensor(rh[None]),
                  ke_int=torch.tensor(net.node_ke), accuracy=acc, status_machine=True,
                  record_schedule=True)
    print(name, "iters", int(out["iters"][0]), "brute force", bf["iters"], flush=True)
    assert int(out["iters"][0]) == bf["iters"]
    assert out["status"][0].cpu().numpy().tolist() == list(bf["status"])
    ev = []
    for e in out["schedule"]:
        if "conv" in e:
            br = "conv" if e["conv"][0] else "per"
            ev += [(e["it"], int(k), br) for k in np.nonzero(e["chg"][0])[0]]
    assert ev == [(it, k, br) for it, k, _, _, br in bf["events"]]
    assert np.abs(out["head_ft"][0].cpu().numpy() - np.array(bf["head"])).max() < 1e-8


NETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "networks")
SHIPPED = ["random_main/rand_%04d.inp" % i for i in range(20)] + \
    ["random_small/rand_%04d.inp" % i for i in range(3)]


@pytest.mark.parametrize("rel", SHIPPED)
def test_shipped_networks_match_oracle(cuda, rel):
    # each network at its own INP options (accuracy 1e-8, 500 trials), four
    # demand scalings per launch
    net = read_inp(os.path.join(NETS, rel))
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    sc = np.array([0.8, 1.0, 1.15, 1.3])
    d = net.demand_cfs[None] * sc[:, None]
    rh = np.broadcast_to(res_heads(net), d.shape).copy()
    acc = float(net.meta["accuracy"])
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=acc, status_machine=True)
    print(rel, "iters", out["iters"].tolist(), flush=True)
    check_forward_against_oracle(net, m, out, d, rh, acc)
