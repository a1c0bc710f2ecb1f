# -*- coding: utf-8 -*-
"""Emitter rows in the Newton update: a scenario whose lanes hold at most two
emitter rows each reads them from its per-scenario list, one with more walks
every row; both give the oracle's iteration counts, statuses and heads, and a
scenario's bits do not depend on the emitters of the others in its batch."""
import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd.generators import planar_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

from netutil import check_forward_against_oracle

pytestmark = pytest.mark.gpu


def _ke(net, rows, value=0.02):
    ke = np.zeros(net.N)
    junc = np.nonzero(np.asarray(net.node_type) == 0)[0]
    ke[junc[rows]] = value
    return ke


@pytest.mark.parametrize("which", ["none", "few", "shared_residue", "every"])
def test_emitter_rows_match_oracle(cuda, which):
    net = planar_network(1)
    Nj = int((np.asarray(net.node_type) == 0).sum())
    rows = {"none": [], "few": [3, 70, 200, 511],
            # three rows of residue 5 mod 64: more than the list holds
            "shared_residue": [5, 69, 133, 300],
            "every": list(range(Nj))}[which]
    print("emitters: planar782 %s (%d rows)" % (which, len(rows)), flush=True)
    net.node_ke = _ke(net, np.asarray(rows, dtype=np.int64))
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 4, seed=31)
    rh = np.nan_to_num(rh)
    kb = np.broadcast_to(net.node_ke, d.shape).copy()
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(kb), accuracy=1e-6,
                  status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, 1e-6)


def test_emitter_list_is_per_scenario(cuda):
    # scenario 0 (three emitters: the list) beside a scenario with an emitter
    # on every junction (the row walk) gives the bits it gives alone
    net = planar_network(1)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 2, seed=32)
    rh = np.nan_to_num(rh)
    Nj = int((np.asarray(net.node_type) == 0).sum())
    few = _ke(net, np.array([3, 70, 200]))
    many = _ke(net, np.arange(Nj))
    a = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(np.stack([few, many])),
                accuracy=1e-6, status_machine=True)
    b = s.solve(torch.tensor(d[:1]), torch.tensor(rh[:1]), ke_int=torch.tensor(few[None]),
                accuracy=1e-6, status_machine=True)
    assert torch.equal(a["head_ft"][:1], b["head_ft"])
    assert torch.equal(a["emitter_cfs"][:1], b["emitter_cfs"])
    assert torch.equal(a["iters"][:1], b["iters"])
