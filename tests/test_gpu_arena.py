# -*- coding: utf-8 -*-
"""Arena mode: the scenario kernels with each scenario's factor and vectors
in a global-memory arena instead of LDS (networks larger than a CU's LDS).
The same programs run on the same byte offsets, so on a network that fits
both ways the two modes give the same bits; a 5000-junction network, which
only fits the arena, is checked against the oracle."""
import time

import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.autodiff import implicit_solve
from wdsgpu_amd.generators import planar_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

pytestmark = pytest.mark.gpu

TOL_H_FT = 1e-10 / 0.3048


def _grads(s, d, rh, ke, r, W):
    dt = torch.tensor(d, device=s.device, requires_grad=True)
    rht = torch.tensor(rh, device=s.device, requires_grad=True)
    ket = torch.tensor(ke, device=s.device, requires_grad=True)
    rt = torch.tensor(r, device=s.device, requires_grad=True)
    H, Q, E = implicit_solve(s, dt, rht, ke=ket, r_hw=rt, accuracy=1e-8, polish_steps=3,
                             adjoint="gpu", status_machine=True)
    (H * torch.tensor(W, device=s.device)).sum().backward()
    return [H.detach(), Q.detach(), E.detach(), dt.grad, rht.grad, ket.grad, rt.grad]


def test_arena_mode_gives_the_lds_modes_bits(cuda):
    net = planar_network(1)
    lds = GGASolver(net, device=cuda, dense_status_machine=True)
    are = GGASolver(net, device=cuda, dense_status_machine=True, arena=True)
    assert lds.stats["arena_bytes"] == 0 and are.stats["arena_bytes"] > 0
    B = 8
    d, rh = scenario_batch(net, B, seed=5)
    ke = torch.tensor(net.node_ke)
    a = lds.solve(torch.tensor(d), torch.tensor(rh), ke_int=ke, accuracy=1e-6,
                  status_machine=True)
    b = are.solve(torch.tensor(d), torch.tensor(rh), ke_int=ke, accuracy=1e-6,
                  status_machine=True)
    for k in ("head_ft", "flow_cfs", "emitter_cfs", "iters", "status", "relerr"):
        assert torch.equal(a[k], b[k]), k
    u = are.solve_unfused(torch.tensor(d), torch.tensor(rh), ke_int=ke, accuracy=1e-6,
                          status_machine=True)
    assert torch.equal(u["head_ft"], a["head_ft"])
    m = gga.GGAModel(net)
    W = np.random.default_rng(2).normal(size=(B, net.N))
    kb = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    ga = _grads(lds, d, rh, kb, m.r_hw.copy(), W)
    gb = _grads(are, d, rh, kb, m.r_hw.copy(), W)
    for x, y in zip(ga, gb):
        assert torch.equal(x, y)
    # the factor and the solve phases on their own
    q = np.broadcast_to(m.init_flow(), (B, m.L)).copy()
    H0 = m.full_heads(np.zeros((B, m.Nj)), rh[:, m.fixed_nodes])
    S = np.broadcast_to(m.init_status_int, (B, m.L)).copy()
    P, Y = m.link_PY(q, S, m.r_hw[None])
    csr, F, *_ = m.assemble(P, Y, q, H0, np.where(kb[:, m.junc_nodes] > 0, 1.0, 0.0),
                            kb[:, m.junc_nodes], d[:, m.junc_nodes], S, csr=True)
    fa, fb = lds.factor(torch.tensor(csr)), are.factor(torch.tensor(csr))
    assert torch.equal(fa, fb)
    assert torch.equal(lds.trisolve(fa, torch.tensor(csr), torch.tensor(F)),
                       are.trisolve(fb, torch.tensor(csr), torch.tensor(F)))


@pytest.fixture(scope="module")
def big(cuda):
    net = planar_network(3, n_junc=5000, n_pipes=5787)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    return net, gga.GGAModel(net), s


def test_5000_junction_network_matches_oracle(big):
    net, m, s = big
    assert s.stats["lds_bytes"] > 160 * 1024 and s.stats["arena_bytes"] > 0
    B = 6
    d, rh = scenario_batch(net, B, seed=7)
    t0 = time.time()
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-6, status_machine=True)
    torch.cuda.synchronize()
    t1 = time.time() - t0
    ext = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, status_machine=True,
                    extended=True)
    assert out["iters"].cpu().numpy().tolist() == ext["iters"].tolist()
    assert np.array_equal(out["status"].cpu().numpy(), ext["status"])
    dh = np.abs(out["head_ft"].cpu().numpy() - ext["head_ft"]).max()
    print("Nj=5000 B=%d: %.1f ms, iterations %s, max|dH| %.3e ft"
          % (B, 1e3 * t1, out["iters"].tolist(), dh), flush=True)
    assert dh <= TOL_H_FT


def test_5000_junction_adjoint_matches_oracle(big):
    net, m, s = big
    B = 2
    d, rh = scenario_batch(net, B, seed=8)
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    W = np.zeros((B, net.N))
    W[:, m.junc_nodes[::50]] = 1.0
    H, Q, E, gd, grh, gke, gr = (x.cpu().numpy() for x in _grads(s, d, rh, ke, m.r_hw.copy(), W))
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=1e-8, linsolve="sparse")
    assert np.abs(H - sol["head"]).max() <= TOL_H_FT
    cd, crh, cke, cr = adj.cpu_adjoint(m, sol, W, np.zeros((B, net.L)), np.zeros((B, net.N)))
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    errs = (rel(gd, cd), rel(grh, crh), rel(gke, cke), rel(gr, cr.sum(0)))
    print("Nj=5000 adjoint vs oracle:", errs, flush=True)
    # the oracle's SuperLU solve of the 10.8k-unknown Jacobian is itself
    # accurate to about 1e-10 here (no refinement), which bounds the match
    assert max(errs) < 1e-9
