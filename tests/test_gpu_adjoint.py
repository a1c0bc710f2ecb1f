# -*- coding: utf-8 -*-
"""GPU implicit adjoint against the oracle's full-Jacobian SuperLU adjoint
and against central finite differences."""
import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.autodiff import implicit_solve
from wdsgpu_amd.generators import grid_network, planar_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def planar(cuda):
    net = planar_network(1)
    return net, gga.GGAModel(net), GGASolver(net, device=cuda, dense_status_machine=True)


def _gpu_grads(s, d, rh, ke, r, W, WQ, acc=1e-8, polish=3):
    dt = torch.tensor(d, device=s.device, requires_grad=True)
    rht = torch.tensor(rh, device=s.device, requires_grad=True)
    ket = torch.tensor(ke, device=s.device, requires_grad=True)
    rt = torch.tensor(r, device=s.device, requires_grad=True)
    H, Q, E = implicit_solve(s, dt, rht, ke=ket, r_hw=rt, accuracy=acc, polish_steps=polish,
                             adjoint="gpu", status_machine=True)
    loss = (H * torch.tensor(W, device=s.device)).sum() + \
        (Q * torch.tensor(WQ, device=s.device)).sum() + (E * 0.5).sum()
    loss.backward()
    return (H.detach().cpu().numpy(), Q.detach().cpu().numpy(), dt.grad.cpu().numpy(),
            rht.grad.cpu().numpy(), ket.grad.cpu().numpy(), rt.grad.cpu().numpy(),
            float(loss.detach()))


def test_adjoint_matches_cpu_reference_adjoint(planar):
    net, m, s = planar
    B = 3
    d, rh = scenario_batch(net, B, seed=11)
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    r = m.r_hw.copy()
    g = np.random.default_rng(1)
    W = g.normal(size=(B, net.N))
    WQ = g.normal(size=(B, net.L)) * 0.1
    H, Q, gd, grh, gke, gr, _ = _gpu_grads(s, d, rh, ke, r, W, WQ)
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=1e-8)
    assert np.abs(H - sol["head"]).max() < 1e-9
    cd, crh, cke, cr = adj.cpu_adjoint(m, sol, W, WQ, np.full((B, net.N), 0.5))
    assert _rel(gd, cd) < 1e-10
    assert _rel(grh, crh.sum(0)) < 1e-10 if grh.ndim == 1 else _rel(grh, crh) < 1e-10
    assert _rel(gke, cke) < 1e-10
    assert _rel(gr, cr.sum(0)) < 1e-10


@pytest.mark.parametrize("headloss,seed,n_custom", [("H-W", 0, 0), ("H-W", 2, 0),
                                                     ("H-W", 0, 1), ("D-W", 0, 0),
                                                     ("D-W", 2, 0)])
def test_adjoint_with_pumps_matches_cpu_reference_adjoint(cuda, headloss, seed, n_custom):
    # feature network: power-function pump OPEN (seed 0) / XHEAD (seed 2) or
    # a CUSTOM pump OPEN inside a segment, ACTIVE PRV branches, a TCV, a
    # check valve, a tank and leak emitters; Darcy-Weisbach networks hold
    # turbulent, transitional and laminar pipes
    from wdsgpu_amd.generators import feature_network
    net = feature_network(seed, headloss=headloss, n_prv=2, n_custom=n_custom)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    B = 3
    d, rh = scenario_batch(net, B, seed=13)
    rh = np.nan_to_num(rh)
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    r = m.r_hw.copy()
    g = np.random.default_rng(2)
    W = g.normal(size=(B, net.N))
    WQ = g.normal(size=(B, net.L)) * 0.1
    H, Q, gd, grh, gke, gr, _ = _gpu_grads(s, d, rh, ke, r, W, WQ)
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=1e-8)
    pump = int(m.pump_links[0])
    assert (sol["status"][:, pump] == (3 if seed == 0 else 0)).all()
    assert np.abs(H - sol["head"]).max() < 1e-9
    cd, crh, cke, cr = adj.cpu_adjoint(m, sol, W, WQ, np.full((B, net.N), 0.5))
    assert _rel(gd, cd) < 1e-10
    assert _rel(grh, crh) < 1e-10
    assert _rel(gke, cke) < 1e-10
    assert _rel(gr, cr.sum(0)) < 1e-10


def test_adjoint_counters_reuse_the_forward_factor(planar):
    net, m, s = planar
    s.counters(reset=True)
    d, rh = scenario_batch(net, 2, seed=12)
    _gpu_grads(s, d, rh, np.broadcast_to(net.node_ke, (2, net.N)).copy(), m.r_hw.copy(),
               np.ones((2, net.N)), np.zeros((2, net.L)))
    c = s.counters()
    assert c["bwd_reuse"] == 1 and c["bwd_refactorize"] == 0


def test_adjoint_matches_central_differences(cuda):
    # grid network, 80 seeded coordinates (demands, pipe r, reservoir head),
    # Richardson-extrapolated central differences of the polished GPU forward
    net = grid_network(0)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    m = gga.GGAModel(net)
    d = net.demand_cfs[None].copy()
    rh = np.nan_to_num(net.reservoir_head_ft())[None]
    ke = net.node_ke[None].copy()
    r = m.r_hw.copy()
    g = np.random.default_rng(7)
    W = np.zeros((1, net.N))
    W[0, m.junc_nodes] = g.normal(size=m.Nj)
    WQ = np.zeros((1, net.L))
    _, _, gd, grh, gke, gr, _ = _gpu_grads(s, d, rh, ke, r, W, WQ, acc=1e-9)

    def loss(dd, rr, hh):
        with torch.no_grad():
            H, Q, E = implicit_solve(s, torch.tensor(dd, device=cuda), torch.tensor(hh, device=cuda),
                                     ke=torch.tensor(ke, device=cuda), r_hw=torch.tensor(rr, device=cuda),
                                     accuracy=1e-9, polish_steps=3, adjoint="gpu",
                                     status_machine=True)
            return float((H.cpu().numpy() * W).sum() + 0.5 * E.cpu().numpy().sum())

    def cd(kind, j, rel):
        def f(h):
            if kind == "d":
                p, q = d.copy(), d.copy()
                p[0, j] += h
                q[0, j] -= h
                return (loss(p, r, rh) - loss(q, r, rh)) / (2 * h)
            if kind == "r":
                p, q = r.copy(), r.copy()
                p[j] += h
                q[j] -= h
                return (loss(d, p, rh) - loss(d, q, rh)) / (2 * h)
            p, q = rh.copy(), rh.copy()
            p[0, j] += h
            q[0, j] -= h
            return (loss(d, r, p) - loss(d, r, q)) / (2 * h)
        x0 = d[0, j] if kind == "d" else (r[j] if kind == "r" else rh[0, j])
        h = max(rel * abs(x0), 1e-5)
        return (4 * f(h / 2) - f(h)) / 3.0
    coords = [("d", int(j)) for j in g.choice(m.junc_nodes, 40, replace=False)]
    coords += [("r", int(k)) for k in g.choice(net.L, 39, replace=False)]
    coords += [("h", int(m.fixed_nodes[0]))]
    # relative error against max(|fd|, 1e-3 * largest gradient of that kind)
    scale = {"d": np.abs(gd).max(), "r": np.abs(gr).max(), "h": np.abs(grh).max()}
    worst = 0.0
    for kind, j in coords:
        fd = cd(kind, j, 1e-3)
        an = gd[0, j] if kind == "d" else (gr[j] if kind == "r" else grh[0, j])
        err = abs(an - fd) / max(abs(fd), 1e-3 * scale[kind])
        worst = max(worst, err)
        assert err < 1e-6, (kind, j, an, fd)
    assert worst < 1e-6


@pytest.mark.parametrize("seed", [2, 7])
def test_adjoint_on_other_planar_topologies(cuda, seed):
    # other elimination trees, so other mixed-group factor and solve
    # programs under the polish, the Woodbury paths and the node adjoint
    net = planar_network(seed)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    B = 2
    d, rh = scenario_batch(net, B, seed=17)
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    r = m.r_hw.copy()
    g = np.random.default_rng(3)
    W = g.normal(size=(B, net.N))
    WQ = g.normal(size=(B, net.L)) * 0.1
    H, Q, gd, grh, gke, gr, _ = _gpu_grads(s, d, rh, ke, r, W, WQ)
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=1e-8)
    assert np.abs(H - sol["head"]).max() < 1e-9
    cd, crh, cke, cr = adj.cpu_adjoint(m, sol, W, WQ, np.full((B, net.N), 0.5))
    assert _rel(gd, cd) < 1e-10
    assert _rel(gke, cke) < 1e-10
    assert _rel(gr, cr.sum(0)) < 1e-10


def test_unconverged_forward_raises_now_or_at_backward(planar):
    # without a backward to come the forward checks at once; with one, the
    # check waits for the backward's synchronisation and raises there
    net, m, s = planar
    d, rh = scenario_batch(net, 2, seed=12)
    args = (torch.tensor(d, device=s.device), torch.tensor(rh, device=s.device))
    with torch.no_grad(), pytest.raises(RuntimeError, match="did not converge"):
        implicit_solve(s, *args, accuracy=1e-8, max_iter=1, adjoint="gpu",
                       status_machine=True)
    dt = args[0].clone().requires_grad_(True)
    H, Q, E = implicit_solve(s, dt, args[1], accuracy=1e-8, max_iter=1, adjoint="gpu",
                             status_machine=True)
    with pytest.raises(RuntimeError, match="did not converge"):
        H.sum().backward()


def test_lazy_checks_raise_at_the_next_call_or_check(cuda):
    # lazy_checks: a step's status is checked behind an event; an
    # unconverged forward surfaces at check() (or the next call), a good
    # step leaves nothing behind
    net = planar_network(1)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    s.lazy_checks = True
    d, rh = scenario_batch(net, 2, seed=12)
    rht = torch.tensor(rh, device=cuda)

    def step(max_iter):
        dt = torch.tensor(d, device=cuda, requires_grad=True)
        H, Q, E = implicit_solve(s, dt, rht, accuracy=1e-8, max_iter=max_iter,
                                 adjoint="gpu", status_machine=True)
        H.sum().backward()
        return dt.grad

    g = step(200)
    s.check()
    assert torch.isfinite(g).all()
    step(1)
    with pytest.raises(RuntimeError, match="did not converge"):
        s.check()
    s.check()          # nothing left queued
