# -*- coding: utf-8 -*-
"""The differentiable batched EPS on the GPU (wdsgpu_amd.eps_adjoint):
heads of every step and the discrete-adjoint gradients against the oracle
(oracle/eps_adjoint.py, itself pinned by finite differences), the same bits
for every checkpoint interval (the recomputed steps reproduce the forward),
and gradients against finite differences of the GPU forward on a
2000-junction network with tanks, pumps and pump controls."""
import numpy as np
import pytest
import torch

from oracle import eps_adjoint as ea
from wdsgpu_amd.eps_adjoint import EpsModel
from wdsgpu_amd.generators import eps_network, geometric_network

pytestmark = pytest.mark.gpu

TOL_H_FT = 1e-10 / 0.3048


def _inputs(net, B, seed):
    g = np.random.default_rng(seed)
    dem = net.demand_cfs * np.exp(0.2 * g.standard_normal((B, net.N)))
    rh = np.nan_to_num(net.reservoir_head_ft())
    h0 = np.broadcast_to(net.tank_h0, (B, net.tank_node.size)).copy()
    return dem, rh, net.node_ke.copy(), net.r_hw.copy(), h0


def _gpu_grads(model, inputs, W, obs):
    dev = model.device
    leaves = [torch.tensor(x, device=dev, requires_grad=True) for x in inputs]
    H = model.simulate(*leaves, observe=torch.tensor(obs))
    loss = (torch.tensor(W, device=dev) * H).sum()
    loss.backward()
    return H.detach(), [x.grad for x in leaves]


@pytest.mark.parametrize("which,T,B", [("geometric120", 8, 3), ("eps6x6", 12, 2)])
def test_heads_and_gradients_match_oracle(cuda, which, T, B):
    net = geometric_network(2, n_junc=120, n_res=2, n_tanks=2, n_pumps=2, n_leaks=3) \
        if which == "geometric120" else eps_network(0)
    inputs = _inputs(net, B, 1)
    obs = np.arange(net.N)
    g = np.random.default_rng(2)
    W = np.zeros((T, B, net.N))
    pick = np.concatenate([g.choice(net.Nj, 6, replace=False), net.tank_node])
    W[:, :, pick] = g.normal(size=(T, B, pick.size))
    m = EpsModel(net, device=cuda, steps=T, accuracy=1e-10, checkpoint_every=4)
    m.record_decisions = True
    H, grads = _gpu_grads(m, inputs, W, obs)
    fw = ea.forward(net, *inputs, T, accuracy=1e-10)
    S_in, clamp = (x.cpu().numpy() for x in m.decisions)
    assert np.array_equal(S_in, fw["S_in"]) and np.array_equal(clamp, fw["clamp"])
    dh = np.abs(H.cpu().numpy() - fw["H"]).max()
    ref = ea.adjoint(net, fw, W)
    rel = []
    for name, a, b in zip(("demand", "res_head", "ke", "r", "tank_h0"), grads, ref):
        a = a.cpu().numpy()
        b = b if a.ndim == 2 else b.sum(0)
        rel.append((name, float(np.abs(a - b).max() / np.abs(b).max())))
    print("%s T=%d B=%d: max|dH| %.2e ft, status changes %d, clamps %d, gradient rel %s"
          % (which, T, B, dh, int((np.diff(S_in.astype(int), axis=0) != 0).sum()),
             int(clamp.sum()), ["%s %.1e" % x for x in rel]), flush=True)
    assert dh <= TOL_H_FT
    for name, r in rel:
        assert r < 1e-9, (name, r)


def test_checkpoint_interval_keeps_the_bits(cuda):
    net = geometric_network(2, n_junc=120, n_res=2, n_tanks=2, n_pumps=2, n_leaks=3)
    T, B = 9, 3
    inputs = _inputs(net, B, 5)
    obs = np.concatenate([np.arange(0, net.Nj, 9), net.tank_node])
    W = np.random.default_rng(6).normal(size=(T, B, obs.size))
    res = []
    for k in (1, 3, 4, 9):
        m = EpsModel(net, device=cuda, steps=T, accuracy=1e-8, checkpoint_every=k)
        res.append(_gpu_grads(m, inputs, W, obs))
        # forward solves T, recomputed T - ceil(T / k)
        assert m.solves == T + T - -(-T // k), (k, m.solves)
    for H, grads in res[1:]:
        assert torch.equal(H, res[0][0])
        for a, b in zip(grads, res[0][1]):
            assert torch.equal(a, b)


def _fd(f, h, base, levels=5):
    """Richardson-extrapolated central differences within the base run's
    discrete piece (statuses and clamps unchanged), steps h, h/4, ...; the
    estimate whose successor differs least."""
    est = []
    for _ in range(levels):
        runs = [f(x) for x in (h, -h, h / 2, -h / 2)]
        if all(torch.equal(s, base[0]) and torch.equal(c, base[1]) for _, (s, c) in runs):
            (a, _), (b, _), (c, _), (d, _) = runs
            est.append((4 * (c - d) / h - (a - b) / (2 * h)) / 3)
        h /= 4
    assert len(est) >= 2, "no perturbation keeps the discrete decisions"
    k = min(range(len(est) - 1), key=lambda i: abs(est[i] - est[i + 1]))
    return est[k]


def test_gradients_match_finite_differences_at_2000_junctions(cuda):
    net = geometric_network(3, n_junc=2000, n_res=4, n_tanks=6, n_pumps=8)
    T, B = 8, 4
    dem, rh, ke, r, h0 = _inputs(net, B, 7)
    g = np.random.default_rng(8)
    obs = np.concatenate([g.choice(net.Nj, 30, replace=False), net.tank_node])
    W = g.normal(size=(T, B, obs.size))
    m = EpsModel(net, device=cuda, steps=T, accuracy=1e-10, checkpoint_every=4)
    m.record_decisions = True
    _, grads = _gpu_grads(m, (dem, rh, ke, r, h0), W, obs)
    base = m.decisions
    Wt = torch.tensor(W, device=cuda)
    obs_t = torch.tensor(obs)

    def loss(dem=dem, rh=rh, ke=ke, r=r, h0=h0):
        with torch.no_grad():
            H = m.simulate(*(torch.tensor(x, device=cuda) for x in (dem, rh, ke, r, h0)),
                           observe=obs_t)
        return float((Wt * H).sum()), m.decisions

    j = int(obs[0])
    k = int(np.argmax(np.where(np.asarray(net.link_type) == 1, net.len_ft, 0.0)))
    res = int(np.nonzero(np.asarray(net.node_type) == 1)[0][0])
    lk = int(np.nonzero(ke > 0)[0][0])
    e_d = np.zeros_like(dem)
    e_d[1, j] = 1.0
    e_h = np.zeros_like(h0)
    e_h[2, 3] = 1.0
    checks = [
        ("demand", grads[0][1, j], lambda h: loss(dem=dem + e_d * h), 1e-2 * dem[1, j]),
        ("r", grads[3][k], lambda h: loss(r=r + np.eye(1, net.L, k)[0] * h), 1e-2 * r[k]),
        ("res_head", grads[1][res], lambda h: loss(rh=rh + np.eye(1, net.N, res)[0] * h),
         1e-3 * rh[res]),
        ("ke", grads[2][lk], lambda h: loss(ke=ke + np.eye(1, net.N, lk)[0] * h),
         1e-2 * ke[lk]),
        ("tank_h0", grads[4][2, 3], lambda h: loss(h0=h0 + e_h * h), 1e-3 * h0[2, 3])]
    changes = int((np.diff(base[0].cpu().numpy().astype(int), axis=0) != 0).sum())
    print("2000 junctions, T=%d B=%d: status changes %d, clamps %d"
          % (T, B, changes, int(base[1].sum())), flush=True)
    for name, ga, f, h in checks:
        ga = float(ga)
        fd = _fd(f, h, base)
        rel = abs(ga - fd) / abs(fd)
        print("  %s: adjoint %.10e fd %.10e rel %.1e" % (name, ga, fd, rel), flush=True)
        assert rel < 1e-6, (name, ga, fd)
