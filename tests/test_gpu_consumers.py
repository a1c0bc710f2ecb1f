# -*- coding: utf-8 -*-
"""The calibration consumers on the HIP path against the oracle: the
sensitivity matrix (one forward, one multi-right-hand-side adjoint) against
the oracle's SuperLU adjoint per sensor, the kernels' link coefficients
against the oracle's, the roughness parametrisation through the implicit
adjoint against central differences, multistart batches against separate
calls, and identifiability on a network with a dead branch and a closed
pipe."""
import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd import calib, optim2, sensitivity
from wdsgpu_amd.autodiff import implicit_solve, solve_polished
from wdsgpu_amd.generators import planar_network, scenario_batch, sensor_nodes
from wdsgpu_amd.network import JUNCTION, RESERVOIR
from wdsgpu_amd.solver import GGASolver

from netutil import make_net

pytestmark = pytest.mark.gpu
# the planar network's relerr levels off near 1e-10: the GGA stopping test
# before the polish (the reference's calibration practice)
ACC = 1e-8


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def planar(cuda):
    net = planar_network(1)
    return net, gga.GGAModel(net), GGASolver(net, device=cuda, dense_status_machine=True)


def test_sensitivity_matrix_matches_oracle_adjoint_per_sensor(planar):
    net, m, s = planar
    B = 3
    d, rh = scenario_batch(net, B, seed=21)
    sens = sensor_nodes(net, 6, 5)
    print("sensitivity: planar782 B=%d sensors %s" % (B, sens.tolist()), flush=True)
    S, info = sensitivity.sensitivity_matrix(s, d, rh, None, sens, wrt="r", frames=[0, 2],
                                             accuracy=ACC, return_info=True)
    assert S.shape == (2 * 6, net.L)
    assert info["n_rhs"] == 12 and info["frames"] == [0, 2]
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=ACC)
    for i, b in enumerate([0, 2]):
        for j, node in enumerate(sens):
            W = np.zeros((B, net.N))
            W[b, node] = 1.0
            _, _, _, gr = adj.cpu_adjoint(m, sol, W, np.zeros((B, net.L)), np.zeros((B, net.N)))
            assert _rel(S[i * 6 + j], gr[b]) < 1e-10, (b, node)
    SC = sensitivity.sensitivity_matrix(s, d, rh, None, sens, wrt="C", frames=[0, 2],
                                        accuracy=ACC)
    chain = calib.dr_dC(s)
    assert np.array_equal(SC, S * chain[None, :])


def test_sensitivity_equals_autograd_rows(planar):
    net, m, s = planar
    d, rh = scenario_batch(net, 2, seed=22)
    sens = sensor_nodes(net, 3, 6)
    S = sensitivity.sensitivity_matrix(s, d, rh, None, sens, wrt="r", accuracy=ACC)
    for b in range(2):
        for j, node in enumerate(sens):
            rt = s.r_hw.clone().requires_grad_(True)
            H, _, _ = implicit_solve(s, torch.tensor(d[b:b + 1], device=s.device),
                                     torch.tensor(rh[b:b + 1], device=s.device), r_hw=rt,
                                     accuracy=ACC)
            H[0, node].backward()
            assert _rel(S[b * 3 + j], rt.grad.cpu().numpy()) < 1e-12


def test_link_coefficients_match_oracle(planar):
    net, m, s = planar
    d, rh = scenario_batch(net, 2, seed=23)
    ke = np.broadcast_to(net.node_ke, (2, net.N)).copy()
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=ACC)
    q = sol["q"].copy()
    q[:, :20] = 1e-9                  # some pipes on the RQtol clamp
    P, hl, dr = (x.cpu().numpy() for x in s.link_coeffs(q, None, sol["status"]))
    for b in range(2):
        ohl, ohg, odr, act = adj.link_coeffs(m, q[b], m.r_hw, sol["status"][b])
        keep = ~act
        assert np.array_equal(dr[b], odr)
        assert np.allclose(P[b][keep], 1.0 / ohg[keep], rtol=1e-15, atol=0)
        assert np.allclose(hl[b][keep], ohl[keep], rtol=1e-14, atol=1e-300)
        assert (P[b][act] == 0.0).all()
        # the clamp branch is exercised (and agrees, above)
        assert (odr[:20] == 0.0).sum() >= 10


def test_schur_precond_is_the_jacobi_formula(planar):
    net, m, s = planar
    d, rh = scenario_batch(net, 2, seed=24)
    st = solve_polished(s, d, rh, accuracy=ACC)
    free = np.nonzero(np.asarray(net.link_type) == 1)[0][:50]
    C = np.asarray(net.roughness)[free]
    M, info = optim2.schur_diag_precond(s, st["q"], st["r_hw"], C, free, floor_rel=0.0,
                                        status=st["status"][0])
    # restated from the oracle's link coefficients
    jidx = np.full(net.N, -1)
    jidx[m.junc_nodes] = np.arange(m.Nj)
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
    acc = np.zeros(free.size)
    for b in range(2):
        _, hg, dr, act = adj.link_coeffs(m, st["q"][b], st["r_hw"], st["status"][0])
        w = np.where(act, 0.0, 1.0 / np.where(act, 1.0, hg))
        A = np.zeros(m.Nj)
        for k in range(net.L):
            for n in (n1[k], n2[k]):
                if jidx[n] >= 0:
                    A[jidx[n]] += w[k]
        for i, k in enumerate(free):
            pre = (1.852 * st["r_hw"][k] / C[i]) * (dr[k] / hg[k])
            t = sum(1.0 / A[jidx[n]] ** 2 for n in (n1[k], n2[k]) if jidx[n] >= 0)
            acc[i] += pre ** 2 * t
    assert _rel(M, acc / 2) < 1e-12
    assert info["kind"] == "schur_jacobi" and np.isfinite(info["cond"])


def test_roughness_parameter_gradient_matches_central_differences(planar):
    net, m, s = planar
    d, rh = scenario_batch(net, 2, seed=25)
    free = np.nonzero(np.asarray(net.link_type) == 1)[0][::60]
    p = calib.RoughnessParam(s, pipe_idx=free)
    r0 = p.r_hw().detach().cpu().numpy()
    base = s.r_hw.cpu().numpy()
    assert np.abs(r0 - base).max() <= 4e-16 * np.abs(base).max()
    sens = torch.as_tensor(sensor_nodes(net, 10, 7), device=s.device)
    dt = torch.tensor(d, device=s.device)
    rht = torch.tensor(rh, device=s.device)
    # a linear functional of the sensor heads: central differences of it
    # carry no large constant to cancel
    w = torch.tensor(np.random.default_rng(8).normal(size=(2, 10)), device=s.device)

    def loss(theta=None):
        if theta is not None:
            with torch.no_grad():
                p.theta.copy_(theta)
        H, _, _ = implicit_solve(s, dt, rht, r_hw=p.r_hw(), accuracy=ACC)
        return (H.index_select(1, sens) * w).sum()

    th0 = p.theta.detach().clone()
    L0 = loss()
    L0.backward()
    g = p.theta.grad.cpu().numpy()
    with torch.no_grad():
        for i in range(0, free.size, 3):
            h = 1e-5 * max(1.0, abs(float(th0[i])))
            e = torch.zeros_like(th0)
            e[i] = h
            fd = (float(loss(th0 + e)) - float(loss(th0 - e))) / (2 * h)
            assert abs(g[i] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, g[i], fd)
        p.theta.copy_(th0)


def test_multistart_batch_equals_separate_calls(planar):
    net, m, s = planar
    d, rh = scenario_batch(net, 3, seed=26)
    rng = np.random.default_rng(3)
    base = s.r_hw.cpu().numpy()
    r_starts = torch.tensor(base[None, :] * rng.uniform(0.8, 1.2, (2, net.L)),
                            device=s.device, requires_grad=True)
    H, Q = optim2.multistart_heads(s, r_starts, d, rh, accuracy=1e-6, max_iter=50)
    assert H.shape == (2, 3, net.N) and Q.shape == (2, 3, net.L)
    w = torch.tensor(rng.normal(size=(2, 3, net.N)), device=s.device)
    (H * w).sum().backward()
    for b in range(2):
        rb = r_starts[b].detach().clone().requires_grad_(True)
        Hb, _, _ = implicit_solve(s, torch.tensor(d, device=s.device),
                                  torch.tensor(rh, device=s.device), r_hw=rb.expand(3, -1),
                                  accuracy=1e-6, max_iter=50, polish_steps=4, adjoint="gpu")
        assert torch.equal(Hb, H[b].detach())
        (Hb * w[b]).sum().backward()
        assert _rel(r_starts.grad[b].cpu().numpy(), rb.grad.cpu().numpy()) < 1e-13


def test_identifiability_attributes_zero_columns(cuda):
    nodes = [(RESERVOIR, 100.0, 0.0), (JUNCTION, 10.0, 0.1), (JUNCTION, 10.0, 0.1),
             (JUNCTION, 10.0, 0.1), (JUNCTION, 10.0, 0.0), (JUNCTION, 10.0, 0.0),
             (JUNCTION, 10.0, 0.1)]
    pipe = lambda a, b, **kw: dict(n1=a, n2=b, len_ft=500.0, diam_ft=0.5, C=100.0, **kw)
    links = [pipe(0, 1), pipe(1, 2), pipe(2, 3), pipe(3, 1), pipe(3, 4), pipe(4, 5),
             pipe(1, 6, status=0), pipe(2, 6)]
    net = make_net(nodes, links)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    rh = np.where(np.asarray(net.node_type) == RESERVOIR, 100.0, 0.0)
    # the dead chain's zero flows keep relerr near 1e-7: stop the GGA at 1e-6
    # and let the polish converge
    out = calib.identifiability(s, net.demand_cfs, rh, [2, 6], accuracy=1e-6, return_S=True)
    reason = out["reason"].tolist()
    assert reason[4] == "dead_branch" and reason[5] == "dead_branch"
    assert reason[6] == "closed"
    assert reason[0] == "informative" and reason[7] == "informative"
    assert out["counts"]["dead_branch"] == 2 and out["counts"]["closed"] == 1
    assert out["S"].shape == (2, net.L)
    assert np.all(out["S"][:, 4:6] == 0.0) and np.all(out["S"][:, 6] == 0.0)


def test_leak_inversion_training_loop_lowers_the_loss(cuda):
    from wdsgpu_amd.workloads import LeakInversion
    wl = LeakInversion(B=32, device=cuda)
    losses = [float(wl.step()) for _ in range(40)]
    print("leak inversion: loss %.4e -> %.4e, top-3 %s, true %s"
          % (losses[0], losses[-1], wl.ranking()[:3].tolist(), wl.true_leaks.tolist()),
          flush=True)
    assert losses[-1] < 0.2 * losses[0]
    assert set(wl.ranking()[:5].tolist()) & set(wl.true_leaks.tolist())
    assert wl.solver.counters()["bwd_refactorize"] == 0
