# -*- coding: utf-8 -*-
"""The benchmark's own settings: the 782-junction planar network with
1024 lognormal demand scenarios at accuracy 1e-6 (forward), and the
calibration workload at accuracy 1e-6 with 3 polish steps (forward +
implicit adjoint), against the oracle and central differences."""
import time

import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from oracle import gga
from wdsgpu_amd.generators import planar_network, scenario_batch
from wdsgpu_amd.network import emitter_ke_dC
from wdsgpu_amd.solver import GGASolver

pytestmark = pytest.mark.gpu

TOL_H_FT = 1e-10 / 0.3048


def test_full_batch_forward_matches_oracle(cuda):
    # every one of the 1024 scenarios: identical Newton iteration counts and
    # final statuses, heads within 1e-10 m of the oracle's correctly rounded
    # iterate
    net = planar_network(1)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    B = 1024
    d, rh = scenario_batch(net, B, seed=4)
    rh = np.nan_to_num(rh)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-6, status_machine=True)
    it_g = out["iters"].cpu().numpy()
    t0 = time.time()
    ext = gga.solve(m, d, rh, ke=net.node_ke, accuracy=1e-6, status_machine=True,
                    extended=True)
    print("oracle: %d scenarios in %.1f s" % (B, time.time() - t0), flush=True)
    assert it_g.tolist() == ext["iters"].tolist()
    assert np.array_equal(out["status"].cpu().numpy(), ext["status"])
    dh = np.abs(out["head_ft"].cpu().numpy() - ext["head_ft"]).max()
    dq = np.abs(out["flow_cfs"].cpu().numpy() - ext["flow_cfs"]).max()
    print("B=1024 accuracy 1e-6: iterations %d..%d, max|dH| %.3e ft (%.3e m), max|dQ| %.3e cfs"
          % (it_g.min(), it_g.max(), dh, dh * 0.3048, dq), flush=True)
    assert dh <= TOL_H_FT
    assert bool(out["converged"].all())


@pytest.fixture(scope="module")
def calib(cuda):
    from wdsgpu_amd.workloads import Calibration
    return Calibration(B=16, device=cuda)


def _gpu_grads(wl):
    loss, gd, grh, gC, gCe = wl.step()
    return (float(loss), gd.cpu().numpy(), grh.cpu().numpy(), gC.cpu().numpy(),
            gCe.cpu().numpy())


def test_calibration_gradients_match_oracle_adjoint(calib):
    # 16 scenarios at the benchmark's accuracy 1e-6 and 3 polish steps:
    # gradients with respect to demands, pipe C, reservoir heads and leak C
    # within 1e-10 (relative to each kind's largest) of the oracle's
    # full-Jacobian SuperLU adjoint, chained through r(C) and Ke(C)
    wl = calib
    net, s = wl.net, wl.solver
    m = gga.GGAModel(net)
    _, gd, grh, gC, gCe = _gpu_grads(wl)
    B = wl.B
    d = wl.d0.cpu().numpy()
    rh = np.broadcast_to(wl.rh0.cpu().numpy(), (B, net.N)).copy()
    with torch.no_grad():
        r = wl.r_of(wl.C0).cpu().numpy()
        ke1 = wl.ke_of(wl.Ce0).cpu().numpy()
    ke = np.broadcast_to(ke1, (B, net.N)).copy()
    sol = adj.solve_polished(m, d, rh, ke, r, accuracy=wl.accuracy, polish_steps=3)
    sens = wl.sensors.cpu().numpy()
    gH = np.zeros((B, net.N))
    gH[:, sens] = 2.0 * (sol["head"][:, sens] - wl.H_true.cpu().numpy())
    cd, crh, cke, cr = adj.cpu_adjoint(m, sol, gH, np.zeros((B, net.L)), np.zeros((B, net.N)))
    C0 = wl.C0.cpu().numpy()
    Ce0 = wl.Ce0.cpu().numpy()
    pipe = np.asarray(net.link_type) == 1
    c_gC = np.where(pipe, cr.sum(0) * (-1.852 * r / C0), 0.0)
    c_gCe = cke.sum(0) * emitter_ke_dC(Ce0, ke1, 2.0)
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    errs = dict(demand=rel(gd, cd), res_head=rel(grh, crh.sum(0)), C=rel(gC, c_gC),
                leak_C=rel(gCe, c_gCe))
    print("GPU vs oracle adjoint (relative to each kind's largest):", errs, flush=True)
    assert all(v < 1e-10 for v in errs.values()), errs


def test_calibration_gradients_match_central_differences(cuda):
    # 80 seeded coordinates of the planar network (30 demands, 36 pipe C,
    # both reservoir heads, 10 leak C, 2 more demands) through the
    # 30-sensor loss: Richardson-extrapolated central differences of the
    # polished forward at relative steps 1e-4 and 5e-5
    from wdsgpu_amd.autodiff import implicit_solve
    from wdsgpu_amd.workloads import Calibration
    wl = Calibration(B=4, device=cuda)
    net = wl.net
    loss0, gd, grh, gC, gCe = _gpu_grads(wl)
    d0, rh0 = wl.d0.clone(), wl.rh0.clone()
    C0, Ce0 = wl.C0.clone(), wl.Ce0.clone()

    def loss(d, rh, C, Ce):
        with torch.no_grad():
            return float(wl.loss(d, rh, C, Ce))

    g = np.random.default_rng(21)
    jn = np.nonzero(np.asarray(net.node_type) == 0)[0]
    pipes = np.nonzero(np.asarray(net.link_type) == 1)[0]
    leaks = np.nonzero(Ce0.cpu().numpy() > 0)[0]
    fixed = np.nonzero(np.asarray(net.node_type) != 0)[0]
    coords = [("d", int(b), int(j)) for b, j in zip(g.integers(0, 4, 32),
                                                  g.choice(jn, 32, replace=False))]
    coords += [("C", 0, int(k)) for k in g.choice(pipes, 36, replace=False)]
    coords += [("h", 0, int(f)) for f in fixed] + [("e", 0, int(j)) for j in leaks]
    assert len(coords) == 80
    an = {"d": gd, "C": gC, "h": grh, "e": gCe}
    scale = {k: np.abs(v).max() for k, v in an.items()}

    def f_at(kind, b, j, h):
        d, rh, C, Ce = d0.clone(), rh0.clone(), C0.clone(), Ce0.clone()
        x = {"d": d, "h": rh, "C": C, "e": Ce}[kind]
        if kind == "d":
            x[b, j] += h
        else:
            x[j] += h
        return loss(d, rh, C, Ce)

    worst = 0.0
    for kind, b, j in coords:
        x0 = float((d0[b, j] if kind == "d" else {"h": rh0, "C": C0, "e": Ce0}[kind][j]))
        h = 1e-4 * abs(x0)

        def cdiff(hh):
            return (f_at(kind, b, j, hh) - f_at(kind, b, j, -hh)) / (2 * hh)
        fd = (4.0 * cdiff(h / 2) - cdiff(h)) / 3.0
        a = float(an[kind][b, j] if kind == "d" else an[kind][j])
        err = abs(a - fd) / max(abs(fd), 1e-3 * scale[kind])
        worst = max(worst, err)
        assert err < 1e-6, (kind, b, j, a, fd)
    print("80 coordinates: worst relative error vs central differences %.3e" % worst,
          flush=True)
