# -*- coding: utf-8 -*-
"""Oracle of the differentiable fixed-step EPS (oracle/eps_adjoint.py):
its discrete adjoint against Richardson-extrapolated central differences of
the whole multi-step loss, and its forward against the event-driven driver
(oracle/eps.py run_eps) where that driver takes whole-hour steps only.
CPU only."""
import numpy as np
import pytest

from oracle import eps_adjoint as ea
from oracle.eps import run_eps
from wdsgpu_amd.generators import eps_network, geometric_network


def _small_geometric():
    return geometric_network(2, n_junc=120, n_res=2, n_tanks=2, n_pumps=2, n_leaks=3)


def _setup(net, B, seed):
    g = np.random.default_rng(seed)
    dem = net.demand_cfs * np.exp(0.2 * g.standard_normal((B, net.N)))
    rh = np.nan_to_num(net.reservoir_head_ft())
    h0 = np.broadcast_to(net.tank_h0, (B, net.tank_node.size)).copy()
    return dem, rh, net.node_ke.copy(), net.r_hw.copy(), h0


def _same_piece(fw, base):
    return np.array_equal(fw["S_in"], base["S_in"]) and np.array_equal(fw["clamp"],
                                                                        base["clamp"])


def _fd(f, h, base, levels=6):
    """Central differences of the loss within the forward's current piece
    (every status input and tank clamp as in the base run: the model is
    piecewise smooth in its discrete decisions), Richardson-extrapolated
    over steps h, h/4, h/16, ...; returns the estimate whose successor
    differs least from it (the step where truncation and rounding balance).
    f(h) -> (loss, forward dict)."""
    cache = {}

    def ev(x):
        if x not in cache:
            cache[x] = f(x)
        return cache[x]

    est = []
    for _ in range(levels):
        runs = [ev(x) for x in (h, -h, h / 2, -h / 2)]
        if all(_same_piece(fw, base) for _, fw in runs):
            (a, _), (b, _), (c, _), (d, _) = runs
            est.append((4 * (c - d) / h - (a - b) / (2 * h)) / 3)
        h /= 4
    assert len(est) >= 2, "no perturbation keeps the discrete decisions"
    k = min(range(len(est) - 1), key=lambda i: abs(est[i] - est[i + 1]))
    return est[k]


@pytest.mark.parametrize("which,T", [("geometric", 8), ("eps6x6", 12)])
def test_discrete_adjoint_matches_finite_differences(which, T):
    net = _small_geometric() if which == "geometric" else eps_network(0)
    B = 2
    dem, rh, ke, r, h0 = _setup(net, B, 1)
    g = np.random.default_rng(2)
    obs = np.concatenate([g.choice(net.Nj, 5, replace=False), net.tank_node])
    W = np.zeros((T, B, net.N))
    W[:, :, obs] = g.normal(size=(T, B, obs.size))

    def loss(dem=dem, rh=rh, ke=ke, r=r, h0=h0):
        fw = ea.forward(net, dem, rh, ke, r, h0, T, accuracy=1e-10)
        return float((W * fw["H"]).sum()), fw

    _, fw = loss()
    gd, grh, gke, gr, gh0 = ea.adjoint(net, fw, W)
    j = int(obs[0])
    k = int(np.nonzero(r > 0)[0][3])
    res = int(np.nonzero(np.asarray(net.node_type) == 1)[0][-1])
    lk = int(np.nonzero(ke > 0)[0][0])
    e_d = np.zeros((B, net.N))
    e_d[0, j] = 1.0
    e_h = np.zeros_like(h0)
    e_h[1, 0] = 1.0
    checks = [
        ("demand", gd[0, j], lambda h: loss(dem=dem + e_d * h), 1e-2 * dem[0, j]),
        ("r", gr[:, k].sum(), lambda h: loss(r=r + np.eye(1, net.L, k)[0] * h), 1e-2 * r[k]),
        ("res_head", grh[:, res].sum(),
         lambda h: loss(rh=rh + np.eye(1, net.N, res)[0] * h), 1e-2 * rh[res]),
        ("ke", gke[:, lk].sum(), lambda h: loss(ke=ke + np.eye(1, net.N, lk)[0] * h),
         1e-2 * ke[lk]),
        ("tank_h0", gh0[1, 0], lambda h: loss(h0=h0 + e_h * h), 1e-2 * h0[1, 0])]
    for name, ga, f, h in checks:
        fd = _fd(f, h, fw)
        rel = abs(ga - fd) / abs(fd)
        print("%s %s: adjoint %.10e fd %.10e rel %.1e" % (which, name, ga, fd, rel))
        assert rel < 1e-6, (name, ga, fd)
    if which == "geometric":
        # the controls switch pumps within the horizon and no tank clamps
        assert (np.diff(fw["S_in"].astype(int), axis=0) != 0).any()
    else:
        # the small tank of the 6x6 network empties and fills: clamped
        # steps pass no derivative through the tank volume
        assert fw["clamp"].any()


def test_tank_flow_couples_the_steps():
    # dLoss/dh_0 through later steps only: a loss on step 3's heads still
    # depends on the initial tank level, through three tank updates
    net = _small_geometric()
    T, B = 4, 1
    dem, rh, ke, r, h0 = _setup(net, B, 3)
    W = np.zeros((T, B, net.N))
    W[3, 0, :net.Nj] = 1.0
    fw = ea.forward(net, dem, rh, ke, r, h0, T, accuracy=1e-10)
    gh0 = ea.adjoint(net, fw, W)[4]
    assert np.abs(gh0).min() > 0.0
    # and a loss on step 0's heads sees h_0 only through step 0's solve
    W0 = np.zeros_like(W)
    W0[0, 0, :net.Nj] = 1.0
    gd0, _, _, _, gh00 = ea.adjoint(net, fw, W0)
    fw1 = ea.forward(net, dem, rh, ke, r, h0, 1, accuracy=1e-10)
    assert np.allclose(gh00, ea.adjoint(net, fw1, W0[:1])[4], rtol=1e-12, atol=0)


def test_forward_equals_event_driven_driver_on_whole_hour_steps():
    # without controls and with tanks far from their bounds, the
    # event-driven driver steps exactly one hour at a time: the two models
    # coincide up to the fixed-step model's polish of every step (the
    # driver's heads are the Newton iterate at relerr <= 1e-10)
    net = _small_geometric()
    net.ctl_link = net.ctl_link[:0]
    net.ctl_node = net.ctl_node[:0]
    net.ctl_type = net.ctl_type[:0]
    net.ctl_grade = net.ctl_grade[:0]
    net.ctl_time = net.ctl_time[:0]
    net.ctl_status = net.ctl_status[:0]
    T = 6
    net.meta["duration_sec"] = 3600 * (T - 1)
    ref = run_eps(net, accuracy=1e-10, extended=True)
    assert ref["t_sec"].tolist() == [3600 * t for t in range(T)]
    dem = net.demand_cfs[None, :]
    fw = ea.forward(net, dem, np.nan_to_num(net.reservoir_head_ft()), net.node_ke, net.r_hw,
                    net.tank_h0[None, :], T, accuracy=1e-10)
    assert not fw["clamp"].any()
    dh = np.abs(fw["H"][:, 0] - ref["head_ft"]).max()
    print("fixed-step model vs event-driven driver: max|dH| %.2e ft" % dh)
    assert dh < 1e-9
    assert np.array_equal(fw["S_in"][1:, 0], ref["status"][:-1])
