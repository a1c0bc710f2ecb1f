# -*- coding: utf-8 -*-
"""Host logic of the calibration consumers (wdsgpu_amd.calib,
.sensitivity, .optim2) that runs without a GPU: the optimiser, the
post-processing of sensitivity matrices and the topological masks."""
import numpy as np
import pytest
import torch

from wdsgpu_amd import calib, optim2, sensitivity
from wdsgpu_amd.network import JUNCTION, PIPE, RESERVOIR
from wdsgpu_amd.solver import GGASolver

from netutil import make_net


def _quadratic(h):
    def fun(x):
        return 0.5 * float(x @ (h * x)) - float(x.sum()), h * x - 1.0
    return fun


def test_lbfgs_reaches_the_minimum_of_a_quadratic():
    h = np.array([1.0, 10.0, 100.0, 1000.0])
    # the line search's worst case must fit the evaluation budget
    # (max_eval defaults to 5 max_iter)
    assert optim2.lbfgs_precond(_quadratic(h), np.zeros(4), max_iter=5)["stop"] == "max_eval"
    res = optim2.lbfgs_precond(_quadratic(h), np.zeros(4), max_iter=200, tol_grad=1e-12)
    assert res["stop"] in ("tol_grad", "tol_change")
    assert np.allclose(res["x"], 1.0 / h, rtol=1e-9)
    assert res["f_hist"][-1] <= res["f_hist"][0]


def test_exact_diagonal_preconditioner_solves_a_diagonal_quadratic_in_one_step():
    # with M = H (diagonal) the first direction is the Newton step and the
    # unit step satisfies the strong-Wolfe conditions
    h = np.array([1.0, 10.0, 100.0, 1000.0, 1e4])
    res = optim2.lbfgs_precond(_quadratic(h), np.zeros(5), minv=1.0 / h, max_iter=5,
                               max_eval=100, tol_grad=1e-12)
    assert res["n_iter"] == 1 and res["stop"] == "tol_grad"
    assert np.allclose(res["x"], 1.0 / h, rtol=1e-12)
    plain = optim2.lbfgs_precond(_quadratic(h), np.zeros(5), max_iter=5, max_eval=100,
                                 tol_grad=1e-12)
    assert plain["n_iter"] > 1


def test_lbfgs_counts_every_evaluation():
    calls = [0]

    def rosen(x):
        calls[0] += 1
        f = (1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2
        g = np.array([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2),
                      200 * (x[1] - x[0] ** 2)])
        return f, g

    res = optim2.lbfgs_precond(rosen, np.array([-1.2, 1.0]), max_iter=200, max_eval=1000)
    assert res["n_eval"] == calls[0]
    assert np.allclose(res["x"], [1.0, 1.0], atol=1e-6)


def test_fisher_information_and_spectrum_closed_forms():
    S = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    F = sensitivity.fisher_information(S, 0.5)
    assert np.allclose(F, np.diag([40.0, 16.0, 0.0]))
    Fv = sensitivity.fisher_information(S, np.array([1.0, 2.0, 1.0, 1.0]))
    assert np.allclose(np.diag(Fv), [10.0, 1.0, 0.0])
    sp = sensitivity.svd_spectrum(S, full=True)
    assert sp["rank"] == 2
    assert np.allclose(sp["sv"], [np.sqrt(10.0), 2.0, 0.0])
    assert np.isclose(sp["cond"], np.sqrt(10.0) / 2.0)
    assert np.isclose(sp["energy"][-1], 1.0)
    assert np.allclose(np.abs(sp["null_dirs"]), [[0.0, 0.0, 1.0]])
    assert sensitivity.identifiability_mask(S).tolist() == [True, True, False]
    with pytest.raises(ValueError):
        sensitivity.fisher_information(S, np.ones(3))
    with pytest.raises(ValueError):
        sensitivity.fisher_information(S, 0.0)


def test_gauss_newton_diagonal_and_floor():
    S = np.array([[1.0, 2.0, 0.0], [3.0, 0.0, 0.0]])
    M, info = optim2.gn_diag_from_S(S, 2, w_reg=0.0, floor_rel=0.01)
    assert np.allclose(M, [5.0, 2.0, 0.05])
    assert info["n_floored"] == 1 and info["max"] == 5.0


def test_latin_hypercube_has_one_point_per_stratum():
    u = optim2.lhs_starts(16, 0.0, 1.0, 3, seed=4)
    assert u.shape == (16, 3)
    for j in range(3):
        assert sorted(np.floor(u[:, j] * 16).astype(int).tolist()) == list(range(16))


def _branch_net():
    # reservoir 0 - junctions 1..3 in a loop, a dead chain 3-4-5 with no
    # demand, a closed pipe 1-6 to a demand node fed also by 2-6
    nodes = [(RESERVOIR, 100.0, 0.0), (JUNCTION, 10.0, 0.1), (JUNCTION, 10.0, 0.1),
             (JUNCTION, 10.0, 0.1), (JUNCTION, 10.0, 0.0), (JUNCTION, 10.0, 0.0),
             (JUNCTION, 10.0, 0.1)]
    pipe = lambda a, b, **kw: dict(n1=a, n2=b, len_ft=500.0, diam_ft=0.5, C=100.0, **kw)
    links = [pipe(0, 1), pipe(1, 2), pipe(2, 3), pipe(3, 1), pipe(3, 4), pipe(4, 5),
             pipe(1, 6, status=0), pipe(2, 6)]
    return make_net(nodes, links)


def test_dead_branches_are_pruned_leaf_chains():
    net = _branch_net()
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    dead = calib.dead_branch_mask(s, demand=net.demand_cfs)
    assert np.nonzero(dead)[0].tolist() == [4, 5]
    # a demand at the chain's end keeps it alive
    d = net.demand_cfs.copy()
    d[5] = 0.01
    assert not calib.dead_branch_mask(s, demand=d).any()
    # an emitter at its end too
    ke = np.zeros(net.N)
    ke[5] = 1.0
    assert not calib.dead_branch_mask(s, demand=net.demand_cfs, ke=ke).any()


def test_clamped_mask_follows_the_rqtol_branch():
    net = _branch_net()
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    r = np.asarray(net.r_hw, dtype=np.float64)
    q = np.full((2, net.L), 0.5)
    q[0, 4] = 0.0                     # clamped in frame 0 only
    q[:, 5] = 0.0                     # clamped in both frames
    cm = calib.clamped_mask(s, dict(q=q, r_hw=r))
    assert cm["per_frame"][0, 4] and not cm["per_frame"][1, 4]
    assert cm["mask"][5] and not cm["mask"][4]
    assert not cm["mask"][6]          # closed pipes are not counted as clamped
    assert cm["n_per_frame"].tolist() == [2, 1]


def test_roughness_parameter_checks_and_inverse():
    C = np.array([60.0, 100.0, 130.0, 150.0])
    th = calib.RoughnessParam.theta_of_C(C, 40.0, 160.0)
    back = 40.0 + 120.0 / (1.0 + np.exp(-th))
    assert np.allclose(back, C, rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        calib.RoughnessParam.theta_of_C(np.array([40.0]), 40.0, 160.0)
    net = _branch_net()
    with pytest.raises(ValueError):
        calib.RoughnessParam(net, pipe_idx=[0, 0], device="cpu")
    with pytest.raises(ValueError):
        calib.RoughnessParam(net, C_lo=100.0, C_hi=50.0, device="cpu")
    with pytest.raises(ValueError):
        calib.RoughnessParam(net, init=170.0, device="cpu")
    p = calib.RoughnessParam(net, pipe_idx=[1, 3], device="cpu")
    assert p.n_free == 2 and p.theta.shape == (2,)
    assert np.allclose(p.C().detach().numpy(), [100.0, 100.0])
    # the C -> r map runs on the HIP device only
    with pytest.raises(NotImplementedError):
        p.r_hw()


def test_chain_factor_is_minus_hexp_r_over_c():
    net = _branch_net()
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    r = np.asarray(net.r_hw, dtype=np.float64)
    g = calib.dr_dC(s, r)
    assert np.allclose(g, -1.852 * r / 100.0)


def test_sensitivity_matrix_refuses_bad_arguments():
    net = _branch_net()
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    with pytest.raises(ValueError):
        sensitivity.sensitivity_matrix(s, net.demand_cfs, np.zeros(net.N), None, [0])
    with pytest.raises(ValueError):
        sensitivity.sensitivity_matrix(s, net.demand_cfs, np.zeros(net.N), None, [1], wrt="x")
    with pytest.raises(ValueError):
        sensitivity.sensitivity_matrix(s, net.demand_cfs, np.zeros(net.N),
                                       torch.zeros(2, net.L), [1])
