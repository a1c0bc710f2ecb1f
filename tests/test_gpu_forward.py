# -*- coding: utf-8 -*-
"""GPU parity of the batched GGA forward solve against the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd.generators import grid_network, planar_network, scenario_batch
from wdsgpu_amd.network import JUNCTION, RESERVOIR
from wdsgpu_amd.solver import GGASolver

from netutil import make_net

# tolerance on heads: 1e-10 m
TOL_H_FT = 1e-10 / 0.3048

pytestmark = pytest.mark.gpu


def _case(name, B):
    if name == "grid":
        net = grid_network(0)
        d = np.broadcast_to(net.demand_cfs, (B, net.N)).copy()
        rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), (B, net.N)).copy()
        return net, d, rh, None
    net = planar_network(1)
    d, rh = scenario_batch(net, B, seed=5)
    return net, d, rh, 1e-6


@pytest.fixture(scope="module")
def solvers(cuda):
    out = {}
    for name in ("grid", "planar"):
        net = _case(name, 1)[0]
        out[name] = (net, gga.GGAModel(net), GGASolver(net, device=cuda,
                                                       dense_status_machine=True))
    return out


def test_factor_solve_matches_dense_cholesky(solvers):
    for name in ("grid", "planar"):
        net, m, s = solvers[name]
        B = 4
        _, d, rh, _ = _case(name, B)
        q = np.broadcast_to(m.init_flow(), (B, m.L)).copy()
        e = np.where(net.node_ke[m.junc_nodes] > 0, 1.0, 0.0)[None].repeat(B, 0)
        H = m.full_heads(np.zeros((B, m.Nj)), rh[:, m.fixed_nodes])
        S = np.broadcast_to(m.init_status_int, (B, m.L)).copy()
        P, Y = m.link_PY(q, S, m.r_hw[None])
        csr, F, *_ = m.assemble(P, Y, q, H, e, net.node_ke[m.junc_nodes][None].repeat(B, 0),
                                d[:, m.junc_nodes], S, csr=True)
        Hg_t, fac = s.factor_solve(torch.tensor(csr), torch.tensor(F), want_factor=True)
        Hg = Hg_t.cpu().numpy()
        Hx = gga.chol_solve_refined(m.csr_to_dense(csr), F, extended=True)
        assert np.abs(Hg - Hx).max() <= 1e-9 * np.abs(Hx).max() * 1e-3
        # the two phases on their own give the same bits
        fac2 = s.factor(torch.tensor(csr))
        assert torch.equal(fac, fac2)
        H2 = s.trisolve(fac2, torch.tensor(csr), torch.tensor(F))
        assert torch.equal(H2, Hg_t)
        # one factor serves several right-hand sides
        F3 = 0.5 * F + 0.1
        H3 = s.trisolve(fac2, torch.tensor(csr), torch.tensor(F3)).cpu().numpy()
        Hx3 = gga.chol_solve_refined(m.csr_to_dense(csr), F3, extended=True)
        assert np.abs(H3 - Hx3).max() <= 1e-9 * np.abs(Hx3).max() * 1e-3


@pytest.mark.parametrize("name,B", [("grid", 1), ("planar", 8)])
def test_forward_matches_oracle(solvers, name, B):
    net, m, s = solvers[name]
    _, d, rh, acc = _case(name, B)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=acc, status_machine=True)
    Hg = out["head_ft"].cpu().numpy()
    ext = gga.solve(m, d, rh, ke=net.node_ke, accuracy=acc, status_machine=True,
                    extended=True)
    ref = gga.solve(m, d, rh, ke=net.node_ke, accuracy=acc, status_machine=True)
    # identical Newton iteration counts and final statuses
    assert out["iters"].cpu().numpy().tolist() == ref["iters"].tolist()
    assert out["iters"].cpu().numpy().tolist() == ext["iters"].tolist()
    assert np.array_equal(out["status"].cpu().numpy(), ref["status"])
    assert bool(out["converged"].all())
    # Both paths evaluate the power laws with the same deterministic pow/log,
    # so the iterates differ only by the linear solves' roundings: heads
    # within 1e-10 m of the oracle's extended-precision-refined iterate
    dh = np.abs(Hg - ext["head_ft"]).max()
    dq = np.abs(out["flow_cfs"].cpu().numpy() - ext["flow_cfs"]).max()
    print("%s B=%d: max|dH| %.3e ft, max|dQ| %.3e cfs" % (name, B, dh, dq), flush=True)
    assert dh <= TOL_H_FT
    assert dq <= 1e-9
    if m.n_pcv:
        dr = out["diag_ratio"].cpu().numpy()
        assert np.allclose(dr, ref["diag_ratio"], rtol=1e-9)


def test_polished_forward_within_1e10_m(solvers):
    # the forward of the implicit solve (GGA + full-Newton polish) is the
    # state the adjoint differentiates; there both paths sit on the root of
    # the residual system to ~1e-13 and the 1e-10 m bar applies directly
    from oracle import adjoint as adj
    from wdsgpu_amd.autodiff import implicit_solve
    net, m, s = solvers["planar"]
    _, d, rh, _ = _case("planar", 4)
    ke = np.broadcast_to(net.node_ke, d.shape).copy()
    with torch.no_grad():
        H, Q, E = implicit_solve(s, torch.tensor(d), torch.tensor(rh), ke=torch.tensor(ke),
                                 accuracy=1e-8, polish_steps=3, adjoint="gpu",
                                 status_machine=True)
    sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=1e-8)
    assert (sol["resid_inf"] < 1e-10).all()
    assert np.abs(H.cpu().numpy() - sol["head"]).max() <= TOL_H_FT
    # flows through open valves (P = 1/CSMALL = 1e6) turn 1e-16 head
    # rounding into ~1e-10 cfs; every other link agrees to ~1e-15
    assert np.abs(Q.cpu().numpy() - sol["q"]).max() <= 1e-9
    assert float(s._last_resid.max()) < 1e-11


def test_polish_drives_the_residual_down(solvers):
    # Newton from the converged GGA iterate must not grow any residual row:
    # more polish steps leave the residual at rounding level, and the flows
    # of ACTIVE valves (closed by their downstream mass balance) agree with
    # the full-Jacobian CPU polish
    from oracle import adjoint as adj
    from wdsgpu_amd.autodiff import implicit_solve
    net, m, s = solvers["planar"]
    _, d, rh, _ = _case("planar", 4)
    ke = np.broadcast_to(net.node_ke, d.shape).copy()
    for steps in (3, 5):
        with torch.no_grad():
            H, Q, E = implicit_solve(s, torch.tensor(d), torch.tensor(rh),
                                     ke=torch.tensor(ke), accuracy=1e-8, polish_steps=steps,
                                     adjoint="gpu", status_machine=True)
        resid = s._last_resid.cpu().numpy()
        assert (resid < 1e-11).all(), (steps, resid)
        sol = adj.solve_polished(m, d, rh, ke, m.r_hw, accuracy=1e-8, polish_steps=steps)
        act = sol["status"][:, m.prv_k] == gga.ST_ACTIVE
        assert act.any()
        dq = np.abs(Q.cpu().numpy()[:, m.prv_k] - sol["q"][:, m.prv_k])
        assert dq[act].max() <= 1e-12 * np.abs(sol["q"][:, m.prv_k]).max()


def test_unfused_equals_fused_bitwise(solvers):
    net, m, s = solvers["planar"]
    _, d, rh, acc = _case("planar", 4)
    args = (torch.tensor(d), torch.tensor(rh))
    a = s.solve(*args, ke_int=torch.tensor(net.node_ke), accuracy=acc, status_machine=True)
    b = s.solve_unfused(*args, ke_int=torch.tensor(net.node_ke), accuracy=acc,
                        status_machine=True)
    assert torch.equal(a["head_ft"], b["head_ft"])
    assert torch.equal(a["flow_cfs"], b["flow_cfs"])
    assert torch.equal(a["iters"], b["iters"])


def test_batch_equals_single_bitwise(solvers):
    net, m, s = solvers["planar"]
    _, d, rh, acc = _case("planar", 5)
    ke = torch.tensor(net.node_ke)
    a = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=ke, accuracy=acc, status_machine=True)
    for b in (0, 3):
        o = s.solve(torch.tensor(d[b]), torch.tensor(rh[b]), ke_int=ke, accuracy=acc,
                    status_machine=True)
        assert torch.equal(o["head_ft"], a["head_ft"][b])
        assert int(o["iters"]) == int(a["iters"][b])


def test_batch_beyond_one_wavefront_per_simd_is_bitwise(solvers):
    # 1100 scenarios: more workgroups than SIMDs, so some SIMDs run two
    # scenarios; every scenario's result is still its own solve's bits
    net, m, s = solvers["planar"]
    _, d, rh, acc = _case("planar", 1100)
    ke = torch.tensor(net.node_ke)
    a = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=ke, accuracy=acc, status_machine=True)
    pick = [0, 549, 1023, 1024, 1099]
    o = s.solve(torch.tensor(d[pick]), torch.tensor(rh[pick]), ke_int=ke, accuracy=acc,
                status_machine=True)
    assert torch.equal(o["head_ft"], a["head_ft"][pick])
    assert torch.equal(o["flow_cfs"], a["flow_cfs"][pick])
    assert o["iters"].tolist() == a["iters"][pick].tolist()
    assert bool(a["converged"].all())


def test_unconverged_reports_maxtrials_plus_one(solvers):
    net, m, s = solvers["grid"]
    _, d, rh, _ = _case("grid", 2)
    out = s.solve(torch.tensor(d), torch.tensor(rh), max_iter=1, status_machine=True)
    assert out["iters"].tolist() == [2, 2]
    assert not bool(out["converged"].any())
    ref = gga.solve(m, d, rh, max_iter=1, status_machine=True)
    assert ref["iters"].tolist() == [2, 2]


def test_singular_matrix_raises_naming_samples(cuda):
    # an isolated junction with no emitter has a zero diagonal (dgga/solver.py:342-349)
    nodes = [(RESERVOIR, 100.0, 0.0), (JUNCTION, 0.0, 0.1), (JUNCTION, 0.0, 0.0)]
    links = [dict(n1=0, n2=1, len_ft=100.0, diam_ft=0.5, C=100.0)]
    net = make_net(nodes, links)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d = np.broadcast_to(net.demand_cfs, (2, 3)).copy()
    rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), (2, 3)).copy()
    with pytest.raises(RuntimeError, match=r"samples \[0, 1\]"):
        s.solve(torch.tensor(d), torch.tensor(rh), status_machine=True)


def test_cache_api_of_the_reference(cuda):
    # cudss_plan / cudss_cache_info / cudss_cache_max / cudss_free /
    # cudss_session / cudss_counters (dgga/solver.py:2514-2603)
    net = grid_network(0)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    rh = np.nan_to_num(net.reservoir_head_ft())
    assert s.cudss_plan(3) >= 0.0
    info = s.cudss_cache_info()
    assert [e[0] for e in info] == [3] and info[0][5] > 0.0
    a = s.solve(torch.tensor(np.tile(net.demand_cfs, (3, 1))), torch.tensor(np.tile(rh, (3, 1))),
                status_machine=True)
    assert len(s.cudss_cache_info()) == 1
    s.cudss_cache_max = 1
    s.solve(torch.tensor(np.tile(net.demand_cfs, (2, 1))), torch.tensor(np.tile(rh, (2, 1))),
            status_machine=True)
    assert [e[0] for e in s.cudss_cache_info()] == [2]
    with s.cudss_session():
        b = s.solve(torch.tensor(np.tile(net.demand_cfs, (3, 1))),
                    torch.tensor(np.tile(rh, (3, 1))), status_machine=True)
    assert s.cudss_cache_info() == []
    assert torch.equal(a["head_ft"], b["head_ft"])
    assert s.cudss_counters() == s.counters()
    s.cudss_cache_max = 0
    with pytest.raises(ValueError):
        s.cudss_plan(4)


def test_hub_network_matches_oracle(cuda):
    # one junction piped to 48 others: the top level of the forward program
    # is a single dot product spread over a 64-lane group (cross-row
    # butterfly stages), the factor's top levels likewise
    from netutil import check_forward_against_oracle, hub_network, res_heads
    net = hub_network(seed=0)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    B = 4
    rng = np.random.default_rng(7)
    d = net.demand_cfs[None] * (1.0 + 0.2 * rng.standard_normal((B, net.N)))
    rh = np.broadcast_to(res_heads(net), (B, net.N)).copy()
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-6, status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, 1e-6)


@pytest.mark.parametrize("seed", [2, 3, 7])
def test_other_planar_topologies_match_oracle(cuda, seed):
    # other seeded planar networks: other elimination trees (23 to 39 levels)
    # and so other mixed-group schedules of the factor and solve programs
    from netutil import check_forward_against_oracle
    net = planar_network(seed)
    m = gga.GGAModel(net)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 3, seed=11)
    out = s.solve(torch.tensor(d), torch.tensor(rh), ke_int=torch.tensor(net.node_ke),
                  accuracy=1e-6, status_machine=True)
    check_forward_against_oracle(net, m, out, d, rh, 1e-6)


def test_mismatched_input_shapes_raise_before_any_launch(solvers):
    # every per-scenario array is checked against B, N and L on the host
    net, m, s = solvers["planar"]
    _, d, rh, acc = _case("planar", 8)
    dt, rht = torch.tensor(d), torch.tensor(rh)
    bad = [dict(r_hw=torch.tensor(np.tile(m.r_hw, (2, 1)))),       # [2, L] for B = 8
           dict(r_hw=torch.tensor(m.r_hw[:-1])),                    # short r
           dict(ke_int=torch.tensor(net.node_ke[:-1])),             # short ke
           dict(ke_int=torch.tensor(np.tile(net.node_ke, (3, 1)))), # [3, N] for B = 8
           dict(q0=torch.tensor(m.init_flow()[:-2])),               # short warm start
           dict(e0=torch.tensor(np.zeros((2, net.N))))]
    for kw in bad:
        with pytest.raises(ValueError):
            s.solve(dt, rht, accuracy=acc, status_machine=True, **kw)
    with pytest.raises(ValueError):
        s.solve(dt[:, :-1], rht, accuracy=acc, status_machine=True)
    # a per-scenario r with 1-D demand fixes B from r in the implicit solve
    from wdsgpu_amd.autodiff import implicit_solve
    r2 = torch.tensor(np.tile(m.r_hw, (2, 1)))
    with torch.no_grad():
        H, Q, E = implicit_solve(s, torch.tensor(d[0]), torch.tensor(rh[0]), r_hw=r2,
                                 accuracy=1e-8, adjoint="gpu", status_machine=True)
    assert tuple(H.shape) == (2, net.N)
