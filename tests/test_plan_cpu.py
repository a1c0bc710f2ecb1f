# -*- coding: utf-8 -*-
"""CPU checks of libwds.so: exported symbols, host symbolic analysis, and the
level-scheduled programs executed by a numpy emulation of the wavefront
(same instruction semantics as the HIP executor) against dense solves."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from oracle import gga
from wdsgpu_amd import _lib
from wdsgpu_amd.generators import grid_network, planar_network, scenario_batch
from wdsgpu_amd.network import CVPIPE, GPV, JUNCTION, PIPE, RESERVOIR
from wdsgpu_amd.solver import GGASolver

from netutil import hub_network, make_net

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIRST, LAST, DIAG, NOP, LEVEL_END = 1 << 30, 1 << 29, 1 << 28, 1 << 27, 1 << 26
WIDTH_SHIFT, ZERO_INIT = 23, 1 << 22
TMASK = (1 << 22) - 1


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "wds.h")).read()
    names = set(re.findall(r"\b(wds_[a-z_]+)\s*\(", header))
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names <= set(_lib.SIGNATURES)


def run_program(s, p, V, X, mode):
    """Emulate the wavefront executors on the exported (index-form) programs:
    per-lane accumulators, stores on LAST, and at a level end a sum over each
    lane group (log2 width per lane in bits 23-25, aligned groups) that the
    group's first lane stores into the target of its slice (the slice started
    from the target's value on that lane, from zero on the others).
    mode 0: LDL^T factor (w_ij -= w_ik w_jk invD_k, reciprocal pivots);
    mode 1: unit-factor dot solve acc = X[t] - sum V[a] X[b]."""
    ins = s.plan_export(8 + 3 * p).reshape(-1, 64, 4)
    bad = False
    acc = np.zeros(64)
    level_ends, grouped_levels = 0, 0
    fin = [None] * 64
    for step in ins:
        for l in range(64):
            a, b, c, w = (int(v) for v in step[l])
            t = w & TMASK
            if w & FIRST:
                acc[l] = 0.0 if w & ZERO_INIT else (V[t] if mode == 0 else X[t])
                fin[l] = (t, bool(w & DIAG))
            if w & NOP:
                continue
            if mode == 0:
                acc[l] -= V[a] * (V[b] * V[c])
                if w & LAST:
                    if w & DIAG:
                        bad |= not acc[l] > 0
                        V[t] = 1.0 / acc[l]
                    else:
                        V[t] = acc[l]
            else:
                acc[l] -= V[a] * X[b]
                if w & LAST:
                    X[t] = acc[l]
        ends = (step[:, 3] & LEVEL_END) != 0
        assert ends.all() or not ends.any()   # uniform across lanes
        if ends[0]:
            level_ends += 1
            lw = (step[:, 3].astype(np.int64) >> WIDTH_SHIFT) & 7
            # lane 0 is in the level's widest group (the executor reads the
            # number of butterfly stages from it)
            assert lw[0] == lw.max()
            if lw.max():
                grouped_levels += 1
            for g in range(64):
                W = 1 << int(lw[g])
                if lw[g] == 0 or g % W:
                    continue
                assert (lw[g:g + W] == lw[g]).all()   # aligned, one width per group
                t, diag = fin[g]
                red = acc[g:g + W].sum()
                if mode == 0:
                    bad |= diag and not red > 0
                    V[t] = 1.0 / red if diag else red
                else:
                    X[t] = red
    return bad, level_ends, grouped_levels


@pytest.mark.parametrize("name", ["grid", "planar", "hub"])
def test_programs_factor_and_solve_like_dense(name):
    net = {"grid": lambda: grid_network(0), "planar": lambda: planar_network(1),
           "hub": lambda: hub_network(seed=0)}[name]()
    m = gga.GGAModel(net)
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    st = s.stats
    # symbolic sanity: permutation, etree parent above child, pattern of A in L
    perm = s.plan_export(0)
    assert sorted(perm.tolist()) == list(range(m.Nj))
    par = s.plan_export(3)
    assert all(p == -1 or p > j for j, p in enumerate(par))
    d, rh = scenario_batch(net, 1, seed=5)
    q = m.init_flow()[None]
    e = np.where(net.node_ke[m.junc_nodes] > 0, 1.0, 0.0)[None]
    H = m.full_heads(np.zeros((1, m.Nj)), rh[:, m.fixed_nodes])
    S = m.init_status_int[None]
    P, Y = m.link_PY(q, S, m.r_hw[None])
    csr, F, *_ = m.assemble(P, Y, q, H, e, net.node_ke[m.junc_nodes][None], d[:, m.junc_nodes],
                            S, csr=True)
    assert st["nnz_csr"] == csr.shape[1]
    c2l = s.plan_export(7)
    slot = np.argsort(perm)          # original junction row -> elimination slot
    V = np.zeros(st["nnzL"] + 1)
    V[c2l[c2l >= 0]] = csr[0][c2l >= 0]
    bad, lev, wide_f = run_program(s, 0, V, None, 0)
    assert not bad and lev == st["nlevels"]
    # unit factor: l_ij = w_ij / D_j below the diagonal
    cp, ri = s.plan_export(1), s.plan_export(2)
    for j in range(m.Nj):
        V[cp[j] + 1:cp[j + 1]] *= V[cp[j]]
    X = np.zeros(m.Nj)
    X[slot] = F[0]
    _, _, wide_s = run_program(s, 1, V, X, 1)
    X *= V[cp[:-1]]                  # z = D^-1 y
    run_program(s, 2, V, X, 1)
    if name == "hub":
        # the hub row's dot product spans a whole 64-lane group in one step
        ins = s.plan_export(8 + 3 * 1).reshape(-1, 64, 4)
        ends = ins[(ins[:, 0, 3] & LEVEL_END) != 0]
        assert ((ends[:, 0, 3] >> WIDTH_SHIFT) & 7).max() == 6
    if name == "planar":
        # the top of the elimination tree has few long dot products: those
        # levels must be split across lane groups
        assert wide_f > 0 and wide_s > 0
        # and the mixed group schedule keeps the programs short
        assert st["factor_steps"] < 80 and st["fwd_steps"] < 60 and st["bwd_steps"] < 60
    A = m.csr_to_dense(csr)[0]
    xr = np.linalg.solve(A, F[0])
    assert np.abs(X[slot] - xr).max() <= 1e-10 * np.abs(xr).max()


ROW_END = 1 << 30


@pytest.mark.parametrize("name", ["grid", "planar", "tiny"])
def test_lane_layout_multiplies_like_dense(name):
    # the lane layout of the symmetric rows (rs_share / rs_ell) that the
    # batched residual and matrix-vector product walk: lane l's rows l,
    # l + 64, ... packed in order, every row's entries in column order with
    # the last one flagged, pads (0, -1) only after a lane's range; the two
    # halves of a pair of wavefronts split every lane at a row boundary; and
    # A x equal to the dense product
    from netutil import make_net
    if name == "tiny":
        nodes = [(RESERVOIR, 100.0, 0.0)] + [(JUNCTION, 0.0, 0.1)] * 5
        links = [dict(n1=i, n2=i + 1, len_ft=500.0, diam_ft=0.5, C=100.0) for i in range(5)]
        net = make_net(nodes, links)
    else:
        net = grid_network(0) if name == "grid" else planar_network(1)
    m = gga.GGAModel(net)
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    st = s.stats
    d, rh = scenario_batch(net, 1, seed=5)
    q = m.init_flow()[None]
    e = np.where(net.node_ke[m.junc_nodes] > 0, 1.0, 0.0)[None]
    H = m.full_heads(np.zeros((1, m.Nj)), rh[:, m.fixed_nodes])
    S = m.init_status_int[None]
    P, Y = m.link_PY(q, S, m.r_hw[None])
    csr, F, *_ = m.assemble(P, Y, q, H, e, net.node_ke[m.junc_nodes][None], d[:, m.junc_nodes],
                            S, csr=True)
    c2l = s.plan_export(7)
    AL = np.zeros(st["nnzL"] + 1)
    AL[c2l[c2l >= 0]] = csr[0][c2l >= 0]
    share = s.plan_export(17).reshape(3, 64, 4)
    ell = s.plan_export(18).reshape(-1, 2)
    assert ell.shape[0] % 64 == 0
    T = ell.shape[0] // 64
    slot = np.argsort(s.plan_export(0))
    x = np.random.default_rng(0).normal(size=m.Nj)
    xs = np.zeros(m.Nj)
    xs[slot] = x
    ys = np.zeros(m.Nj)
    seen = np.zeros(m.Nj, dtype=int)
    rowlen = np.zeros(m.Nj, dtype=int)
    for lane in range(64):
        t0, t1, j0, rows = share[0, lane]
        assert t0 == 0 and j0 == lane and t1 <= T
        assert rows == (len(range(lane, m.Nj, 64)))
        j, cols = j0, []
        for t in range(t1):
            pos, cw = ell[t * 64 + lane]
            assert cw >= 0
            col = cw & (ROW_END - 1)
            cols.append(col)
            ys[j] += AL[pos] * xs[col]
            if cw & ROW_END:
                assert cols == sorted(cols) and j in cols
                seen[j] += 1
                rowlen[j] = len(cols)
                j, cols = j + 64, []
        assert not cols and j == lane + 64 * rows
        for t in range(t1, T):
            assert tuple(ell[t * 64 + lane]) == (0, -1)
        # the halves: [0, tk) with the first k rows, [tk, t1) with the rest
        a, b = share[1, lane], share[2, lane]
        assert a[0] == 0 and a[1] == b[0] and b[1] == t1
        assert a[2] == lane and b[2] == lane + 64 * a[3] and a[3] + b[3] == rows
    assert (seen == 1).all()
    A = m.csr_to_dense(csr)[0]
    assert np.allclose(ys[slot], A @ x, rtol=1e-13, atol=1e-12 * np.abs(A).max())
    # no lane holds more slots than 64-row blocks as wide as their longest
    # row would give it
    blocks = [rowlen[b:b + 64].max() for b in range(0, m.Nj, 64)]
    assert T <= sum(blocks)


def test_minimum_degree_beats_natural_ordering():
    net = planar_network(1)
    a = GGASolver(net, device="cpu", dense_status_machine=True, ordering="mmd").stats
    b = GGASolver(net, device="cpu", dense_status_machine=True, ordering="natural").stats
    assert a["nnzL"] < b["nnzL"] and a["factor_updates"] < b["factor_updates"]


def test_plan_errors_map_to_reference_exceptions():
    nodes = [(RESERVOIR, 100.0, 0.0), (JUNCTION, 0.0, 0.1)]
    with pytest.raises(ValueError, match="self-loop"):
        GGASolver(make_net(nodes, [dict(n1=1, n2=1, len_ft=10.0, diam_ft=0.5, C=100.0)]),
                  device="cpu")
    with pytest.raises(NotImplementedError):
        GGASolver(make_net(nodes, [dict(n1=0, n2=1, type=GPV, diam_ft=0.5)]), device="cpu",
                  dense_status_machine=True)
    with pytest.raises(NotImplementedError):
        # CV pipes need the status machine (dgga/solver.py:2779-2785)
        GGASolver(make_net(nodes, [dict(n1=0, n2=1, type=CVPIPE, len_ft=10.0, diam_ft=0.5,
                                        C=100.0)]), device="cpu")


def test_custom_pump_curves_reach_the_plan():
    from netutil import custom_pump_tree
    net = custom_pump_tree([(2.0, 60.0), (6.0, 50.0), (10.0, 32.0)], 4.0)
    GGASolver(net, device="cpu", dense_status_machine=True)
    # the batched segment search needs strictly increasing flows
    # (dgga/solver.py:989-996)
    net.pump_curve_q = np.array([2.0, 2.0, 10.0])
    with pytest.raises(NotImplementedError, match="strictly increasing"):
        GGASolver(net, device="cpu", dense_status_machine=True)


def test_solver_refuses_cpu_execution():
    net = grid_network(0)
    s = GGASolver(net, device="cpu", dense_status_machine=True)
    with pytest.raises(NotImplementedError, match="HIP device"):
        s.solve(net.demand_cfs, np.nan_to_num(net.reservoir_head_ft()), status_machine=True)


def test_parameter_maps_refuse_cpu_tensors():
    from wdsgpu_amd.autodiff import emitter_ke, hw_resistance
    one = torch.ones(3, dtype=torch.float64)
    with pytest.raises(NotImplementedError, match="HIP device"):
        hw_resistance(one, one, one)
    with pytest.raises(NotImplementedError, match="HIP device"):
        emitter_ke(one)


_PSWITCH_INP = """[JUNCTIONS]
 J1 10 5
 J2 8 3
[RESERVOIRS]
 R1 60
[TANKS]
 T1 20 5 2 10 30 0
[PIPES]
 P1 R1 J1 1000 12 100 0 Open
 P2 J1 J2 1000 10 100 0 Open
 P3 J2 T1 1000 10 100 0 Open
[CONTROLS]
%s
[OPTIONS]
 Units GPM
 Headloss H-W
[END]
"""


@pytest.mark.parametrize("control,refused", [
    ("LINK P2 CLOSED IF NODE J1 BELOW 20", True),
    ("LINK P3 OPEN IF NODE J2 ABOVE 10", True),
    ("LINK P3 CLOSED IF NODE T1 ABOVE 12", False),
    ("LINK P2 CLOSED AT TIME 2", False),
    ("", False),
])
def test_junction_head_controls_are_refused(tmp_path, control, refused):
    # pswitch guard of the dense path (dgga/solver.py:617-644): a simple
    # control watching a junction's head would change a link after
    # convergence, which the batched status machine does not do
    from wdsgpu_amd.network import read_inp
    p = tmp_path / "ctl.inp"
    p.write_text(_PSWITCH_INP % control)
    net = read_inp(p)
    if refused:
        with pytest.raises(NotImplementedError, match="pswitch"):
            GGASolver(net, device="cpu", dense_status_machine=True)
        with pytest.raises(NotImplementedError, match="pswitch"):
            GGASolver(net, device="cpu")
    else:
        GGASolver(net, device="cpu", dense_status_machine=True)
