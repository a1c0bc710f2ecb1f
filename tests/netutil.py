# -*- coding: utf-8 -*-
"""Tiny hand-built networks for closed-form and brute-force checks."""
import numpy as np

from wdsgpu_amd.network import (Net, JUNCTION, RESERVOIR, PIPE, PRV, TCV, CVPIPE,
                                hw_resistance)


def make_net(nodes, links, accuracy=1e-3, ke=None):
    """nodes: list of (type, elev_ft, demand_cfs); links: list of dicts with
    n1, n2 and type PIPE (len_ft, diam_ft, C, [km]) / PRV (diam_ft,
    setting_ft, [status]) / TCV (diam_ft, setting) / CVPIPE."""
    N = len(nodes)
    nt = np.array([n[0] for n in nodes], dtype=np.int8)
    elev = np.array([n[1] for n in nodes], dtype=np.float64)
    dem = np.array([n[2] for n in nodes], dtype=np.float64)
    L = len(links)
    lt = np.array([lk.get("type", PIPE) for lk in links], dtype=np.int8)
    ln = np.array([lk.get("len_ft", 0.0) for lk in links])
    dm = np.array([lk.get("diam_ft", 1.0) for lk in links])
    C = np.array([lk.get("C", 0.0) for lk in links])
    km = np.array([lk.get("km", 0.0) for lk in links])
    r = np.zeros(L)
    pipe = lt <= PIPE
    r[pipe] = hw_resistance(C[pipe], ln[pipe], dm[pipe])
    st = np.array([lk.get("status", 2 if lk.get("type", PIPE) in (PRV, TCV) else 1)
                   for lk in links], dtype=np.int8)
    # PRV settings are in user pressure units; flow units CFS -> US units (psi):
    # setting_ft is converted with Ucf[PRESSURE] = 0.4333 psi/ft.
    setting = np.array([lk.get("setting_ft", 0.0) * 0.4333 if lk.get("type") == PRV
                        else lk.get("setting", 0.0) for lk in links])
    meta = dict(flow_units="CFS", headloss="H-W", accuracy=accuracy, trials=200)
    return Net(node_id=["N%d" % i for i in range(N)], node_type=nt, elev_ft=elev,
               node_ke=np.zeros(N) if ke is None else np.asarray(ke, dtype=np.float64),
               demand_cfs=dem, link_id=["L%d" % k for k in range(L)], link_type=lt,
               link_n1=np.array([lk["n1"] for lk in links], dtype=np.int32),
               link_n2=np.array([lk["n2"] for lk in links], dtype=np.int32),
               diam_ft=dm, len_ft=ln, roughness=C, km_int=km, r_hw=r,
               init_status=st, valve_setting_user=setting, meta=meta)


def res_heads(net):
    return np.nan_to_num(net.reservoir_head_ft())


def check_forward_against_oracle(net, m, out, d, rh, acc, tol_h_ft=1e-10 / 0.3048,
                                 tol_q=1e-9):
    """GPU forward output vs the oracle on the same inputs: identical Newton
    iteration counts and final statuses, heads within 1e-10 m of the
    oracle's refined iterate and flows within tol_q cfs.  Both paths evaluate
    the power laws with the same deterministic pow/log, so their iterates
    differ only by the roundings of the linear solves (the oracle refines in
    extended precision, the kernels in double-double).  The measured
    differences and, for scale, the oracle's own change under a one-ulp
    perturbation of its power laws are printed."""
    from oracle import gga
    ke = net.node_ke
    ext = gga.solve(m, d, rh, ke=ke, accuracy=acc, status_machine=True, extended=True)
    ref = gga.solve(m, d, rh, ke=ke, accuracy=acc, status_machine=True)
    assert out["iters"].cpu().numpy().tolist() == ref["iters"].tolist()
    assert out["iters"].cpu().numpy().tolist() == ext["iters"].tolist()
    assert np.array_equal(out["status"].cpu().numpy(), ref["status"])
    assert bool(out["converged"].all())
    m.pow_jitter = np.random.default_rng(123)
    try:
        jit = gga.solve(m, d, rh, ke=ke, accuracy=acc, status_machine=True, extended=True)
    finally:
        m.pow_jitter = None
    noise = np.abs(jit["head_ft"] - ext["head_ft"]).max()
    Hg = out["head_ft"].cpu().numpy()
    dh = np.abs(Hg - ext["head_ft"]).max()
    dq = np.abs(out["flow_cfs"].cpu().numpy() - ext["flow_cfs"]).max()
    print("max|dH| %.3e ft (1-ulp pow jitter of the oracle %.3e ft), max|dQ| %.3e cfs"
          % (dh, noise, dq), flush=True)
    assert dh <= tol_h_ft
    assert dq <= tol_q
    return ref


# feature_network cases (headloss, seed, tank level m): together they cover
# pumps OPEN and XHEAD, check valves OPEN and CLOSED, ACTIVE PRVs and tank
# links OPEN and TEMPCLOSED (checked on the CPU in test_oracle.py)
FEATURE_CASES = [("H-W", 0, 18.0), ("H-W", 2, 18.0), ("H-W", 1, 25.0),
                 ("D-W", 0, 18.0), ("D-W", 3, 18.0), ("D-W", 2, 25.0)]


def custom_pump_tree(points, demand_cfs, speed=1.0, H0=100.0, flow_units="CFS"):
    """reservoir (head H0 ft) -> CUSTOM pump with the given curve points
    (user units of flow_units) -> junction with a demand: the pump flow is
    the demand and the junction head is H0 plus the pump's lift."""
    from wdsgpu_amd.network import PUMP, PUMP_CUSTOM, SI_FLOW_UNITS, pump_custom_curve
    c = pump_custom_curve(points, flow_units=flow_units, si=flow_units in SI_FLOW_UNITS)
    net = make_net([(RESERVOIR, H0, 0.0), (JUNCTION, 0.0, demand_cfs)],
                   [dict(n1=0, n2=1, type=PUMP, diam_ft=0.0)])
    net.meta["flow_units"] = flow_units
    net.pump_link = np.array([0], dtype=np.int32)
    net.pump_ptype = np.array([PUMP_CUSTOM], dtype=np.int8)
    for k in ("h0", "r", "n", "q0", "qmax", "hmax"):
        setattr(net, "pump_" + k, np.array([c[k]]))
    net.pump_curve_q = c["x"]
    net.pump_curve_h = c["y"]
    net.pump_curve_ptr = np.array([0, len(c["x"])], dtype=np.int32)
    net.valve_setting_user[0] = speed
    return net


def curve_lift(points, q):
    """Head gain of a piecewise-linear pump curve at flow q: linear
    interpolation between the points, the end segments extended beyond."""
    x = np.array([p[0] for p in points])
    y = np.array([p[1] for p in points])
    if q < x[0]:
        return y[0] + (y[1] - y[0]) / (x[1] - x[0]) * (q - x[0])
    if q > x[-1]:
        return y[-1] + (y[-1] - y[-2]) / (x[-1] - x[-2]) * (q - x[-1])
    return float(np.interp(q, x, y))


def hub_network(n_ring=60, n_spokes=48, seed=0):
    """A reservoir, a ring of junctions with a few chords, and one hub
    junction piped to n_spokes ring junctions: the hub row of the factor
    holds about n_spokes updates, so the top of the elimination tree has one
    long dot product (a 64-lane group at one step per level)."""
    rng = np.random.default_rng(seed)
    nodes = [(RESERVOIR, 300.0, 0.0)]
    nodes += [(JUNCTION, 100.0 + 5.0 * rng.random(), 0.05 + 0.1 * rng.random())
              for _ in range(n_ring)]
    hub = len(nodes)
    nodes.append((JUNCTION, 102.0, 0.2))

    def pipe(a, b, d=0.8, L=800.0):
        return dict(n1=a, n2=b, len_ft=L * (0.5 + rng.random()), diam_ft=d, C=110.0)

    links = [pipe(0, 1, 1.5, 300.0), pipe(0, hub, 1.5, 300.0)]
    links += [pipe(i, i + 1) for i in range(1, n_ring)] + [pipe(n_ring, 1)]
    for _ in range(6):
        a, b = rng.choice(np.arange(1, n_ring + 1), 2, replace=False)
        links.append(pipe(int(a), int(b), 0.5))
    for j in rng.choice(np.arange(1, n_ring + 1), n_spokes, replace=False):
        links.append(pipe(hub, int(j), 0.6))
    return make_net(nodes, links, accuracy=1e-6)
