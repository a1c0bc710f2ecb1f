# -*- coding: utf-8 -*-
"""Seeded synthetic networks and scenario batches for the GGA workloads.

Every generator takes an integer seed and is deterministic for it (numpy
``default_rng``).  Quantities are drawn in SI user units (m, mm, L/s) and
converted to EPANET internal units (ft, cfs) exactly as the reference's INP
reader does (dgga/parse.py:557-611): lengths / 0.3048, diameters / 304.8,
flows / 28.317, emitter C -> Ke via (28.317^Qexp / 0.3048) / C^Qexp.

Workloads:
  * ``grid_network``    -- 10x10 grid, 180 grid + 20 chord pipes, one
    reservoir at 100 m joined to corner (0,0).
  * ``planar_network``  -- planar random network: Delaunay graph of random
    points, a random breadth-first spanning tree grown from the reservoir
    mains, chords to loop density ~0.15, 3 PRVs on trunk mains, 10 leaks.
  * ``feature_network`` -- grid network with every link and node kind of the
    path: a pumped supply (3-point head curve, shutoff U[40,80] m, design
    flow U[20,200] L/s), a second reservoir, a tank with level bounds, TCVs,
    a check-valve pipe, PRVs and leak emitters; Hazen-Williams or
    Darcy-Weisbach friction (roughness U[0.01,0.5] mm).
  * ``scenario_batch``  -- per-scenario i.i.d. lognormal demand multipliers.
"""
from __future__ import annotations

import numpy as np

from .network import (Net, JUNCTION, RESERVOIR, TANK, PIPE, CVPIPE, PUMP, PRV, PSV, FCV, TCV,
                      M_PER_FT,
                      LPS_PER_CFS, PUMP_POWER_FUNC, PUMP_CUSTOM, hw_resistance, dw_resistance,
                      emitter_ke, pump_power_curve, pump_custom_curve)

DIAMS_MM = np.array([100.0, 150.0, 200.0, 250.0, 300.0, 400.0])


def _mk_net(node_type, elev_m, demand_lps, links, meta, ke=None):
    """links: list of dicts with n1, n2, type, len_m, diam_mm, C, setting_m."""
    N = len(node_type)
    L = len(links)
    lt = np.array([lk["type"] for lk in links], dtype=np.int8)
    ln_ft = np.array([lk.get("len_m", 0.0) for lk in links]) / M_PER_FT
    d_ft = np.array([lk["diam_mm"] for lk in links]) / (1000.0 * M_PER_FT)
    C = np.array([lk.get("C", 0.0) for lk in links])
    r = np.zeros(L)
    pipe = lt == PIPE
    r[pipe] = hw_resistance(C[pipe], ln_ft[pipe], d_ft[pipe])
    st = np.array([2 if lk["type"] == PRV else 1 for lk in links], dtype=np.int8)
    setting = np.array([lk.get("setting_m", 0.0) for lk in links])
    nid = [("R%d" % i if t != JUNCTION else "J%d" % i) for i, t in enumerate(node_type)]
    return Net(node_id=nid, node_type=np.asarray(node_type, dtype=np.int8),
               elev_ft=np.asarray(elev_m, dtype=np.float64) / M_PER_FT,
               node_ke=np.zeros(N) if ke is None else ke,
               demand_cfs=np.asarray(demand_lps, dtype=np.float64) / LPS_PER_CFS,
               link_id=["P%d" % k for k in range(L)], link_type=lt,
               link_n1=np.array([lk["n1"] for lk in links], dtype=np.int32),
               link_n2=np.array([lk["n2"] for lk in links], dtype=np.int32),
               diam_ft=d_ft, len_ft=ln_ft, roughness=C, km_int=np.zeros(L),
               r_hw=r, init_status=st, valve_setting_user=setting, meta=meta)


def grid_network(seed=0, nx=10, ny=10, n_chords=20):
    """10x10 grid network (100 junctions, elevation U[0,50] m, demand
    U[0.1,2.0] L/s), one reservoir at 100 m joined to corner junction (0,0),
    180 grid pipes plus ``n_chords`` cell-diagonal chords; pipe length
    U[100,1000] m, diameter from {100..400} mm, C U[90,140].  The reservoir
    main is a 100 m, 400 mm, C=130 pipe."""
    g = np.random.default_rng(seed)
    nj = nx * ny
    elev = np.concatenate([g.uniform(0.0, 50.0, nj), [100.0]])
    dem = np.concatenate([g.uniform(0.1, 2.0, nj), [0.0]])
    ntype = [JUNCTION] * nj + [RESERVOIR]
    res = nj
    jid = lambda i, j: i * ny + j
    pairs = []
    for i in range(nx):
        for j in range(ny):
            if i + 1 < nx:
                pairs.append((jid(i, j), jid(i + 1, j)))
            if j + 1 < ny:
                pairs.append((jid(i, j), jid(i, j + 1)))
    diag = [(jid(i, j), jid(i + 1, j + 1)) for i in range(nx - 1) for j in range(ny - 1)]
    diag += [(jid(i, j + 1), jid(i + 1, j)) for i in range(nx - 1) for j in range(ny - 1)]
    pick = g.choice(len(diag), size=n_chords, replace=False)
    pairs += [diag[int(p)] for p in pick]
    links = [dict(n1=res, n2=jid(0, 0), type=PIPE, len_m=100.0, diam_mm=400.0, C=130.0)]
    for a, b in pairs:
        links.append(dict(n1=a, n2=b, type=PIPE, len_m=g.uniform(100.0, 1000.0),
                          diam_mm=float(g.choice(DIAMS_MM)), C=g.uniform(90.0, 140.0)))
    meta = dict(flow_units="LPS", headloss="H-W", accuracy=1e-3, trials=200,
                seed=int(seed), workload="grid10x10")
    return _mk_net(ntype, elev, dem, links, meta)


def feature_network(seed=0, nx=6, ny=6, headloss="H-W", n_pumps=1, n_prv=1, n_tcv=1,
                    n_cv=1, tank=True, n_leaks=2, tank_level_m=18.0, n_custom=0, n_psv=0,
                    n_fcv=0, psv_m=(20.0, 40.0), fcv_lps=(1.0, 6.0)):
    """Grid of nx*ny junctions (elevation U[0,30] m, demand U[0.2,1.5] L/s,
    pipes 100-600 m) with: ``n_pumps`` pumps lifting from low reservoirs
    (head 5 m) into corner junctions, each with a 3-point curve (0, h0),
    (qd, 0.75 h0), (2 qd, 0) -- shutoff h0 U[40,80] m, design flow qd
    U[20,200] L/s; a reservoir at 70 m on the opposite corner; a tank
    (bottom 40 m, level ``tank_level_m`` in [2, 25] m) on an edge junction; ``n_tcv``
    grid pipes turned into TCVs (loss coefficient U[2,10]), ``n_cv`` into
    check-valve pipes; ``n_prv`` PRVs (setting U[25,35] m) each feeding a
    two-junction branch of its own; ``n_leaks`` leak emitters.  The first
    ``n_custom`` pumps get a 4-point CUSTOM curve instead, (0.25 qd, 0.97 h0),
    (qd, 0.75 h0), (1.6 qd, 0.42 h0), (2.2 qd, 0.08 h0).  ``n_fcv`` grid
    pipes become FCVs (flow setting U[fcv_lps] L/s); ``n_psv`` PSVs
    (upstream pressure setting U[psv_m] m) each let a grid junction drain
    through a junction 5 m lower to a reservoir 10 m below it.  Control
    valves start ACTIVE.
    headloss "H-W" (C U[90,140]) or "D-W" (roughness U[0.01,0.5] mm)."""
    g = np.random.default_rng(seed)
    nj = nx * ny
    jid = lambda i, j: i * ny + j
    elev = list(g.uniform(0.0, 30.0, nj))
    dem = list(g.uniform(0.2, 1.5, nj))
    ntype = [JUNCTION] * nj
    links = []
    pairs = []
    for i in range(nx):
        for j in range(ny):
            if i + 1 < nx:
                pairs.append((jid(i, j), jid(i + 1, j)))
            if j + 1 < ny:
                pairs.append((jid(i, j), jid(i, j + 1)))
    for a, b in pairs:
        links.append(dict(n1=a, n2=b, type=PIPE, len_m=g.uniform(100.0, 600.0),
                          diam_mm=float(g.choice(DIAMS_MM[1:])), C=g.uniform(90.0, 140.0),
                          rough_mm=g.uniform(0.01, 0.5)))

    def add_node(t, el, d=0.0):
        ntype.append(t)
        elev.append(el)
        dem.append(d)
        return len(ntype) - 1

    def main(n1, n2, diam=300.0, length=50.0):
        links.append(dict(n1=n1, n2=n2, type=PIPE, len_m=length, diam_mm=diam, C=130.0,
                          rough_mm=0.05))

    pumps = []
    corners = [jid(0, 0), jid(nx - 1, 0), jid(0, ny - 1)]
    for k in range(n_pumps):
        r = add_node(RESERVOIR, 5.0)
        h0 = g.uniform(40.0, 80.0)
        qd = g.uniform(20.0, 200.0)
        if k < n_custom:
            cu = pump_custom_curve([(0.25 * qd, 0.97 * h0), (qd, 0.75 * h0),
                                    (1.6 * qd, 0.42 * h0), (2.2 * qd, 0.08 * h0)])
            pumps.append((len(links), cu))
        else:
            pumps.append((len(links), pump_power_curve(h0, 0.75 * h0, qd, 0.0, 2.0 * qd)))
        links.append(dict(n1=r, n2=corners[k % 3], type=PUMP, diam_mm=0.0))
    r2 = add_node(RESERVOIR, 70.0)
    main(r2, jid(nx - 1, ny - 1))
    tank_node = None
    if tank:
        tank_node = add_node(TANK, 40.0)
        main(tank_node, jid(nx // 2, ny - 1), diam=200.0, length=150.0)
    # PRVs feed pressure zones of their own: grid junction -> PRV -> branch
    # junction (5 m lower) -> pipe -> end junction (8 m lower)
    taps = g.choice(nj, size=n_prv, replace=False) if n_prv else []
    for a in taps:
        a = int(a)
        b = add_node(JUNCTION, elev[a] - 5.0, 2.0)
        c = add_node(JUNCTION, elev[a] - 8.0, 1.0)
        links.append(dict(n1=a, n2=b, type=PRV, diam_mm=150.0, setting_m=g.uniform(25.0, 35.0)))
        links.append(dict(n1=b, n2=c, type=PIPE, len_m=g.uniform(100.0, 300.0), diam_mm=150.0,
                          C=g.uniform(90.0, 140.0), rough_mm=g.uniform(0.01, 0.5)))
    # TCVs and check-valve pipes replace grid pipes (keeping the diameter)
    grid_links = list(range(len(pairs)))
    g.shuffle(grid_links)
    for k in grid_links[:n_tcv]:
        links[k].update(type=TCV, setting=g.uniform(2.0, 10.0))
    for k in grid_links[n_tcv:n_tcv + n_cv]:
        links[k].update(type=CVPIPE)
    # control valves share no node (EPANET leaves every controlled valve out
    # of the nodal inflow an ACTIVE valve's flow is set from, so valves on
    # one node would not balance it): FCVs on grid pipes away from the PRV
    # taps, PSVs on junctions that touch no other control valve
    busy = {int(a) for a in taps}
    fcv = []
    for k in grid_links[n_tcv + n_cv:]:
        if len(fcv) == n_fcv:
            break
        a, b = pairs[k]
        if a in busy or b in busy:
            continue
        fcv.append(k)
        busy |= {a, b}
    for k in fcv:
        links[k].update(type=FCV, setting=g.uniform(*fcv_lps))
    # PSVs sustain the pressure of a grid junction that drains to a low
    # reservoir: grid junction -> PSV -> junction (5 m lower) -> pipe ->
    # reservoir (head 10 m below the grid junction's elevation)
    free = [a for a in range(nj) if a not in busy]
    for a in (g.choice(free, size=n_psv, replace=False) if n_psv else []):
        a = int(a)
        b = add_node(JUNCTION, elev[a] - 5.0, 0.5)
        r = add_node(RESERVOIR, elev[a] - 10.0)
        links.append(dict(n1=a, n2=b, type=PSV, diam_mm=150.0, setting_m=g.uniform(*psv_m)))
        links.append(dict(n1=b, n2=r, type=PIPE, len_m=200.0, diam_mm=150.0, C=120.0,
                          rough_mm=0.05))
    N, L = len(ntype), len(links)
    lt = np.array([lk["type"] for lk in links], dtype=np.int8)
    ln_ft = np.array([lk.get("len_m", 0.0) for lk in links]) / M_PER_FT
    d_ft = np.array([lk["diam_mm"] for lk in links]) / (1000.0 * M_PER_FT)
    pipe = (lt == PIPE) | (lt == CVPIPE)
    if headloss == "H-W":
        rough = np.array([lk.get("C", 0.0) for lk in links])
        r = np.zeros(L)
        r[pipe] = hw_resistance(rough[pipe], ln_ft[pipe], d_ft[pipe])
    elif headloss == "D-W":
        # roughness in internal units: mm -> ft, stored as the pipe's Kc
        rough = np.array([lk.get("rough_mm", 0.0) for lk in links]) / (1000.0 * M_PER_FT)
        r = np.zeros(L)
        r[pipe] = dw_resistance(ln_ft[pipe], d_ft[pipe])
    else:
        raise ValueError("headloss must be 'H-W' or 'D-W'")
    rough = np.where(pipe, rough, 0.0)
    st = np.array([2 if t in (PRV, PSV, FCV, TCV) else 1 for t in lt], dtype=np.int8)
    setting = np.array([lk.get("setting_m", lk.get("setting", 0.0)) for lk in links])
    for k, _ in pumps:
        setting[k] = 1.0          # relative speed
    nt = np.asarray(ntype, dtype=np.int8)
    elev_ft = np.asarray(elev) / M_PER_FT
    ke = np.zeros(N)
    junc = np.where(nt == JUNCTION)[0]
    leaks = g.choice(junc, size=n_leaks, replace=False) if n_leaks else np.zeros(0, int)
    c_leak = np.zeros(N)
    c_leak[leaks] = g.uniform(0.05, 0.5, n_leaks)
    ke = emitter_ke(c_leak, "LPS", 2.0, M_PER_FT)
    kw = {}
    if pumps:
        kw.update(pump_link=np.array([k for k, _ in pumps], dtype=np.int32),
                  pump_ptype=np.array([PUMP_CUSTOM if "x" in c else PUMP_POWER_FUNC
                                       for _, c in pumps], dtype=np.int8),
                  pump_h0=np.array([c["h0"] for _, c in pumps]),
                  pump_r=np.array([c["r"] for _, c in pumps]),
                  pump_n=np.array([c["n"] for _, c in pumps]),
                  pump_q0=np.array([c["q0"] for _, c in pumps]),
                  pump_qmax=np.array([c["qmax"] for _, c in pumps]),
                  pump_hmax=np.array([c["hmax"] for _, c in pumps]),
                  pump_curve_q=np.concatenate([c.get("x", np.zeros(0)) for _, c in pumps]),
                  pump_curve_h=np.concatenate([c.get("y", np.zeros(0)) for _, c in pumps]),
                  pump_curve_ptr=np.concatenate(
                      [[0], np.cumsum([len(c.get("x", ())) for _, c in pumps])]).astype(np.int32))
    if tank_node is not None:
        kw.update(tank_node=np.array([tank_node], dtype=np.int32),
                  tank_h0=np.array([(40.0 + tank_level_m) / M_PER_FT]),
                  tank_hmin=np.array([(40.0 + 2.0) / M_PER_FT]),
                  tank_hmax=np.array([(40.0 + 25.0) / M_PER_FT]),
                  tank_area=np.array([np.pi * 10.0 ** 2 / M_PER_FT ** 2]),
                  tank_overflow=np.zeros(1, dtype=np.int8))
    meta = dict(flow_units="LPS", headloss=headloss, accuracy=1e-3, trials=200,
                seed=int(seed), workload="feature%dx%d-%s" % (nx, ny, headloss),
                emitter_c=c_leak, leak_nodes=leaks)
    nid = [("R%d" % i if t != JUNCTION else "J%d" % i) for i, t in enumerate(ntype)]
    return Net(node_id=nid, node_type=nt, elev_ft=elev_ft, node_ke=ke,
               demand_cfs=np.asarray(dem) / LPS_PER_CFS,
               link_id=["P%d" % k for k in range(L)], link_type=lt,
               link_n1=np.array([lk["n1"] for lk in links], dtype=np.int32),
               link_n2=np.array([lk["n2"] for lk in links], dtype=np.int32),
               diam_ft=d_ft, len_ft=ln_ft, roughness=rough, km_int=np.zeros(L),
               r_hw=r, init_status=st, valve_setting_user=setting, meta=meta, **kw)


def eps_network(seed=0, nx=6, ny=6, hours=24, headloss="H-W"):
    """Extended-period network: a ``feature_network`` grid whose supply pump
    is switched by the tank's level (LOWLEVEL opens it at 12 m, HILEVEL
    closes it at 20 m) and whose junction demands follow a seeded diurnal
    pattern (24 hourly factors in [0.6, 1.4]); duration ``hours``, hydraulic
    and pattern steps 1 h.  The tank (area 12 m^2) starts at 17 m; the
    gravity reservoir is lowered to 52 m.""" 
    from .network import ST_CLOSED, ST_OPEN
    g = np.random.default_rng(seed + 1000)
    net = feature_network(seed, nx=nx, ny=ny, headloss=headloss, n_pumps=1, n_prv=1,
                          n_tcv=1, n_cv=1, tank=True, n_leaks=2, tank_level_m=17.0)
    base = 1.0 + 0.4 * np.sin(2.0 * np.pi * (np.arange(24) - 6.0) / 24.0)
    pat = np.clip(base + g.uniform(-0.05, 0.05, 24), 0.6, 1.4)
    net.patterns = [pat]
    nt = np.asarray(net.node_type)
    net.dem_pat = np.where(nt == JUNCTION, 0, -1).astype(np.int32)
    tank = int(net.tank_node[0])
    net.tank_area = np.array([12.0 / M_PER_FT ** 2])
    net.tank_hmin = np.array([(40.0 + 2.0) / M_PER_FT])
    net.tank_hmax = np.array([(40.0 + 25.0) / M_PER_FT])
    # the gravity reservoir sits below the full tank, so the tank drains at
    # peak demand and the pump controls switch during the day
    el = np.asarray(net.elev_ft).copy()
    el[(nt == RESERVOIR) & (el > 60.0 / M_PER_FT)] = 52.0 / M_PER_FT
    net.elev_ft = el
    pump = int(net.pump_link[0])
    net.ctl_link = np.array([pump, pump], dtype=np.int32)
    net.ctl_node = np.array([tank, tank], dtype=np.int32)
    net.ctl_type = np.array([0, 1], dtype=np.int8)
    net.ctl_grade = np.array([(40.0 + 12.0) / M_PER_FT, (40.0 + 20.0) / M_PER_FT])
    net.ctl_time = np.zeros(2, dtype=np.int64)
    net.ctl_status = np.array([ST_OPEN, ST_CLOSED], dtype=np.int8)
    net.meta.update(duration_sec=3600 * int(hours), hyd_step_sec=3600, pat_step_sec=3600,
                    report_step_sec=3600, pat_start_sec=0, tstart_sec=0,
                    workload="eps-" + net.meta["workload"])
    return net


def planar_network(seed=1, n_junc=782, n_res=2, n_pipes=905, n_prv=3,
                   n_leaks=10, demand_lps=(0.05, 0.5), prv_setting_m=(30.0, 50.0),
                   res_head_m=(60.0, 90.0)):
    """Seeded random planar network.

    * junctions: uniform points in a square of side 120 m * sqrt(n_junc),
      elevation U[0,50] m, base demand U[demand_lps] L/s;
    * graph: Delaunay triangulation (planar); each reservoir (head
      U[res_head_m] m) sits on the boundary and feeds its nearest junction
      through a 400 mm main;
    * spanning tree: multi-source breadth-first search from the reservoir
      mains with a random neighbour order;
    * chords: random non-tree Delaunay edges until the link count reaches
      n_pipes + n_prv (loop density (links - tree links) / tree links ~0.16);
    * lengths lognormal (median 120 m, sigma 0.8); C U[90,140]; diameters drawn
      from {100..400} mm, the tree draws assigned by downstream-demand rank so
      mains carry the large sizes;
    * n_prv trunk tree links become PRVs (setting U[prv_setting_m] m),
      oriented downstream, both ends junctions, distinct downstream nodes and
      no cascade (no PRV's downstream node is another PRV's upstream node);
    * n_leaks junctions get emitter C U[0,0.5] L/s/m^0.5 (Qexp 2).
    """
    from scipy.spatial import Delaunay

    g = np.random.default_rng(seed)
    side = 120.0 * np.sqrt(n_junc)
    pts = g.uniform(0.0, side, (n_junc, 2))
    tri = Delaunay(pts)
    adj = [set() for _ in range(n_junc)]
    for s in tri.simplices:
        for a in range(3):
            u, v = int(s[a]), int(s[(a + 1) % 3])
            adj[u].add(v)
            adj[v].add(u)
    adj = [sorted(a) for a in adj]
    # reservoir attachment junctions: nearest junction to random boundary points
    roots = []
    for r in range(n_res):
        edge = g.integers(4)
        t = g.uniform(0.1, 0.9) * side
        bp = [(t, 0.0), (side, t), (t, side), (0.0, t)][edge]
        d2 = ((pts - np.asarray(bp)) ** 2).sum(1)
        for j in np.argsort(d2):
            if int(j) not in roots:
                roots.append(int(j))
                break
    # multi-source BFS tree with random neighbour order
    parent = np.full(n_junc, -2, dtype=np.int64)
    depth = np.zeros(n_junc, dtype=np.int64)
    frontier = list(roots)
    for r in roots:
        parent[r] = -1
    while frontier:
        nxt = []
        for u in frontier:
            nb = list(adj[u])
            g.shuffle(nb)
            for v in nb:
                if parent[v] == -2:
                    parent[v] = u
                    depth[v] = depth[u] + 1
                    nxt.append(v)
        g.shuffle(nxt)
        frontier = nxt
    assert (parent != -2).all(), "Delaunay graph is connected"
    tree = [(int(parent[v]), v) for v in range(n_junc) if parent[v] >= 0]
    # subtree demand (for trunk ranking and diameter assignment)
    dem = g.uniform(demand_lps[0], demand_lps[1], n_junc)
    order = np.argsort(-depth)
    sub = dem.copy()
    subn = np.ones(n_junc, dtype=np.int64)
    for v in order:
        p = parent[v]
        if p >= 0:
            sub[p] += sub[v]
            subn[p] += subn[v]
    # node arrays: junctions 0..n_junc-1, reservoirs after
    elev = np.concatenate([g.uniform(0.0, 50.0, n_junc),
                           g.uniform(res_head_m[0], res_head_m[1], n_res)])
    ntype = [JUNCTION] * n_junc + [RESERVOIR] * n_res
    demand = np.concatenate([dem, np.zeros(n_res)])
    # PRV trunk selection: tree links with a meaningful downstream zone
    trunk = [i for i, (p, v) in enumerate(tree) if 15 <= subn[v] <= 150 and parent[p] >= 0]
    g.shuffle(trunk)
    prv_idx, used_up, used_dn = [], set(), set()
    for i in trunk:
        if len(prv_idx) == n_prv:
            break
        p, v = tree[i]
        if v in used_dn or v in used_up or p in used_dn or p in used_up:
            continue
        prv_idx.append(i)
        used_up.add(p)
        used_dn.add(v)
    assert len(prv_idx) == n_prv
    prv_set = set(prv_idx)
    # pressure zones: a junction belongs to the zone of its nearest PRV
    # ancestor (0 = the main zone); chords never cross zones, so every PRV
    # zone is fed through its valve only (a district metered area)
    zone = np.zeros(n_junc, dtype=np.int64)
    zone_of_dn = {tree[i][1]: z + 1 for z, i in enumerate(prv_idx)}
    for v in np.argsort(depth):
        p = parent[v]
        if v in zone_of_dn:
            zone[v] = zone_of_dn[v]
        elif p >= 0:
            zone[v] = zone[p]
    tset = set((min(a, b), max(a, b)) for a, b in tree)
    cand = sorted(set((min(u, v), max(u, v)) for u in range(n_junc) for v in adj[u]
                      if zone[u] == zone[v]) - tset)
    n_links_total = n_pipes + n_prv
    n_tree = len(tree) + n_res
    n_ch = n_links_total - n_tree
    pick = g.choice(len(cand), size=n_ch, replace=False)
    chords = [cand[int(i)] for i in sorted(pick)]
    # diameters: tree draws sorted by subtree demand rank, chords random
    dt_draw = np.sort(g.choice(DIAMS_MM, size=len(tree)))
    rank = np.argsort(np.argsort([sub[v] for _, v in tree]))
    tree_diam = dt_draw[rank]
    links = []
    for r in range(n_res):
        links.append(dict(n1=n_junc + r, n2=roots[r], type=PIPE,
                          len_m=float(np.exp(np.log(120.0) + 0.8 * g.standard_normal())),
                          diam_mm=400.0, C=float(g.uniform(90.0, 140.0))))
    for i, (p, v) in enumerate(tree):
        lk = dict(n1=p, n2=v, type=PIPE,
                  len_m=float(np.exp(np.log(120.0) + 0.8 * g.standard_normal())),
                  diam_mm=float(tree_diam[i]), C=float(g.uniform(90.0, 140.0)))
        if i in prv_set:
            lk["type"] = PRV
            lk["setting_m"] = float(g.uniform(prv_setting_m[0], prv_setting_m[1]))
            lk["C"] = 0.0
            lk["len_m"] = 0.0
        links.append(lk)
    for a, b in chords:
        if g.random() < 0.5:
            a, b = b, a
        links.append(dict(n1=a, n2=b, type=PIPE,
                          len_m=float(np.exp(np.log(120.0) + 0.8 * g.standard_normal())),
                          diam_mm=float(g.choice(DIAMS_MM)), C=float(g.uniform(90.0, 140.0))))
    leaks = g.choice(n_junc, size=n_leaks, replace=False)
    c_leak = np.zeros(n_junc + n_res)
    c_leak[leaks] = g.uniform(0.0, 0.5, n_leaks)
    ke = emitter_ke(c_leak, "LPS", 2.0, M_PER_FT)
    meta = dict(flow_units="LPS", headloss="H-W", accuracy=1e-3, trials=200,
                seed=int(seed), workload="planar%d" % n_junc,
                leak_nodes=np.sort(leaks).tolist(), emitter_c=c_leak)
    return _mk_net(ntype, elev, demand, links, meta, ke=ke)


SIZES_MM = np.array([100.0, 150.0, 200.0, 250.0, 300.0, 400.0, 500.0, 600.0, 800.0, 1000.0,
                     1200.0, 1500.0])


def _hw_loss_m(q_m3s, len_m, d_m, c):
    """Hazen-Williams head loss in m (SI form 10.67 L Q^1.852 / C^1.852 /
    D^4.871); only used to place tanks and pump suctions."""
    return 10.67 * len_m * np.abs(q_m3s) ** 1.852 / (c ** 1.852 * d_m ** 4.871)


def geometric_network(seed=0, n_junc=20000, n_res=4, n_tanks=6, n_pumps=8, n_links=None,
                      k=3, n_leaks=10, hours=24):
    """Seeded extended-period network on a random geometric graph.

    * junctions: uniform points in a square of side 120 m * sqrt(n_junc),
      elevation U[0,50] m, base demand U[0.05,0.5] L/s following one seeded
      diurnal pattern (24 hourly factors in [0.6, 1.4]);
    * pipes: the Euclidean minimum spanning tree of the points (connected),
      plus chords drawn from the k nearest-neighbour pairs until the network
      has ``n_links`` links (default 1.2 n_junc); lengths are the point
      distances, C U[90,140];
    * n_res reservoirs (head U[90,110] m) on the square's boundary, each
      feeding its nearest junction through a 100 m main;
    * diameters: a multi-source breadth-first tree from the reservoir mains
      carries each junction's demand to its source; a tree pipe gets the
      smallest standard size (100-1500 mm) with velocity <= 1 m/s at its
      subtree demand, chords 100-200 mm;
    * n_pumps pumps lift from low reservoirs into random junctions through a
      3-point curve (0, h0), (qd, 0.75 h0), (2 qd, 0), shutoff h0 U[40,80] m,
      design flow qd U[20,200] L/s; the suction head is the junction's head
      estimate (tree path losses from its source) minus 0.75 h0;
    * n_tanks cylindrical tanks (diameter U[10,30] m, initial level U[2,6] m,
      levels 0.5..10 m) joined to random junctions by a 500 m, 150 mm pipe,
      the bottom placed at the junction's head estimate minus the initial
      level;
    * controls: pump i is opened when tank i mod n_tanks falls 0.5 m below
      its initial level (LOWLEVEL) and closed when it rises 0.5 m above it
      (HILEVEL);
    * n_leaks junctions get emitter C U[0,0.5] L/s/m^0.5.
    Hydraulic and pattern steps 1 h, duration ``hours``."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    from scipy.spatial import Delaunay, cKDTree

    from .network import ST_CLOSED, ST_OPEN
    g = np.random.default_rng(seed)
    nj = int(n_junc)
    n_links = int(round(1.2 * nj)) if n_links is None else int(n_links)
    side = 120.0 * np.sqrt(nj)
    pts = g.uniform(0.0, side, (nj, 2))
    # Euclidean MST: a subgraph of the Delaunay triangulation
    tri = Delaunay(pts)
    e = np.concatenate([tri.simplices[:, [0, 1]], tri.simplices[:, [1, 2]],
                        tri.simplices[:, [0, 2]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    w = np.hypot(*(pts[e[:, 0]] - pts[e[:, 1]]).T)
    mst = minimum_spanning_tree(coo_matrix((w, (e[:, 0], e[:, 1])), shape=(nj, nj))).tocoo()
    tree_e = np.sort(np.stack([mst.row, mst.col], 1), axis=1)
    # chords from the k-nearest-neighbour pairs
    _, nb = cKDTree(pts).query(pts, k + 1)
    knn = np.sort(np.stack([np.repeat(np.arange(nj), k), nb[:, 1:].ravel()], 1), axis=1)
    knn = np.unique(knn, axis=0)
    tkey = set(map(tuple, tree_e.tolist()))
    cand = np.array([p for p in knn.tolist() if tuple(p) not in tkey], dtype=np.int64)
    n_fixed_links = n_res + n_pumps + n_tanks
    n_ch = max(0, min(len(cand), n_links - n_fixed_links - len(tree_e)))
    chords = cand[np.sort(g.choice(len(cand), size=n_ch, replace=False))]
    pipes = np.concatenate([tree_e, chords])
    is_tree = np.arange(len(pipes)) < len(tree_e)
    adj = [[] for _ in range(nj)]
    for i, (a, b) in enumerate(pipes.tolist()):
        adj[a].append((b, i))
        adj[b].append((a, i))
    # sources on the boundary
    roots = []
    for r in range(n_res):
        t = g.uniform(0.1, 0.9) * side
        bp = [(t, 0.0), (side, t), (t, side), (0.0, t)][r % 4]
        d2 = ((pts - np.asarray(bp)) ** 2).sum(1)
        roots.append(int(next(j for j in np.argsort(d2) if int(j) not in roots)))
    res_head = g.uniform(90.0, 110.0, n_res)
    elev = g.uniform(0.0, 50.0, nj)
    dem = g.uniform(0.05, 0.5, nj)
    C = g.uniform(90.0, 140.0, len(pipes))
    # multi-source BFS over the pipe graph
    parent = np.full(nj, -2, dtype=np.int64)
    plink = np.full(nj, -1, dtype=np.int64)
    src = np.full(nj, -1, dtype=np.int64)
    order = []
    frontier = []
    for r, j in enumerate(roots):
        parent[j], src[j] = -1, r
        frontier.append(j)
    while frontier:
        order.extend(frontier)
        nxt = []
        for u in frontier:
            for v, i in adj[u]:
                if parent[v] == -2:
                    parent[v], plink[v], src[v] = u, i, src[u]
                    nxt.append(v)
        frontier = nxt
    assert (parent != -2).all(), "the spanning tree reaches every junction"
    sub = dem.copy()
    for v in reversed(order):
        if parent[v] >= 0:
            sub[parent[v]] += sub[v]
    # sizes: velocity <= 1 m/s at the subtree demand (BFS-tree pipes)
    diam = g.choice(SIZES_MM[:3], size=len(pipes))
    qsub = np.zeros(len(pipes))
    for v in range(nj):
        if plink[v] >= 0:
            qsub[plink[v]] = sub[v] / 1000.0
    in_bfs = qsub > 0.0
    need = np.sqrt(4.0 * qsub / np.pi) * 1000.0
    idx = np.minimum(np.searchsorted(SIZES_MM, need), len(SIZES_MM) - 1)
    diam = np.where(in_bfs, np.maximum(SIZES_MM[idx], diam), diam)
    length = np.hypot(*(pts[pipes[:, 0]] - pts[pipes[:, 1]]).T)
    length = np.maximum(length, 1.0)
    # head estimate along the BFS tree
    main_q = np.array([sub[j] for j in roots]) / 1000.0
    main_d = SIZES_MM[np.minimum(np.searchsorted(SIZES_MM, np.sqrt(4.0 * main_q / np.pi)
                                                 * 1000.0), len(SIZES_MM) - 1)]
    h_est = np.zeros(nj)
    for r, j in enumerate(roots):
        h_est[j] = res_head[r] - _hw_loss_m(main_q[r], 100.0, main_d[r] / 1000.0, 130.0)
    for v in order:
        p = parent[v]
        if p >= 0:
            i = plink[v]
            h_est[v] = h_est[p] - _hw_loss_m(sub[v] / 1000.0, length[i], diam[i] / 1000.0, C[i])
    # node and link lists
    ntype = [JUNCTION] * nj
    elev_l = list(elev)
    dem_l = list(dem)
    links = []
    for i, (a, b) in enumerate(pipes.tolist()):
        if g.random() < 0.5:
            a, b = b, a
        links.append(dict(n1=a, n2=b, type=PIPE, len_m=float(length[i]),
                          diam_mm=float(diam[i]), C=float(C[i])))

    def add_node(t, el):
        ntype.append(t)
        elev_l.append(el)
        dem_l.append(0.0)
        return len(ntype) - 1

    for r, j in enumerate(roots):
        n = add_node(RESERVOIR, float(res_head[r]))
        links.append(dict(n1=n, n2=j, type=PIPE, len_m=100.0, diam_mm=float(main_d[r]),
                          C=130.0))
    interior = np.setdiff1d(np.arange(nj), roots)
    pick = g.choice(interior, size=n_pumps + n_tanks, replace=False)
    pumps = []
    for i, j in enumerate(pick[:n_pumps]):
        h0 = g.uniform(40.0, 80.0)
        qd = g.uniform(20.0, 200.0)
        n = add_node(RESERVOIR, float(h_est[j] - 0.75 * h0))
        pumps.append((len(links), pump_power_curve(h0, 0.75 * h0, qd, 0.0, 2.0 * qd)))
        links.append(dict(n1=n, n2=int(j), type=PUMP, diam_mm=0.0))
    tanks, level0, tdiam, tmax = [], [], [], []
    for i, j in enumerate(pick[n_pumps:]):
        lv = g.uniform(2.0, 6.0)
        n = add_node(TANK, float(h_est[j] - lv))
        tanks.append(n)
        level0.append(lv)
        tdiam.append(g.uniform(10.0, 30.0))
        links.append(dict(n1=n, n2=int(j), type=PIPE, len_m=500.0, diam_mm=150.0,
                          C=float(g.uniform(90.0, 140.0))))
    N = len(ntype)
    leaks = g.choice(nj, size=min(n_leaks, nj), replace=False)
    c_leak = np.zeros(N)
    c_leak[leaks] = g.uniform(0.0, 0.5, leaks.size)
    ke = emitter_ke(c_leak, "LPS", 2.0, M_PER_FT)
    meta = dict(flow_units="LPS", headloss="H-W", accuracy=1e-3, trials=200, seed=int(seed),
                workload="geometric%d" % nj, leak_nodes=np.sort(leaks).tolist(),
                emitter_c=c_leak, duration_sec=3600 * int(hours), hyd_step_sec=3600,
                pat_step_sec=3600, report_step_sec=3600, pat_start_sec=0, tstart_sec=0)
    net = _mk_net(ntype, elev_l, dem_l, links, meta, ke=ke)
    for k_, c in pumps:
        net.valve_setting_user[k_] = 1.0
    if pumps:
        net.pump_link = np.array([k_ for k_, _ in pumps], dtype=np.int32)
        net.pump_ptype = np.full(len(pumps), PUMP_POWER_FUNC, dtype=np.int8)
        for key in ("h0", "r", "n", "q0", "qmax", "hmax"):
            setattr(net, "pump_" + key, np.array([c[key] for _, c in pumps]))
        net.pump_curve_ptr = np.zeros(len(pumps) + 1, dtype=np.int32)
    if tanks:
        el_t = np.asarray(elev_l)[tanks]
        net.tank_node = np.asarray(tanks, dtype=np.int32)
        net.tank_h0 = (el_t + np.asarray(level0)) / M_PER_FT
        net.tank_hmin = (el_t + 0.5) / M_PER_FT
        net.tank_hmax = (el_t + 10.0) / M_PER_FT
        net.tank_area = np.pi * (np.asarray(tdiam) / M_PER_FT) ** 2 / 4.0
        net.tank_overflow = np.zeros(len(tanks), dtype=np.int8)
    base = 1.0 + 0.4 * np.sin(2.0 * np.pi * (np.arange(24) - 6.0) / 24.0)
    net.patterns = [np.clip(base + g.uniform(-0.05, 0.05, 24), 0.6, 1.4)]
    net.dem_pat = np.where(np.asarray(ntype) == JUNCTION, 0, -1).astype(np.int32)
    if pumps and tanks:
        cl, cn, ct, cg, cs = [], [], [], [], []
        for i, (k_, _) in enumerate(pumps):
            j = i % len(tanks)
            el_j = float(np.asarray(elev_l)[tanks[j]])
            for typ, dl, st in ((0, -0.5, ST_OPEN), (1, 0.5, ST_CLOSED)):
                cl.append(k_)
                cn.append(tanks[j])
                ct.append(typ)
                cg.append((el_j + level0[j] + dl) / M_PER_FT)
                cs.append(st)
        net.ctl_link = np.asarray(cl, dtype=np.int32)
        net.ctl_node = np.asarray(cn, dtype=np.int32)
        net.ctl_type = np.asarray(ct, dtype=np.int8)
        net.ctl_grade = np.asarray(cg)
        net.ctl_time = np.zeros(len(cl), dtype=np.int64)
        net.ctl_status = np.asarray(cs, dtype=np.int8)
    return net


def ragged_member(seed):
    """One member of the ragged multi-network batch: a seeded planar network
    (planar_network) of log-uniform size Nj in [50, 5000], Hazen-Williams or
    Darcy-Weisbach friction (roughness U[0.01, 0.5] mm), 0-4 pumps lifting
    from low reservoirs (head U[10, 30] m) into random junctions, each with a
    3-point head curve (0, h0), (qd, 0.75 h0), (2 qd, 0), shutoff h0 U[40, 80]
    m and design flow qd U[20, 200] L/s, and 0-3 valves: PRVs on trunk mains
    and TCVs (loss coefficient U[2, 10]) on random pipes."""
    from .network import TCV as _TCV
    g = np.random.default_rng(10_000 + int(seed))
    nj = int(round(np.exp(g.uniform(np.log(50.0), np.log(5000.0)))))
    headloss = "D-W" if g.random() < 0.5 else "H-W"
    n_pumps = int(g.integers(0, 5))
    n_valves = int(g.integers(0, 4))
    n_prv = int(g.integers(0, n_valves + 1))
    n_tcv = n_valves - n_prv
    while True:
        try:
            net = planar_network(int(seed), n_junc=nj, n_pipes=int(round(nj * 905 / 782)),
                                 n_prv=n_prv, n_leaks=min(10, max(1, nj // 50)))
            break
        except AssertionError:      # too few trunk mains for the PRVs
            if n_prv == 0:
                raise
            n_prv -= 1
            n_tcv += 1
    lt = np.asarray(net.link_type).copy()
    st = np.asarray(net.init_status).copy()
    setting = np.asarray(net.valve_setting_user, dtype=np.float64).copy()
    rough = np.asarray(net.roughness, dtype=np.float64).copy()
    r = np.asarray(net.r_hw, dtype=np.float64).copy()
    pipes = np.nonzero(lt == PIPE)[0]
    for k in g.choice(pipes, size=n_tcv, replace=False):
        lt[k] = TCV
        st[k] = 2
        setting[k] = g.uniform(2.0, 10.0)
        rough[k] = 0.0
        r[k] = 0.0
    pipe = lt == PIPE
    if headloss == "D-W":
        rough = np.where(pipe, g.uniform(0.01, 0.5, lt.size) / (1000.0 * M_PER_FT), 0.0)
        r = np.where(pipe, dw_resistance(np.where(pipe, net.len_ft, 1.0),
                                         np.where(pipe, net.diam_ft, 1.0)), 0.0)
    nt = list(np.asarray(net.node_type))
    elev = list(np.asarray(net.elev_ft))
    dem = list(np.asarray(net.demand_cfs))
    ke = list(np.asarray(net.node_ke))
    n1, n2 = list(net.link_n1), list(net.link_n2)
    diam, length, km = list(net.diam_ft), list(net.len_ft), list(net.km_int)
    lt, st, setting, rough, r = list(lt), list(st), list(setting), list(rough), list(r)
    link_id, node_id = list(net.link_id), list(net.node_id)
    curves = []
    junc = np.nonzero(np.asarray(net.node_type) == JUNCTION)[0]
    for j in g.choice(junc, size=n_pumps, replace=False):
        res = len(nt)
        nt.append(RESERVOIR)
        elev.append(g.uniform(10.0, 30.0) / M_PER_FT)
        dem.append(0.0)
        ke.append(0.0)
        node_id.append("RP%d" % res)
        h0 = g.uniform(40.0, 80.0)
        qd = g.uniform(20.0, 200.0)
        curves.append((len(lt), pump_power_curve(h0, 0.75 * h0, qd, 0.0, 2.0 * qd)))
        n1.append(res)
        n2.append(int(j))
        lt.append(PUMP)
        st.append(1)
        setting.append(1.0)
        for arr, v in ((diam, 0.0), (length, 0.0), (km, 0.0), (rough, 0.0), (r, 0.0)):
            arr.append(v)
        link_id.append("PU%d" % len(link_id))
    kw = {}
    if curves:
        kw = dict(pump_link=np.array([k for k, _ in curves], dtype=np.int32),
                  pump_ptype=np.full(len(curves), PUMP_POWER_FUNC, dtype=np.int8),
                  pump_h0=np.array([c["h0"] for _, c in curves]),
                  pump_r=np.array([c["r"] for _, c in curves]),
                  pump_n=np.array([c["n"] for _, c in curves]),
                  pump_q0=np.array([c["q0"] for _, c in curves]),
                  pump_qmax=np.array([c["qmax"] for _, c in curves]),
                  pump_hmax=np.array([c["hmax"] for _, c in curves]))
    meta = dict(net.meta)
    meta.update(headloss=headloss, workload="ragged%d-%s" % (nj, headloss), seed=int(seed),
                pumps=n_pumps, prvs=n_prv, tcvs=n_tcv)
    c_leak = np.concatenate([np.asarray(meta.get("emitter_c", np.zeros(net.N))),
                             np.zeros(n_pumps)])
    meta["emitter_c"] = c_leak
    return Net(node_id=node_id, node_type=np.asarray(nt, dtype=np.int8),
               elev_ft=np.asarray(elev), node_ke=np.asarray(ke), demand_cfs=np.asarray(dem),
               link_id=link_id, link_type=np.asarray(lt, dtype=np.int8),
               link_n1=np.asarray(n1, dtype=np.int32), link_n2=np.asarray(n2, dtype=np.int32),
               diam_ft=np.asarray(diam), len_ft=np.asarray(length), roughness=np.asarray(rough),
               km_int=np.asarray(km), r_hw=np.asarray(r),
               init_status=np.asarray(st, dtype=np.int8),
               valve_setting_user=np.asarray(setting), meta=meta, **kw)


def ragged_batch(seed=0, n_net=64, total=4096, b_range=(16, 256)):
    """The ragged workload: n_net ragged_member networks (seeds seed*1000 + i)
    with scenario counts log-uniform in b_range, scaled to about ``total``
    solves.  Returns [(net, B)]."""
    g = np.random.default_rng(seed)
    nets = [ragged_member(seed * 1000 + i) for i in range(n_net)]
    b = np.exp(g.uniform(np.log(b_range[0]), np.log(b_range[1]), n_net))
    b = np.clip(np.round(b * total / b.sum()), b_range[0], b_range[1]).astype(int)
    return list(zip(nets, b.tolist()))


def scenario_batch(net, B, seed, sigma=0.3):
    """Per-scenario demands: every junction base demand times an independent
    lognormal(0, sigma) factor.  Returns (demand_cfs[B,N], res_head_ft[B,N])
    float64; reservoir heads are the network's fixed heads."""
    g = np.random.default_rng(seed)
    base = np.asarray(net.demand_cfs, dtype=np.float64)
    mult = np.exp(sigma * g.standard_normal((B, net.N)))
    d = base[None, :] * mult
    rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), (B, net.N)).copy()
    return d, rh


def roughness_truth(net, seed, sd=5.0):
    """'True' Hazen-Williams C field: C + N(0, sd) on pipes (other links 0)."""
    g = np.random.default_rng(seed)
    C = np.asarray(net.roughness, dtype=np.float64).copy()
    pipe = np.asarray(net.link_type) == PIPE
    C[pipe] = C[pipe] + sd * g.standard_normal(int(pipe.sum()))
    return C


def sensor_nodes(net, n, seed):
    """n distinct random junction indices (sorted)."""
    g = np.random.default_rng(seed)
    junc = np.where(np.asarray(net.node_type) == JUNCTION)[0]
    return np.sort(g.choice(junc, size=n, replace=False))
