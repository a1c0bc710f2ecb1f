# -*- coding: utf-8 -*-
"""INP writer: read -> write -> read gives the same network, for the 23
networks the reference ships and for the seeded networks of the benchmark
configurations (pumps of every curve kind, PRVs, TCVs, check valves, tanks,
emitters, patterns, times and controls).  CPU only."""
import os

import numpy as np
import pytest

from wdsgpu_amd.generators import (eps_network, feature_network, grid_network,
                                   planar_network)
from wdsgpu_amd.network import read_inp, write_inp

from test_inp import NETS, SIZES



NODE_F = ("node_type", "elev_ft", "demand_cfs", "node_ke", "dem_pat")
LINK_F = ("link_type", "init_status", "diam_ft", "len_ft", "roughness", "km_int", "r_hw",
          "valve_setting_user")
TANK_F = ("tank_h0", "tank_hmin", "tank_hmax", "tank_area", "tank_overflow")
PUMP_F = ("pump_ptype", "pump_h0", "pump_r", "pump_n", "pump_q0", "pump_qmax", "pump_hmax")
CTL_F = ("ctl_type", "ctl_grade", "ctl_time", "ctl_status")


def _eq(x, y, where, f):
    x, y = np.asarray(x), np.asarray(y)
    assert x.shape == y.shape, (where, f)
    if x.dtype.kind == "f" and x.size:
        assert np.allclose(x, y, rtol=1e-13, atol=1e-300), (where, f, np.abs(x - y).max())
    else:
        assert np.array_equal(x, y), (where, f)


def _same(a, b, where):
    """a and b describe the same network.  An INP file lists links by
    section (pipes, pumps, valves) and nodes by kind, so b's order may
    differ from a's: b is compared through the id maps."""
    assert sorted(a.node_id) == sorted(b.node_id), where
    assert sorted(a.link_id) == sorted(b.link_id), where
    bn = {n: i for i, n in enumerate(b.node_id)}
    bl = {n: i for i, n in enumerate(b.link_id)}
    pn = np.array([bn[n] for n in a.node_id])            # a node -> b node
    pl = np.array([bl[n] for n in a.link_id])
    inv_n = np.empty_like(pn)
    inv_n[pn] = np.arange(pn.size)
    inv_l = np.empty_like(pl)
    inv_l[pl] = np.arange(pl.size)
    for f in NODE_F:
        _eq(getattr(a, f), np.asarray(getattr(b, f))[pn], where, f)
    pipe = np.asarray(a.link_type) <= 1
    for f in LINK_F:
        x, y = np.asarray(getattr(a, f)), np.asarray(getattr(b, f))[pl]
        if f in ("len_ft", "roughness"):
            # an INP file gives valves and pumps no length or roughness
            x, y = x[pipe], y[pipe]
        _eq(x, y, where, f)
    _eq(a.link_n1, inv_n[np.asarray(b.link_n1)[pl]], where, "link_n1")
    _eq(a.link_n2, inv_n[np.asarray(b.link_n2)[pl]], where, "link_n2")
    # tanks, pumps and controls: the same sets, matched by a's indices
    ta = np.argsort(a.tank_node)
    tb = np.argsort(inv_n[b.tank_node])
    _eq(np.asarray(a.tank_node)[ta], inv_n[b.tank_node][tb], where, "tank_node")
    for f in TANK_F:
        _eq(np.asarray(getattr(a, f))[ta], np.asarray(getattr(b, f))[tb], where, f)
    pa = np.argsort(a.pump_link)
    pb = np.argsort(inv_l[b.pump_link])
    _eq(np.asarray(a.pump_link)[pa], inv_l[b.pump_link][pb], where, "pump_link")
    for f in PUMP_F:
        _eq(np.asarray(getattr(a, f))[pa], np.asarray(getattr(b, f))[pb], where, f)
    for j, i in zip(pa, pb):
        qa = a.pump_curve_q[a.pump_curve_ptr[j]:a.pump_curve_ptr[j + 1]]
        qb = b.pump_curve_q[b.pump_curve_ptr[i]:b.pump_curve_ptr[i + 1]]
        ha = a.pump_curve_h[a.pump_curve_ptr[j]:a.pump_curve_ptr[j + 1]]
        hb = b.pump_curve_h[b.pump_curve_ptr[i]:b.pump_curve_ptr[i + 1]]
        _eq(qa, qb, where, "pump_curve_q")
        _eq(ha, hb, where, "pump_curve_h")
    _eq(a.ctl_link, inv_l[b.ctl_link], where, "ctl_link")
    _eq(a.ctl_node, np.where(b.ctl_node >= 0, inv_n[np.maximum(b.ctl_node, 0)], -1), where,
        "ctl_node")
    for f in CTL_F:
        _eq(getattr(a, f), getattr(b, f), where, f)
    assert len(a.patterns) == len(b.patterns)
    for p, q in zip(a.patterns, b.patterns):
        assert np.array_equal(p, q), where
    for k in ("flow_units", "headloss", "trials", "accuracy", "checkfreq", "maxcheck",
              "extra_iter", "duration_sec", "hyd_step_sec", "pat_step_sec", "pat_start_sec"):
        if k in a.meta or k in b.meta:
            assert a.meta.get(k) == b.meta.get(k), (where, k)
    for k in ("qexp", "spgrav", "demand_multiplier", "viscosity"):
        assert np.isclose(a.meta.get(k, 1.0), b.meta.get(k, 1.0), rtol=1e-15, atol=0), \
            (where, k)


@pytest.mark.parametrize("rel", sorted(SIZES))
def test_shipped_networks_round_trip(rel, tmp_path):
    a = read_inp(os.path.join(NETS, rel))
    p = tmp_path / "net.inp"
    write_inp(a, p)
    _same(a, read_inp(p), rel)


@pytest.mark.parametrize("name,make", [
    ("grid", lambda: grid_network(0)),
    ("planar", lambda: planar_network(1)),
    ("feature-hw", lambda: feature_network(0)),
    ("feature-dw", lambda: feature_network(1, headloss="D-W")),
    ("feature-custom", lambda: feature_network(2, n_custom=1)),
    ("eps", lambda: eps_network(0)),
])
def test_generated_networks_round_trip(name, make, tmp_path):
    a = make()
    p = tmp_path / (name + ".inp")
    write_inp(a, p)
    b = read_inp(p)
    _same(a, b, name)
    # and again from the network read back
    q = tmp_path / (name + "2.inp")
    write_inp(b, q)
    _same(b, read_inp(q), name + " (second pass)")
