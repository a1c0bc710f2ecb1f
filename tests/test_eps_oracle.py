# -*- coding: utf-8 -*-
"""Pins the oracle's extended-period driver with its invariants: tank mass
balance between frames, the time grid (hydraulic steps, pattern
boundaries, control and tank events), and the level controls.  CPU only."""
import numpy as np
import pytest

from oracle.eps import run_eps
from wdsgpu_amd.generators import eps_network
from wdsgpu_amd.network import M_PER_FT, ST_CLOSED, ST_OPEN


@pytest.fixture(scope="module", params=[0, 1, 3])
def eps_case(request):
    net = eps_network(request.param)
    return net, run_eps(net)


def test_every_hour_is_a_frame_and_the_run_ends_at_the_duration(eps_case):
    net, f = eps_case
    t = f["t_sec"]
    assert t[0] == 0 and t[-1] == net.meta["duration_sec"]
    assert np.all(np.diff(t) > 0) and np.all(np.diff(t) <= 3600)
    assert set(range(0, 86401, 3600)) <= set(t.tolist())


def test_tank_volume_follows_the_inflow_of_the_previous_frame(eps_case):
    net, f = eps_case
    vmin, vmax, _ = net.tank_volumes()
    a, hmin = net.tank_area[0], net.tank_hmin[0]
    tank = int(net.tank_node[0])
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
    for i in range(len(f["t_sec"]) - 1):
        q = f["flow_cfs"][i]
        inflow = q[n2 == tank].sum() - q[n1 == tank].sum()
        dt = f["t_sec"][i + 1] - f["t_sec"][i]
        v0 = vmin[0] + (f["tank_head_ft"][i, 0] - hmin) * a
        v1 = vmin[0] + (f["tank_head_ft"][i + 1, 0] - hmin) * a
        expect = np.clip(v0 + inflow * dt, vmin[0], vmax[0])
        # the one-second look-ahead clamp snaps to a bound within one
        # second of flow of it
        if abs(expect - vmax[0]) <= abs(inflow) or abs(expect - vmin[0]) <= abs(inflow):
            continue
        assert abs(v1 - expect) <= 1e-9 * vmax[0]


def test_level_controls_switch_the_pump(eps_case):
    net, f = eps_case
    pump = int(net.pump_link[0])
    lvl = f["tank_head_ft"][:, 0] * M_PER_FT - 40.0
    st = f["status"][:, pump]
    for i in range(len(lvl)):
        if abs(lvl[i] - 20.0) < 1e-6:      # HILEVEL reached: the pump is shut
            assert st[i] <= ST_CLOSED
        if abs(lvl[i] - 12.0) < 1e-6 and i > 0 and st[i - 1] <= ST_CLOSED:
            assert st[i] != ST_CLOSED      # LOWLEVEL reopens it (may then go XHEAD)
    # event frames between the hours exist (control and tank times)
    assert any(t % 3600 for t in f["t_sec"])
