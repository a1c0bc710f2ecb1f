# -*- coding: utf-8 -*-
"""INP reader: the 23 networks the reference ships (networks/random_main,
networks/random_small) against the sizes the reference's own reports hold,
and the sections a user model needs (demands, pumps, curves, patterns,
times, controls) against hand-computed values.  CPU only."""
import os

import numpy as np
import pytest

from oracle import gga
from wdsgpu_amd.network import (PUMP, PUMP_CONST_HP, PUMP_CUSTOM, PUMP_POWER_FUNC,
                                ST_CLOSED, ST_OPEN, pump_power_curve, read_inp)

HERE = os.path.dirname(os.path.abspath(__file__))
# the reference's networks/ directory, copied as data fixtures
NETS = os.path.join(HERE, "golden", "networks")

# (nodes, links) per network: data/benchmark_report.txt:36-58 of the
# reference (rows rand_main_0000 .. rand_small_0002)
SIZES = {"random_main/rand_%04d.inp" % i: nl for i, nl in enumerate(
    [(104, 176), (46, 67), (101, 135), (36, 52), (81, 143), (87, 141), (37, 61), (65, 105),
     (74, 148), (39, 50), (39, 63), (72, 133), (89, 162), (72, 105), (106, 220), (47, 79),
     (98, 142), (89, 178), (117, 235), (85, 168)])}
SIZES.update({"random_small/rand_%04d.inp" % i: nl for i, nl in
              enumerate([(43, 68), (117, 181), (73, 122)])})
# junctions and full symmetric CSR pattern size (data/p4_guards_wip.txt:59)
CSR = {"random_main/rand_0009.inp": (38, 136), "random_main/rand_0014.inp": (104, 540)}


def _shipped(rel):
    return os.path.join(NETS, rel)


@pytest.mark.parametrize("rel", sorted(SIZES))
def test_shipped_networks_read_with_reference_sizes(rel):
    net = read_inp(_shipped(rel))
    assert (net.N, net.L) == SIZES[rel]
    assert net.meta["flow_units"] == "LPS" and net.meta["headloss"] == "H-W"
    assert net.demand_cfs.sum() > 0.0 and len(net.pump_link) == 0
    if rel in CSR:
        m = gga.GGAModel(net)
        assert (m.Nj, m.csr_pos.size) == CSR[rel]


INP = """
[TITLE]
mixed sections
[JUNCTIONS]
 J1   10   2.0   P1
 J2   12   3.0
 J3   11   0
[RESERVOIRS]
 R1   50
[TANKS]
 T1   30   5   1   9   10   0
[PIPES]
 P1   R1  J1  1000  300  120  0  Open
 P2   J1  J2   500  200  110
 P3   J2  J3   400  150  100  0  CV
 P4   J3  T1   300  200  100
[PUMPS]
 PU1  R1  J2  HEAD C1
 PU2  R1  J3  POWER 5
 PU3  R1  J1  HEAD C3  SPEED 1.2
 PU4  R1  J2  HEAD C4
[CURVES]
 C1   10   40
 C3    0   60
 C3   20   50
 C3   40   20
 C4    5   50
 C4   15   45
 C4   25   30
 C4   35    5
[DEMANDS]
 J2   1.5  P1
 J2   0.5  P1
[PATTERNS]
 P1   1.0  1.2
 P1   0.8
[TIMES]
 Duration           24:00
 Hydraulic Timestep 1:00
 Pattern Timestep   2 HOURS
 Pattern Start      0:30
[CONTROLS]
 LINK PU1 CLOSED IF NODE T1 ABOVE 8
 LINK PU1 OPEN AT TIME 6
 LINK P2 CLOSED AT CLOCKTIME 3 PM
[OPTIONS]
 Units  LPS
 Headloss H-W
 Accuracy 0.0001
 Trials 50
[END]
"""


def test_sections_of_a_user_model(tmp_path):
    p = tmp_path / "m.inp"
    p.write_text(INP)
    net = read_inp(str(p))
    idx = {n: i for i, n in enumerate(net.node_id)}
    lps = 28.317
    # [DEMANDS] rows replace J2's junction demand (1.5 + 0.5); J1 keeps its
    # own demand and pattern, J3 has none
    assert net.demand_cfs[idx["J1"]] == pytest.approx(2.0 / lps)
    assert net.demand_cfs[idx["J2"]] == pytest.approx(2.0 / lps)
    assert net.demand_cfs[idx["J3"]] == 0.0
    assert net.dem_pat[idx["J1"]] == 0 and net.patterns[0].tolist() == [1.0, 1.2, 0.8]
    assert net.dem_pat[idx["J2"]] == 0 and net.dem_pat[idx["J3"]] == -1
    # pattern period (t + 30 min) // 2 h: t = 1.5 h is period 1 (factor 1.2)
    assert net.demand_cfs_at(5400)[idx["J1"]] == pytest.approx(1.2 * 2.0 / lps)
    assert net.meta["duration_sec"] == 86400 and net.meta["hyd_step_sec"] == 3600
    assert net.meta["pat_step_sec"] == 7200 and net.meta["pat_start_sec"] == 1800
    assert net.meta["accuracy"] == 1e-4 and net.meta["trials"] == 50
    # pumps: 1-point curve -> power function through (0, 1.33334 h1),
    # (q1, h1), (2 q1, 0); POWER -> constant horsepower, r = -8.814 kW/0.7457;
    # 3-point curve from zero flow -> power function; 4 points -> CUSTOM
    lt = np.asarray(net.link_type)
    assert (lt == PUMP).sum() == 4
    assert net.pump_ptype.tolist() == [PUMP_POWER_FUNC, PUMP_CONST_HP, PUMP_POWER_FUNC,
                                       PUMP_CUSTOM]
    c1 = pump_power_curve(1.33334 * 40.0, 40.0, 10.0, 0.0, 20.0)
    assert net.pump_h0[0] == pytest.approx(c1["h0"]) and net.pump_n[0] == pytest.approx(c1["n"])
    assert net.pump_r[1] == pytest.approx(-8.814 * 5.0 / 0.7457)
    c3 = pump_power_curve(60.0, 50.0, 20.0, 20.0, 40.0)
    assert net.pump_r[2] == pytest.approx(c3["r"])
    assert net.valve_setting_user[net.pump_link[2]] == 1.2
    a, b = net.pump_curve_ptr[3], net.pump_curve_ptr[4]
    assert net.pump_curve_q[a:b].tolist() == [5.0, 15.0, 25.0, 35.0]
    # tank: levels above its bottom, diameter 10 m
    t = int(net.tank_node[0])
    assert net.tank_hmax[0] == pytest.approx((30.0 + 9.0) / 0.3048)
    assert net.tank_area[0] == pytest.approx(3.141592654 * (10 / 0.3048) ** 2 / 4)
    # controls: level control on the tank (El + 8 m), TIMER at 6 h,
    # TIMEOFDAY 3 PM
    assert net.ctl_type.tolist() == [1, 2, 3]
    assert net.ctl_node[0] == t
    assert net.ctl_grade[0] == pytest.approx((30.0 + 8.0) / 0.3048)
    assert net.ctl_time[1:].tolist() == [6 * 3600, 15 * 3600]
    assert net.ctl_status.tolist() == [ST_CLOSED, ST_OPEN, ST_CLOSED]
    assert lt[idx_link(net, "P3")] == 0          # CV pipe


def idx_link(net, lid):
    return net.link_id.index(lid)


@pytest.mark.parametrize("bad,msg", [
    ("[RESERVOIRS]\n R1 50 PAT\n", "reservoir head patterns"),
    ("[PUMPS]\n PU1 R1 J1 HEAD C1 PATTERN P1\n[CURVES]\n C1 10 40\n", "pump speed patterns"),
    ("[DEMANDS]\n J1 1.0\n J1 2.0 P1\n", "different patterns")])
def test_unsupported_sections_are_refused(tmp_path, bad, msg):
    base = "[JUNCTIONS]\n J1 10 1\n[RESERVOIRS]\n R1 50\n[PIPES]\n P1 R1 J1 100 200 100\n"
    if bad.startswith("[RESERVOIRS]"):
        base = base.replace("[RESERVOIRS]\n R1 50\n", "")
    p = tmp_path / "bad.inp"
    p.write_text(base + bad + "[PATTERNS]\n P1 1.0\n[OPTIONS]\n Units LPS\n")
    with pytest.raises(NotImplementedError, match=msg):
        read_inp(str(p))
