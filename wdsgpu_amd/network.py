# -*- coding: utf-8 -*-
"""Water-network data model in EPANET internal units (ft, cfs).

The field names and meanings follow dgga's ``Net`` dataclass
(dgga/parse.py:205-270) so that code written against the reference's network
object reads the same here.  Only the fields the GGA hot path needs are kept.
Arrays are in file order, 0-based.

EPANET constants used by the solver live here too; their values are fixed by
EPANET 2.2 (dgga/solver.py:34-56 quotes them from hydcoeffs.c/types.h) and
therefore must be identical, not merely similar.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------- constants
# EPANET 2.2 values; dgga/solver.py:34-56.
CSMALL = 1.0e-6
CBIG = 1.0e8
BIG = 1.0e10
TINY = 1.0e-6
MISSING = -1.0e10
QZERO = 1.0e-6
RQTOL_DEFAULT = 1.0e-7
HEXP_HW = 1.852
PI_EPANET = 3.141592654          # types.h fallback literal (dgga/solver.py:53)
VISCOS_DEFAULT = 1.1e-5
# Darcy-Weisbach friction-factor constants (dgga/solver.py:40-45)
DW_A1 = 3.14159265358979323850e+03
DW_A2 = 1.57079632679489661930e+03
DW_A8 = 4.61841319859066668690e+00
DW_A9 = -8.68588963806503655300e-01
DW_AB = 3.28895476345399058690e-03
DW_AC = -5.14214965799093883760e-03

# unit factors (dgga/units.py:12-24)
LPS_PER_CFS = 28.317
GPM_PER_CFS = 448.831
M_PER_FT = 0.3048
PSI_PER_FT = 0.4333
FLOW_UCF = {"LPS": LPS_PER_CFS, "GPM": GPM_PER_CFS, "CFS": 1.0,
            "LPM": 1699.0, "MLD": 2.4466, "CMH": 101.94, "CMD": 2446.6,
            "MGD": 0.64632, "IMGD": 0.5382, "AFD": 1.9837}
SI_FLOW_UNITS = ("LPS", "LPM", "MLD", "CMH", "CMD")

# link types (EN_LinkType; dgga/solver.py:72)
CVPIPE, PIPE, PUMP, PRV, PSV, PBV, FCV, TCV, GPV = range(9)
# node types
JUNCTION, RESERVOIR, TANK = 0, 1, 2
# internal status codes (dgga/solver.py:1278-1281)
ST_XHEAD, ST_TEMPCLOSED, ST_CLOSED, ST_OPEN, ST_ACTIVE = 0, 1, 2, 3, 4
ST_XFLOW, ST_XFCV, ST_XPRESSURE = 5, 6, 7
# pump curve types (dgga/parse.py Net.pump_ptype)
PUMP_CONST_HP, PUMP_POWER_FUNC, PUMP_CUSTOM, PUMP_NOCURVE = 0, 1, 2, 3


@dataclass
class Net:
    """Network in EPANET internal units.  Mirrors dgga/parse.py:205-270."""

    node_id: list
    node_type: np.ndarray          # int8[N] 0 junction, 1 reservoir, 2 tank
    elev_ft: np.ndarray            # float64[N] (reservoir: its fixed head)
    node_ke: np.ndarray            # float64[N] emitter coefficient (internal)
    demand_cfs: np.ndarray         # float64[N] base demand (cfs)
    link_id: list
    link_type: np.ndarray          # int8[L]
    link_n1: np.ndarray            # int32[L]
    link_n2: np.ndarray            # int32[L]
    diam_ft: np.ndarray            # float64[L]
    len_ft: np.ndarray             # float64[L]
    roughness: np.ndarray          # float64[L] H-W C, or D-W Kc (ft)
    km_int: np.ndarray             # float64[L] 0.02517*K/D^4
    r_hw: np.ndarray               # float64[L] resistance coefficient
    init_status: np.ndarray        # int8[L] 0 closed, 1 open, 2 active
    valve_setting_user: np.ndarray  # float64[L] valve setting (user units) / pump speed
    meta: dict = field(default_factory=dict)
    # tanks (steady state: fixed-head nodes with level bounds)
    tank_node: np.ndarray = None
    tank_h0: np.ndarray = None
    tank_hmin: np.ndarray = None
    tank_hmax: np.ndarray = None
    tank_area: np.ndarray = None
    tank_overflow: np.ndarray = None
    # pumps (file order of pump links; internal units)
    pump_link: np.ndarray = None
    pump_ptype: np.ndarray = None
    pump_h0: np.ndarray = None
    pump_r: np.ndarray = None
    pump_n: np.ndarray = None
    pump_q0: np.ndarray = None
    pump_qmax: np.ndarray = None
    pump_hmax: np.ndarray = None
    # CUSTOM pump curves: points of pump j in [ptr[j], ptr[j+1]), user units
    # (flow, head) as written in the input file; empty for other pumps
    pump_curve_q: np.ndarray = None
    pump_curve_h: np.ndarray = None
    pump_curve_ptr: np.ndarray = None
    # extended period (dgga/parse.py Net: patterns, controls, tank volumes)
    patterns: list = None          # list of float64 multiplier arrays
    dem_pat: np.ndarray = None     # int32[N] pattern of each node's demand, -1 none
    ctl_link: np.ndarray = None    # simple controls (hydraul.c controls):
    ctl_node: np.ndarray = None    #   controlling tank node or -1
    ctl_type: np.ndarray = None    #   0 LOWLEVEL, 1 HILEVEL, 2 TIMER, 3 TIMEOFDAY
    ctl_grade: np.ndarray = None   #   level controls: head (ft)
    ctl_time: np.ndarray = None    #   time controls: seconds
    ctl_status: np.ndarray = None  #   internal status to set (CLOSED 2 / OPEN 3)

    def __post_init__(self):
        empty = {"tank_node": np.int32, "tank_h0": np.float64,
                 "tank_hmin": np.float64, "tank_hmax": np.float64,
                 "tank_area": np.float64, "tank_overflow": np.int8,
                 "pump_link": np.int32, "pump_ptype": np.int8,
                 "pump_h0": np.float64, "pump_r": np.float64,
                 "pump_n": np.float64, "pump_q0": np.float64,
                 "pump_qmax": np.float64, "pump_hmax": np.float64}
        empty.update({"ctl_link": np.int32, "ctl_node": np.int32, "ctl_type": np.int8,
                      "ctl_grade": np.float64, "ctl_time": np.int64,
                      "ctl_status": np.int8})
        for k, dt in empty.items():
            if getattr(self, k) is None:
                setattr(self, k, np.zeros(0, dtype=dt))
        if self.pump_curve_q is None:
            self.pump_curve_q = np.zeros(0)
            self.pump_curve_h = np.zeros(0)
        if self.pump_curve_ptr is None:
            self.pump_curve_ptr = np.zeros(len(self.pump_link) + 1, dtype=np.int32)
        if self.patterns is None:
            self.patterns = []
        if self.dem_pat is None:
            self.dem_pat = np.full(len(self.node_id), -1, dtype=np.int32)
        defaults = dict(flow_units="LPS", headloss="H-W", trials=200,
                        accuracy=1e-3, demand_model="DDA", qexp=2.0,
                        emitter_exponent=0.5, extra_iter=-1, checkfreq=2,
                        maxcheck=10, damp_limit=0.0, spgrav=1.0,
                        press_units="", htol=0.0005, qtol=0.0001,
                        rqtol=RQTOL_DEFAULT, viscosity=VISCOS_DEFAULT)
        for k, v in defaults.items():
            self.meta.setdefault(k, v)

    @property
    def N(self):
        return len(self.node_id)

    @property
    def L(self):
        return len(self.link_id)

    @property
    def Nj(self):
        return int(np.count_nonzero(np.asarray(self.node_type) == JUNCTION))

    # ---- extended period ----
    def tank_volumes(self):
        """(vmin, vmax, v0) of every tank: cylinders, V = A (h - El)
        (tankvolume, hydraul.c:1038-1068, without volume curves)."""
        el = np.asarray(self.elev_ft)[self.tank_node]
        a = np.asarray(self.tank_area)
        return (a * (self.tank_hmin - el), a * (self.tank_hmax - el),
                a * (self.tank_h0 - el))

    def demand_cfs_at(self, t_sec):
        """Junction demands at time t: base demand times its pattern factor
        for period (t + Pstart) // Pstep (demands, hydraul.c:465-504)."""
        pstep = int(self.meta.get("pat_step_sec", 3600)) or 3600
        p = (int(t_sec) + int(self.meta.get("pat_start_sec", 0))) // pstep
        d = np.asarray(self.demand_cfs, dtype=np.float64).copy()
        for i, pat in enumerate(self.patterns):
            sel = np.asarray(self.dem_pat) == i
            if sel.any():
                d[sel] = d[sel] * float(pat[p % len(pat)])
        return d * float(self.meta.get("demand_multiplier", 1.0))

    def reservoir_head_ft(self):
        """float64[N]: fixed heads at reservoirs/tanks, nan at junctions
        (the res_head_ft convention of dgga/solver.py:2703)."""
        h = np.full(self.N, np.nan)
        nt = np.asarray(self.node_type)
        h[nt == RESERVOIR] = np.asarray(self.elev_ft)[nt == RESERVOIR]
        if self.tank_node.size:
            h[self.tank_node] = self.tank_h0
        return h

    def ucf_pressure(self):
        """Ucf[PRESSURE] (dgga/solver.py:1291-1303)."""
        spg = float(self.meta.get("spgrav", 1.0))
        if str(self.meta["flow_units"]).upper() in SI_FLOW_UNITS:
            press = str(self.meta.get("press_units", "") or "METERS").upper()
            if press == "KPA":
                return 6.895 * PSI_PER_FT * spg
            return M_PER_FT * spg
        return PSI_PER_FT * spg

    def ucf_flow(self):
        return FLOW_UCF[str(self.meta["flow_units"]).upper()]


def hw_resistance(c, len_ft, diam_ft):
    """Hazen-Williams resistance r = 4.727 L / C^1.852 / D^4.871 in EPANET
    internal units; operation order of dgga/calib.py:64-86 (parse.py:603)."""
    return 4.727 * np.asarray(len_ft) / np.asarray(c) ** 1.852 \
        / np.asarray(diam_ft) ** 4.871


def dw_resistance(len_ft, diam_ft):
    """Darcy-Weisbach resistance factor (friction factor applied at solve
    time): r = L / 2 / 32.2 / D / (pi D^2 / 4)^2 (dgga/parse.py:605-606)."""
    d = np.asarray(diam_ft)
    return np.asarray(len_ft) / 2.0 / 32.2 / d / ((PI_EPANET * (d * d) / 4.0) ** 2)


def pump_power_curve(h0, h1, q1, h2, q2, flow_units="LPS", si=True):
    """Fit h = a - b q^c through (0, h0), (q1, h1), (q2, h2) in user units and
    convert to internal units: EPANET's powercurve (input3.c:2101-2129) as
    restated by dgga/parse.py:821-843.  Returns dict(h0, r, n, q0, qmax, hmax)
    with h0 negative (the head gain enters hloss with a minus sign)."""
    if h0 < TINY or h0 - h1 < TINY or h1 - h2 < TINY or q1 < TINY or q2 - q1 < TINY:
        raise ValueError("pump curve fails the powercurve checks")
    a = h0
    h4 = h0 - h1
    h5 = h0 - h2
    c = np.log(h5 / h4) / np.log(q2 / q1)
    if c <= 0.0 or c > 20.0:
        raise ValueError("pump curve exponent %r out of range" % c)
    b = -h4 / q1 ** c
    qcf = FLOW_UCF[flow_units]
    hcf = M_PER_FT if si else 1.0
    return dict(h0=-a / hcf, r=-b * (qcf ** c / hcf), n=c, q0=q1 / qcf,
                qmax=((-a / b) ** (1.0 / c)) / qcf, hmax=h0 / hcf)


def pump_custom_curve(points, flow_units="LPS", si=True):
    """Multi-point (CUSTOM) pump curve, dgga/parse.py:797-820 (EPANET
    input2.c:433-444 and input1.c:644-647): heads must fall strictly from
    point to point (EPANET error 227).  The points stay in user units; the
    solve converts them on every call (curvecoeff, hydcoeffs.c:794-831).
    Returns dict(h0=0, r=0, n=1, q0, qmax, hmax, x, y) in internal units
    except x, y."""
    pts = [(float(q), float(h)) for q, h in points]
    if len(pts) < 2:
        raise ValueError("a CUSTOM pump curve needs at least two points")
    for m in range(1, len(pts)):
        if pts[m][1] >= pts[m - 1][1]:
            raise ValueError("CUSTOM pump curve heads must decrease (EPANET error 227)")
    qcf = FLOW_UCF[flow_units]
    hcf = M_PER_FT if si else 1.0
    qmax = pts[-1][0]
    q0 = (pts[0][0] + qmax) / 2.0
    return dict(h0=0.0, r=0.0, n=1.0, q0=q0 / qcf, qmax=qmax / qcf, hmax=pts[0][1] / hcf,
                x=np.array([p[0] for p in pts]), y=np.array([p[1] for p in pts]))


def emitter_ke(c_user, flow_units="LPS", qexp=2.0, ucf_pressure=M_PER_FT):
    """User emitter coefficient C -> internal Ke = (Ucf_Q^Qexp / Ucf_P) / C^Qexp
    for C>0, else 0 (dgga/parse.py:557-568)."""
    c = np.asarray(c_user, dtype=np.float64)
    qcf = FLOW_UCF[flow_units]
    out = np.zeros_like(c)
    pos = c > 0.0
    out[pos] = (qcf ** qexp / ucf_pressure) / c[pos] ** qexp
    return out


def emitter_ke_dC(c_user, ke, qexp=2.0):
    """dKe/dC = -Qexp * Ke / C (0 where C==0)."""
    c = np.asarray(c_user, dtype=np.float64)
    out = np.zeros_like(c)
    pos = c > 0.0
    out[pos] = -qexp * np.asarray(ke)[pos] / c[pos]
    return out


# ---------------------------------------------------------------- INP reader
def _sections(path):
    sec = None
    out = {}
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for raw in f:
            line = raw.split(";", 1)[0].strip()
            if not line:
                continue
            if line.startswith("["):
                sec = line.upper()
                out.setdefault(sec, [])
                continue
            if sec is not None:
                out[sec].append(line.split())
    return out


def _hours(tok, unit=""):
    """Hours of an INP time field: h[:mm[:ss]] or a number with an optional
    unit word SEC/MIN/HOUR/DAY or AM/PM (EPANET hour(), input3.c)."""
    u = unit.upper()
    if ":" in tok:
        parts = [float(x) for x in tok.split(":")]
        while len(parts) < 3:
            parts.append(0.0)
        h = parts[0] + parts[1] / 60.0 + parts[2] / 3600.0
    else:
        h = float(tok)
        if u.startswith("SEC"):
            return h / 3600.0
        if u.startswith("MIN"):
            return h / 60.0
        if u.startswith("DAY"):
            return h * 24.0
    if u == "AM":
        h = 0.0 if h >= 12.0 and h < 13.0 else h
    elif u == "PM":
        h = h if h >= 12.0 and h < 13.0 else h + 12.0
    return h


def read_inp(path):
    """Read the EPANET INP sections the GGA path and its extended-period
    driver use: [OPTIONS] [TIMES] [PATTERNS] [CURVES] [JUNCTIONS]
    [RESERVOIRS] [TANKS] [PIPES] [PUMPS] [VALVES] [DEMANDS] [EMITTERS]
    [STATUS] [CONTROLS] (simple controls).  Unit conversion follows
    dgga/parse.py:424-1042: lengths and diameters to ft, flows to cfs,
    emitter C to Ke, Hazen-Williams r, pump curves fitted in user units
    (EPANET powercurve for 1- and 3-point curves, CUSTOM otherwise, CONST_HP
    for POWER).  A junction's [DEMANDS] rows replace its [JUNCTIONS] demand
    (input3.c demanddata); junctions without rows keep theirs.  What the
    path cannot represent is refused with NotImplementedError rather than
    dropped: reservoir head patterns, pump speed patterns, tank volume
    curves, several demand categories with different patterns at one node,
    control settings other than OPEN/CLOSED."""
    s = _sections(path)
    opts = {}
    for tk in s.get("[OPTIONS]", []):
        key = tk[0].upper()
        if key in ("UNITS", "HEADLOSS", "PRESSURE") and len(tk) > 1:
            opts[key] = tk[1].upper()
        elif key in ("TRIALS", "ACCURACY", "CHECKFREQ", "MAXCHECK", "VISCOSITY") \
                and len(tk) > 1:
            opts[key] = float(tk[1])
        elif key == "EMITTER" and len(tk) > 2:
            opts["EMITTER"] = float(tk[2])
        elif key == "PATTERN" and len(tk) > 1:
            opts["PATTERN"] = tk[1]
        elif key == "DEMAND" and len(tk) > 2 and tk[1].upper().startswith("MULT"):
            opts["DMULT"] = float(tk[2])
        elif key == "SPECIFIC" and len(tk) > 2:
            opts["SPGRAV"] = float(tk[2])
        elif key == "UNBALANCED" and len(tk) > 1:
            if tk[1].upper().startswith("CONTINUE"):
                opts["EXTRA"] = int(tk[2]) if len(tk) > 2 else 0
            else:
                opts["EXTRA"] = -1
    units = opts.get("UNITS", "GPM")
    si = units in SI_FLOW_UNITS
    qcf = FLOW_UCF[units]
    hcf = M_PER_FT if si else 1.0
    dcf = 1000.0 * M_PER_FT if si else 12.0
    spg = opts.get("SPGRAV", 1.0)
    press = opts.get("PRESSURE", "METERS" if si else "PSI")
    if press == "KPA":
        pcf = 6.895 * PSI_PER_FT * spg
    elif press in ("METERS", "METER"):
        pcf = M_PER_FT * spg
    else:
        pcf = PSI_PER_FT * spg
    headloss = opts.get("HEADLOSS", "H-W")
    # ---- [TIMES] ----
    times = {}
    for tk in s.get("[TIMES]", []):
        key = " ".join(w.upper() for w in tk[:2])
        for name, pre in (("duration_sec", "DURATION"), ("hyd_step_sec", "HYDRAULIC TIMESTEP"),
                          ("pat_step_sec", "PATTERN TIMESTEP"), ("pat_start_sec", "PATTERN START"),
                          ("report_step_sec", "REPORT TIMESTEP"),
                          ("tstart_sec", "START CLOCKTIME")):
            nw = len(pre.split())
            if key.startswith(pre) or (nw == 1 and tk[0].upper() == pre):
                vals = tk[nw:]
                if vals:
                    times[name] = int(round(3600.0 * _hours(vals[0], vals[1] if len(vals) > 1
                                                            else "")))
    # ---- [PATTERNS] / [CURVES] ----
    pat_index, patterns = {}, []
    for tk in s.get("[PATTERNS]", []):
        if tk[0] not in pat_index:
            pat_index[tk[0]] = len(patterns)
            patterns.append([])
        patterns[pat_index[tk[0]]].extend(float(x) for x in tk[1:])
    curves = {}
    for tk in s.get("[CURVES]", []):
        curves.setdefault(tk[0], []).append((float(tk[1]), float(tk[2])))
    default_pat = opts.get("PATTERN", "1")

    def pat_of(name):
        if name is None:
            return pat_index.get(default_pat, -1)
        if name not in pat_index:
            raise ValueError(f"{path}: unknown pattern {name!r}")
        return pat_index[name]

    # ---- nodes ----
    node_id, node_type, elev = [], [], []
    junc_dem = {}
    for tk in s.get("[JUNCTIONS]", []):
        node_id.append(tk[0]); node_type.append(JUNCTION); elev.append(float(tk[1]) / hcf)
        if len(tk) > 2:
            junc_dem[tk[0]] = (float(tk[2]), tk[3] if len(tk) > 3 else None)
    for tk in s.get("[RESERVOIRS]", []):
        if len(tk) > 2:
            raise NotImplementedError(f"{path}: reservoir head patterns are not supported")
        node_id.append(tk[0]); node_type.append(RESERVOIR); elev.append(float(tk[1]) / hcf)
    tanks = []
    for tk in s.get("[TANKS]", []):
        if len(tk) > 7 and tk[7] not in ("*",):
            raise NotImplementedError(f"{path}: tank volume curves are not supported")
        node_id.append(tk[0]); node_type.append(TANK); elev.append(float(tk[1]) / hcf)
        tanks.append(tk)
    idx = {n: i for i, n in enumerate(node_id)}
    N = len(node_id)
    demand = np.zeros(N)
    dem_pat = np.full(N, -1, dtype=np.int32)
    rows = {}
    for tk in s.get("[DEMANDS]", []):
        rows.setdefault(tk[0], []).append((float(tk[1]), tk[2] if len(tk) > 2 else None))
    for nid, (d0, p0) in junc_dem.items():
        if nid not in rows:
            rows[nid] = [(d0, p0)]
    for nid, lst in rows.items():
        i = idx[nid]
        pats = {pat_of(p) for dv, p in lst if dv != 0.0} or {pat_of(lst[0][1])}
        if len(pats) > 1:
            raise NotImplementedError(f"{path}: node {nid} has demand categories with "
                                      "different patterns")
        for dv, _ in lst:
            demand[i] += dv / qcf
        dem_pat[i] = pats.pop()
    qexp = 1.0 / opts.get("EMITTER", 0.5)
    ke = np.zeros(N)
    for tk in s.get("[EMITTERS]", []):
        i = idx[tk[0]]
        c = float(tk[1])
        if node_type[i] == JUNCTION and c > 0:
            ke[i] = (qcf ** qexp / pcf) / c ** qexp
    # ---- links ----
    link_id, lt, n1, n2, dia, ln, rough, km, st, setting = ([] for _ in range(10))
    for tk in s.get("[PIPES]", []):
        link_id.append(tk[0]); n1.append(idx[tk[1]]); n2.append(idx[tk[2]])
        ln.append(float(tk[3]) / hcf); dia.append(float(tk[4]) / dcf)
        rough.append(float(tk[5]) if headloss != "D-W" else float(tk[5]) / (1000.0 * hcf))
        mlk = float(tk[6]) if len(tk) > 6 else 0.0
        km.append(0.02517 * mlk / dia[-1] ** 4)
        status = tk[7].upper() if len(tk) > 7 else "OPEN"
        lt.append(CVPIPE if status == "CV" else PIPE)
        st.append(0 if status == "CLOSED" else 1)
        setting.append(0.0)
    pumps = []
    for tk in s.get("[PUMPS]", []):
        k = len(link_id)
        link_id.append(tk[0]); n1.append(idx[tk[1]]); n2.append(idx[tk[2]])
        dia.append(0.0); ln.append(0.0); rough.append(0.0); km.append(0.0)
        lt.append(PUMP); st.append(1)
        kw = {tk[m].upper()[:4]: tk[m + 1] for m in range(3, len(tk) - 1, 2)}
        if "PATT" in kw:
            raise NotImplementedError(f"{path}: pump speed patterns are not supported")
        setting.append(float(kw.get("SPEE", 1.0)))
        pumps.append((k, kw))
    vt = {"PRV": PRV, "PSV": PSV, "PBV": PBV, "FCV": FCV, "TCV": TCV, "GPV": GPV}
    for tk in s.get("[VALVES]", []):
        link_id.append(tk[0]); n1.append(idx[tk[1]]); n2.append(idx[tk[2]])
        dia.append(float(tk[3]) / dcf); ln.append(0.0); rough.append(0.0)
        lt.append(vt[tk[4].upper()]); setting.append(float(tk[5]))
        mlk = float(tk[6]) if len(tk) > 6 else 0.0
        km.append(0.02517 * mlk / dia[-1] ** 4)
        st.append(2)
    lidx = {n: i for i, n in enumerate(link_id)}
    for tk in s.get("[STATUS]", []):
        k = lidx[tk[0]]
        word = tk[1].upper()
        if word == "CLOSED":
            st[k] = 0
        elif word == "OPEN":
            st[k] = 1
        elif lt[k] == PUMP:
            setting[k] = float(tk[1])
            st[k] = 0 if setting[k] == 0.0 else 1
    ln_a, dia_a = np.asarray(ln), np.asarray(dia)
    rough_a = np.asarray(rough)
    lt_a = np.asarray(lt, dtype=np.int8)
    r = np.zeros(len(link_id))
    pipe = lt_a <= PIPE
    if headloss == "H-W":
        r[pipe] = hw_resistance(rough_a[pipe], ln_a[pipe], dia_a[pipe])
    elif headloss == "D-W":
        r[pipe] = dw_resistance(ln_a[pipe], dia_a[pipe])
    else:
        raise NotImplementedError(f"headloss {headloss}")
    meta = dict(flow_units=units, headloss=headloss, press_units=press,
                trials=int(opts.get("TRIALS", 200)),
                accuracy=float(opts.get("ACCURACY", 1e-3)), qexp=qexp,
                emitter_exponent=1.0 / qexp, spgrav=spg,
                demand_multiplier=float(opts.get("DMULT", 1.0)))
    if "VISCOSITY" in opts:
        meta["viscosity"] = VISCOS_DEFAULT * opts["VISCOSITY"]
    if "EXTRA" in opts:
        meta["extra_iter"] = int(opts["EXTRA"])
    if "CHECKFREQ" in opts:
        meta["checkfreq"] = int(opts["CHECKFREQ"])
    if "MAXCHECK" in opts:
        meta["maxcheck"] = int(opts["MAXCHECK"])
    meta.update(times)
    net = Net(node_id=node_id, node_type=np.asarray(node_type, dtype=np.int8),
              elev_ft=np.asarray(elev), node_ke=ke, demand_cfs=demand,
              link_id=link_id, link_type=lt_a,
              link_n1=np.asarray(n1, dtype=np.int32),
              link_n2=np.asarray(n2, dtype=np.int32), diam_ft=dia_a,
              len_ft=ln_a, roughness=rough_a, km_int=np.asarray(km), r_hw=r,
              init_status=np.asarray(st, dtype=np.int8),
              valve_setting_user=np.asarray(setting, dtype=np.float64), meta=meta,
              patterns=[np.asarray(p, dtype=np.float64) for p in patterns],
              dem_pat=dem_pat)
    if pumps:
        _read_pumps(net, pumps, curves, path, units, si)
    if tanks:
        tn = np.asarray([idx[t[0]] for t in tanks], dtype=np.int32)
        el = net.elev_ft[tn]
        net.tank_node = tn
        net.tank_h0 = el + np.asarray([float(t[2]) for t in tanks]) / hcf
        net.tank_hmin = el + np.asarray([float(t[3]) for t in tanks]) / hcf
        net.tank_hmax = el + np.asarray([float(t[4]) for t in tanks]) / hcf
        dd = np.asarray([float(t[5]) for t in tanks]) / hcf
        net.tank_area = PI_EPANET * dd * dd / 4.0
        net.tank_overflow = np.asarray([1 if len(t) > 8 and t[8].upper().startswith("Y")
                                        else 0 for t in tanks], dtype=np.int8)
    _read_controls(net, s.get("[CONTROLS]", []), idx, lidx, hcf, pcf, path)
    return net


def _read_pumps(net, pumps, curves, path, units, si):
    """Pump curve types and coefficients (dgga/parse.py:740-860; EPANET
    updatepumpparams, input2.c:372-462): POWER -> CONST_HP with
    r = -8.814 P (P in hp; kW / 0.7457 for SI units); a 1-point HEAD curve
    -> power function through (0, 1.33334 h1), (q1, h1), (2 q1, 0); a
    3-point curve starting at zero flow -> power function; any other curve
    -> CUSTOM (points kept in user units)."""
    kinds, h0s, rs, ns, q0s, qmaxs, hmaxs = ([] for _ in range(7))
    cq, ch, cptr = [], [], [0]
    for k, kw in pumps:
        if "POWE" in kw:
            p = float(kw["POWE"])
            if p <= 0.0:
                raise ValueError(f"{path}: pump {net.link_id[k]} power must be positive")
            r = -8.814 * p / (0.7457 if si else 1.0)
            kinds.append(PUMP_CONST_HP); h0s.append(0.0); rs.append(r); ns.append(-1.0)
            q0s.append(1.0); qmaxs.append(BIG); hmaxs.append(BIG)
        elif "HEAD" in kw:
            pts = curves.get(kw["HEAD"])
            if not pts:
                raise ValueError(f"{path}: pump {net.link_id[k]} names unknown curve "
                                 f"{kw['HEAD']!r}")
            if len(pts) == 1 or (len(pts) == 3 and pts[0][0] == 0.0):
                if len(pts) == 1:
                    q1, h1 = pts[0]
                    h0, h2, q2 = 1.33334 * h1, 0.0, 2.0 * q1
                else:
                    h0 = pts[0][1]
                    (q1, h1), (q2, h2) = pts[1], pts[2]
                c = pump_power_curve(h0, h1, q1, h2, q2, flow_units=units, si=si)
                kinds.append(PUMP_POWER_FUNC)
            else:
                c = pump_custom_curve(pts, flow_units=units, si=si)
                kinds.append(PUMP_CUSTOM)
                cq.extend(c["x"].tolist()); ch.extend(c["y"].tolist())
            h0s.append(c["h0"]); rs.append(c["r"]); ns.append(c["n"])
            q0s.append(c["q0"]); qmaxs.append(c["qmax"]); hmaxs.append(c["hmax"])
        else:
            raise ValueError(f"{path}: pump {net.link_id[k]} has neither a curve nor a power")
        cptr.append(len(cq))
    net.pump_link = np.asarray([k for k, _ in pumps], dtype=np.int32)
    net.pump_ptype = np.asarray(kinds, dtype=np.int8)
    net.pump_h0, net.pump_r, net.pump_n = np.asarray(h0s), np.asarray(rs), np.asarray(ns)
    net.pump_q0, net.pump_qmax, net.pump_hmax = np.asarray(q0s), np.asarray(qmaxs), \
        np.asarray(hmaxs)
    net.pump_curve_q, net.pump_curve_h = np.asarray(cq), np.asarray(ch)
    net.pump_curve_ptr = np.asarray(cptr, dtype=np.int32)


def _read_controls(net, rows, idx, lidx, hcf, pcf, path):
    """Simple controls (controldata, input3.c; dgga/parse.py:862-930):
    LINK id OPEN|CLOSED IF NODE id ABOVE|BELOW value, or AT TIME / AT
    CLOCKTIME.  Level thresholds become heads (tank: El + v / Ucf[ELEV];
    junction: El + v / Ucf[PRESSURE])."""
    cl, cn, ct, cg, ctm, cs = ([] for _ in range(6))
    for tk in rows:
        if len(tk) < 6 or tk[0].upper() != "LINK":
            raise NotImplementedError(f"{path}: control {' '.join(tk)!r} is not a simple "
                                      "LINK control")
        k = lidx[tk[1]]
        if net.link_type[k] == CVPIPE:
            raise ValueError(f"{path}: a control cannot act on a check valve")
        word = tk[2].upper()
        if word.startswith("OPEN"):
            status = ST_OPEN
        elif word.startswith("CLOSED"):
            status = ST_CLOSED
        else:
            raise NotImplementedError(f"{path}: control settings other than OPEN/CLOSED")
        u4 = tk[4].upper()
        node, grade, tsec = -1, 0.0, 0
        if u4.startswith("TIME") or u4.startswith("CLOCKTIME"):
            ctype = 2 if u4.startswith("TIME") else 3
            tsec = int(3600.0 * _hours(tk[5], tk[6] if len(tk) > 6 else ""))
            if ctype == 3:
                tsec %= 86400
        else:
            node = idx[tk[5]]
            u6 = tk[6].upper()
            if u6.startswith("BELOW"):
                ctype = 0
            elif u6.startswith("ABOVE"):
                ctype = 1
            else:
                raise ValueError(f"{path}: control condition {tk[6]!r}")
            v = float(tk[7])
            el = float(net.elev_ft[node])
            grade = el + v / (hcf if net.node_type[node] != JUNCTION else pcf)
        cl.append(k); cn.append(node); ct.append(ctype); cg.append(grade)
        ctm.append(tsec); cs.append(status)
    net.ctl_link = np.asarray(cl, dtype=np.int32)
    net.ctl_node = np.asarray(cn, dtype=np.int32)
    net.ctl_type = np.asarray(ct, dtype=np.int8)
    net.ctl_grade = np.asarray(cg, dtype=np.float64)
    net.ctl_time = np.asarray(ctm, dtype=np.int64)
    net.ctl_status = np.asarray(cs, dtype=np.int8)


def _g(x):
    """Shortest text that reads back as the same double."""
    return repr(float(x))


def _time_field(sec):
    return "%d SEC" % int(sec)


def write_inp(net, path):
    """Write a Net as an EPANET INP file that ``read_inp`` reads back to the
    same network (the INP writer half of dgga/parse.py:112-1042; section and
    keyword names are EPANET's).  Internal units go back to the file's flow
    units (meta flow_units) and pressure units (meta press_units); values are
    written with 17 significant digits.  Power-function pump curves are
    written as 3-point curves through (0, shutoff), the design flow and twice
    it, which the reader fits back to the same coefficients; CUSTOM curves
    keep their points; CONST_HP pumps are written as POWER."""
    m = net.meta
    units = str(m.get("flow_units", "LPS")).upper()
    si = units in SI_FLOW_UNITS
    qcf = FLOW_UCF[units]
    hcf = M_PER_FT if si else 1.0
    dcf = 1000.0 * M_PER_FT if si else 12.0
    pcf = net.ucf_pressure()
    headloss = str(m.get("headloss", "H-W")).upper()
    nt = np.asarray(net.node_type)
    lt = np.asarray(net.link_type)
    el = np.asarray(net.elev_ft, dtype=np.float64)
    qexp = float(m.get("qexp", 2.0))
    out = ["[TITLE]", "written by wdsgpu_amd.network.write_inp", ""]
    press = str(m.get("press_units", "") or ("METERS" if si else "PSI")).upper()
    out += ["[OPTIONS]", "UNITS %s" % units, "HEADLOSS %s" % headloss,
            "PRESSURE %s" % press, "TRIALS %d" % int(m.get("trials", 200)),
            "ACCURACY %s" % _g(m.get("accuracy", 1e-3)),
            "EMITTER EXPONENT %s" % _g(1.0 / qexp),
            "SPECIFIC GRAVITY %s" % _g(m.get("spgrav", 1.0)),
            "DEMAND MULTIPLIER %s" % _g(m.get("demand_multiplier", 1.0)),
            "CHECKFREQ %d" % int(m.get("checkfreq", 2)),
            "MAXCHECK %d" % int(m.get("maxcheck", 10)),
            "VISCOSITY %s" % _g(float(m.get("viscosity", VISCOS_DEFAULT)) / VISCOS_DEFAULT)]
    extra = int(m.get("extra_iter", -1))
    out.append("UNBALANCED STOP" if extra < 0 else "UNBALANCED CONTINUE %d" % extra)
    # a pattern name no pattern has: junctions without one stay without
    out += ["PATTERN __none__", ""]
    times = [("DURATION", "duration_sec"), ("HYDRAULIC TIMESTEP", "hyd_step_sec"),
             ("PATTERN TIMESTEP", "pat_step_sec"), ("PATTERN START", "pat_start_sec"),
             ("REPORT TIMESTEP", "report_step_sec"), ("START CLOCKTIME", "tstart_sec")]
    rows = ["%s %s" % (k, _time_field(m[f])) for k, f in times if f in m]
    if rows:
        out += ["[TIMES]"] + rows + [""]
    pname = ["P%d" % (i + 1) for i in range(len(net.patterns))]
    if net.patterns:
        out.append("[PATTERNS]")
        for name, pat in zip(pname, net.patterns):
            out.append(name + " " + " ".join(_g(v) for v in np.asarray(pat)))
        out.append("")
    ids = net.node_id
    dem_pat = np.asarray(net.dem_pat)
    out.append("[JUNCTIONS]")
    for i in np.nonzero(nt == JUNCTION)[0]:
        row = "%s %s %s" % (ids[i], _g(el[i] * hcf), _g(net.demand_cfs[i] * qcf))
        if dem_pat[i] >= 0:
            row += " " + pname[dem_pat[i]]
        out.append(row)
    out += ["", "[RESERVOIRS]"]
    for i in np.nonzero(nt == RESERVOIR)[0]:
        out.append("%s %s" % (ids[i], _g(el[i] * hcf)))
    out.append("")
    if len(net.tank_node):
        out.append("[TANKS]")
        for j, i in enumerate(net.tank_node):
            dd = np.sqrt(4.0 * float(net.tank_area[j]) / PI_EPANET)
            out.append("%s %s %s %s %s %s 0 * %s" % (
                ids[i], _g(el[i] * hcf), _g((net.tank_h0[j] - el[i]) * hcf),
                _g((net.tank_hmin[j] - el[i]) * hcf), _g((net.tank_hmax[j] - el[i]) * hcf),
                _g(dd * hcf), "YES" if net.tank_overflow[j] else "NO"))
        out.append("")
    lids = net.link_id
    n1, n2 = np.asarray(net.link_n1), np.asarray(net.link_n2)
    st = np.asarray(net.init_status)
    dia = np.asarray(net.diam_ft, dtype=np.float64)
    K = np.where(dia > 0, np.asarray(net.km_int) * dia ** 4 / 0.02517, 0.0)
    out.append("[PIPES]")
    for k in np.nonzero(lt <= PIPE)[0]:
        rough = net.roughness[k] if headloss != "D-W" else net.roughness[k] * 1000.0 * hcf
        word = "CV" if lt[k] == CVPIPE else ("CLOSED" if st[k] == 0 else "OPEN")
        out.append("%s %s %s %s %s %s %s %s" % (lids[k], ids[n1[k]], ids[n2[k]],
                                                 _g(net.len_ft[k] * hcf), _g(dia[k] * dcf),
                                                 _g(rough), _g(K[k]), word))
    out.append("")
    curves = []
    status_rows = []
    if len(net.pump_link):
        out.append("[PUMPS]")
        ptr = np.asarray(net.pump_curve_ptr)
        for j, k in enumerate(net.pump_link):
            kind = int(net.pump_ptype[j])
            speed = float(net.valve_setting_user[k])
            if kind == PUMP_CONST_HP:
                p = -float(net.pump_r[j]) * (0.7457 if si else 1.0) / 8.814
                spec = "POWER %s" % _g(p)
            elif kind == PUMP_CUSTOM:
                cname = "C_%s" % lids[k]
                curves += ["%s %s %s" % (cname, _g(x), _g(y)) for x, y in
                           zip(net.pump_curve_q[ptr[j]:ptr[j + 1]],
                               net.pump_curve_h[ptr[j]:ptr[j + 1]])]
                spec = "HEAD %s" % cname
            elif kind == PUMP_POWER_FUNC:
                cname = "C_%s" % lids[k]
                a = -float(net.pump_h0[j]) * hcf
                n = float(net.pump_n[j])
                b = -float(net.pump_r[j]) * hcf / qcf ** n
                q1 = float(net.pump_q0[j]) * qcf
                pts = [(0.0, a), (q1, a + b * q1 ** n), (2.0 * q1, a + b * (2.0 * q1) ** n)]
                curves += ["%s %s %s" % (cname, _g(x), _g(y)) for x, y in pts]
                spec = "HEAD %s" % cname
            else:
                raise NotImplementedError("write_inp: pump %s has no curve to write" % lids[k])
            if speed != 1.0:
                spec += " SPEED %s" % _g(speed)
            out.append("%s %s %s %s" % (lids[k], ids[n1[k]], ids[n2[k]], spec))
            if st[k] == 0:
                status_rows.append("%s CLOSED" % lids[k])
        out.append("")
    vname = {PRV: "PRV", PSV: "PSV", PBV: "PBV", FCV: "FCV", TCV: "TCV", GPV: "GPV"}
    vk = np.nonzero(lt >= PRV)[0]
    if vk.size:
        out.append("[VALVES]")
        for k in vk:
            out.append("%s %s %s %s %s %s %s" % (lids[k], ids[n1[k]], ids[n2[k]],
                                                 _g(dia[k] * dcf), vname[int(lt[k])],
                                                 _g(net.valve_setting_user[k]), _g(K[k])))
            if st[k] == 0:
                status_rows.append("%s CLOSED" % lids[k])
            elif st[k] == 1:
                status_rows.append("%s OPEN" % lids[k])
        out.append("")
    if status_rows:
        out += ["[STATUS]"] + status_rows + [""]
    if curves:
        out += ["[CURVES]"] + curves + [""]
    ke = np.asarray(net.node_ke, dtype=np.float64)
    em = np.nonzero((ke > 0) & (nt == JUNCTION))[0]
    if em.size:
        out.append("[EMITTERS]")
        for i in em:
            out.append("%s %s" % (ids[i], _g(((qcf ** qexp / pcf) / ke[i]) ** (1.0 / qexp))))
        out.append("")
    if len(net.ctl_link):
        out.append("[CONTROLS]")
        for c in range(len(net.ctl_link)):
            k = int(net.ctl_link[c])
            word = "OPEN" if int(net.ctl_status[c]) == ST_OPEN else "CLOSED"
            ct = int(net.ctl_type[c])
            if ct >= 2:
                cond = "AT %s %s" % ("TIME" if ct == 2 else "CLOCKTIME",
                                     _time_field(net.ctl_time[c]))
            else:
                i = int(net.ctl_node[c])
                f = hcf if nt[i] != JUNCTION else pcf
                cond = "IF NODE %s %s %s" % (ids[i], "BELOW" if ct == 0 else "ABOVE",
                                            _g((net.ctl_grade[c] - el[i]) * f))
            out.append("LINK %s %s %s" % (lids[k], word, cond))
        out.append("")
    out.append("[END]")
    with open(path, "w", encoding="utf-8") as fh:
        fh.write("\n".join(out) + "\n")
