# -*- coding: utf-8 -*-
"""Forward + implicit-adjoint throughput of the batched GGA on one MI355X.

A step = one forward solve (fused Newton loop, polish, terminal factor) plus
the implicit adjoint of the calibration loss for B = 1024 scenarios of the
782-junction planar network (synthetic, seeded).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint scenarios/s @B=1024 vs plain per-scenario GGA; grad rel err ≤1e-6"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--B", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--stages", action="store_true",
                    help="also time the standalone factor (k_factor) and triangular-solve "
                         "(k_trisolve) kernels at the same B against their algorithmic bytes")
    ap.add_argument("--workload", choices=("calibration", "ragged", "eps", "leak"),
                    default="calibration",
                    help="calibration: B scenarios of the 782-junction network (the headline "
                         "metric); ragged: the 64-network ragged batch in one launch per "
                         "kernel mode; eps: the differentiable 24-step EPS of the "
                         "20000-junction network at B=256 with checkpointed adjoint; leak: "
                         "the Adam leak-inversion training loop (60 candidate junctions, "
                         "3 true leaks, 33 sensors) on B scenarios of the 782-junction network")
    ap.add_argument("--checkpoint-every", type=int, default=4)
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run one plain 512 MiB read + 512 MiB write elementwise "
                         "kernel, the reference point that scales rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: scenarios are independent and this build runs "
              "on a single MI355X (--gpus 1); multi-GPU runs are not part of it.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "ragged":
        return bench_ragged(args)
    if args.workload == "eps":
        return bench_eps(args)
    if args.workload == "leak":
        return bench_leak(args)
    # the CPU baseline runs first: its worker processes are forked before
    # this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_sample, args.B)
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    from wdsgpu_amd.roofline import adjoint_bytes, forward_bytes
    from wdsgpu_amd.workloads import Calibration

    dev = torch.device("cuda:0")
    B = args.B
    log("building workload B=%d" % B)
    # status checks run one step behind (and once more after the timed
    # steps), so consecutive steps queue without a host round trip
    wl = Calibration(B=B, device=dev, lazy_checks=True)
    s = wl.solver
    st = s.stats
    ne = int(np.count_nonzero(wl.net.node_ke > 0))
    for _ in range(args.warmup):
        wl.step()
    torch.cuda.synchronize()
    wl.check()
    s.event_log = []
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, gd, grh, gC, gCe = wl.step()
    torch.cuda.synchronize()
    wl.check()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    # kernel durations from HIP events recorded on the launching stream
    fw = [(e0.elapsed_time(e1), b, it) for (n, e0, e1, b, it) in s.event_log if n == "k_solve"]
    bw = [(e0.elapsed_time(e1), b) for (n, e0, e1, b, it) in s.event_log if n == "k_adjoint"]
    s.event_log = None
    fw_ms = float(np.mean([x[0] for x in fw]))
    bw_ms = float(np.mean([x[0] for x in bw]))
    iters_sum = float(np.mean([int(x[2].sum()) for x in fw]))
    fbytes = forward_bytes(st, iters_sum, wl.polish_steps, B, ne)
    abytes = adjoint_bytes(st["L"], st["Nj"], st["nnz_lower"], st["nnzL"], ne,
                           st["N"] - st["Nj"]) * B
    achieved = fbytes / (fw_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_solve", "kernel_ms": round(fw_ms, 4),
                "algorithmic_bytes": int(fbytes),
                "adjoint_kernel_ms": round(bw_ms, 4),
                "adjoint_achieved_GBps": round(abytes / (bw_ms * 1e-3) / 1e9, 1),
                "mean_newton_iters": round(iters_sum / B, 3)}
    out = {"metric": METRIC, "value": round(value, 1), "unit": "scenarios/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded planar network, lognormal demand scenarios)",
           "config": {"workload": "planar782 calibration: forward+implicit adjoint, "
                                  "30 sensors, grads wrt demands/C/reservoir heads/leak C",
                      "B": B, "Nj": st["Nj"], "L": st["L"], "nnzL": st["nnzL"],
                      "prvs": st["npcv"], "emitters": ne, "accuracy": wl.accuracy,
                      "polish_steps": wl.polish_steps},
           "roofline": roofline}
    traffic = measured_traffic()
    if traffic is not None:
        roofline["traffic"] = traffic["bytes_per_launch"]
        roofline["traffic_source"] = traffic["source"]
    if args.copy_calibrate:
        n = 512 * 2 ** 20 // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if args.stages:
        out["stages"] = stage_rooflines(s, B)
    # ---- plain per-scenario GPU GGA with finite-difference gradients ----
    if not args.no_plain_baseline:
        out["plain_fd_baseline"] = plain_fd(wl)
        if out["plain_fd_baseline"]["value"] > 0:
            out["speedup_vs_plain_fd"] = round(value / out["plain_fd_baseline"]["value"], 1)
    # ---- CPU baseline: the oracle (reference CPU path restated) ----
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_solve_planar782_B1024_traffic.md")


def measured_traffic():
    """HBM bytes per k_solve launch of this workload from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary (counter KiB scaled by what the
    same counters read for a copy kernel of known size), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    import re
    txt = open(TRAFFIC_FILE).read()
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    if not m:
        return None
    return {"bytes_per_launch": int(float(m.group(1))),
            "source": os.path.relpath(TRAFFIC_FILE, ROOT)}


def stage_rooflines(s, B, reps=20):
    """The factor and solve phases as standalone kernels (wds_factor,
    wds_trisolve with 2 refinements) on B symmetric positive definite
    matrices in the network's CSR pattern (random link weights, diagonally
    dominant): their programs do not depend on the values.  HIP events on
    the launching stream around each launch; algorithmic bytes per scenario
    K3 = 8 (nnzA_lo + nnzL) and K4 + K4b = 8 (2 nnzL + 2 Nj) (1 + r) +
    8 (nnzA_lo + 3 Nj) r (wdsgpu_amd/roofline.py)."""
    st = s.stats
    indptr = s.plan_export(5)
    indices = s.plan_export(6)
    Nj, nnz = st["Nj"], st["nnz_csr"]
    rows = np.repeat(np.arange(Nj), np.diff(indptr))
    g = np.random.default_rng(3)
    w = g.uniform(0.5, 2.0, (B, nnz))
    upper = indices > rows
    # symmetric weights: the value of (i, j) and (j, i) from the upper entry
    key = np.where(upper, rows * Nj + indices, indices * Nj + rows)
    _, inv = np.unique(key, return_inverse=True)
    wu = np.zeros((B, inv.max() + 1))
    wu[:, inv[upper]] = w[:, upper]
    vals = -wu[:, inv]
    diag = rows == indices
    off = np.where(diag, 0.0, vals)
    rowsum = np.zeros((B, Nj))
    np.add.at(rowsum, (slice(None), rows), -off)
    vals = np.where(diag, rowsum[:, rows] + 1.0, off)
    dev = s.device
    csr = torch.tensor(vals, device=dev)
    F = torch.tensor(g.normal(size=(B, Nj)), device=dev)
    fac = s.factor(csr)
    s.trisolve(fac, csr, F)
    torch.cuda.synchronize()
    # the kernels alone: preallocated outputs and workspace, HIP events on
    # the launching stream right around each C-ABI call (the Python methods
    # also allocate and read the pivot status back)
    from wdsgpu_amd import _lib
    lib, plan = s._lib, s._plan
    info = torch.zeros(B, dtype=torch.int32, device=dev)
    H = torch.empty_like(F)
    ws = s._workspace(B)
    strm = torch.cuda.current_stream(dev)
    sp = s._stream()
    calls = (
        ("k_factor", lambda: lib.wds_factor(plan, sp, B, _lib.ptr(csr), _lib.ptr(fac),
                                            _lib.ptr(info), _lib.ptr(ws)),
         8 * (st["nnz_lower"] + st["nnzL"])),
        ("k_trisolve", lambda: lib.wds_trisolve(plan, sp, B, _lib.ptr(fac), _lib.ptr(csr),
                                                _lib.ptr(F), _lib.ptr(H), 2, _lib.ptr(ws)),
         8 * (2 * st["nnzL"] + 2 * Nj) * 3 + 8 * (st["nnz_lower"] + 3 * Nj) * 2))
    out = {}
    for name, fn, nbytes in calls:
        evs = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(strm)
            _lib.check(fn(), name)
            e1.record(strm)
            evs.append((e0, e1))
        torch.cuda.synchronize()
        ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))
        gbs = nbytes * B / (ms * 1e-3) / 1e9
        out[name] = {"kernel_ms": round(ms, 4), "algorithmic_bytes": int(nbytes * B),
                     "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def plain_fd(wl, n_coords=80):
    """Plain per-scenario GPU GGA: the unfused kernels (one launch per Newton
    iteration, host convergence check every iteration) at B=1, with central
    finite-difference gradients: 2P+1 forward solves per gradient scenario,
    P = 782 demands + 905 C + 2 reservoir heads + 10 leak C.  80 coordinates
    are timed for real; the per-scenario time is extrapolated linearly to P."""
    s = wl.solver
    net = wl.net
    P = wl.solver.Nj + int((np.asarray(net.link_type) == 1).sum()) + \
        len(s.fixed_nodes) + int(np.count_nonzero(net.node_ke > 0))
    d = wl.d0[0].clone()
    rh = wl.rh0.clone()
    ke = wl.ke_of(wl.Ce0)
    r = wl.r_of(wl.C0)
    g = np.random.default_rng(5)
    coords = g.choice(s.junc_nodes, n_coords, replace=False)

    def fwd(dd):
        o = s.solve_unfused(dd, rh, ke_int=ke, accuracy=wl.accuracy, status_machine=True)
        return o["head_ft"]
    fwd(d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in coords:
        h = 1e-6 * float(d[j])
        dp = d.clone()
        dp[j] += h
        dm = d.clone()
        dm[j] -= h
        fwd(dp)
        fwd(dm)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    per_solve = t / (2 * n_coords)
    per_grad_scenario = per_solve * (2 * P + 1)
    return {"value": round(1.0 / per_grad_scenario, 4), "unit": "gradient-scenarios/s",
            "forward_ms_B1": round(per_solve * 1e3, 3), "P": int(P),
            "coords_timed": int(n_coords), "extrapolated_to_P": True}


def _cpu_inputs(B):
    """The calibration workload's inputs, built on the CPU: the same seeded
    network, demand scenarios, resistances (r = 4.727 L / C^1.852 /
    D^4.871), leak coefficients and sensors.  The cotangent at the sensors is
    2 (H - H_true) in the workload; its values do not change the work, so
    the baseline uses 2 at every sensor."""
    from wdsgpu_amd.generators import planar_network, scenario_batch, sensor_nodes
    from wdsgpu_amd.network import emitter_ke
    net = planar_network(1)
    d, rh = scenario_batch(net, B, seed=4)
    rh = np.broadcast_to(np.nan_to_num(net.reservoir_head_ft()), (B, net.N)).copy()
    ke = np.broadcast_to(emitter_ke(net.meta.get("emitter_c", np.zeros(net.N))),
                         (B, net.N)).copy()
    gH = np.zeros((B, net.N))
    gH[:, sensor_nodes(net, 30, 5)] = 2.0
    return net, d, rh, ke, gH


def _cpu_chunk(job):
    """Forward + polish + full-Jacobian adjoint of scenarios [a, b), one
    thread (a worker of the all-cores leg)."""
    a, b = job
    from threadpoolctl import threadpool_limits
    from oracle import adjoint as adj
    from oracle import gga
    net, d, rh, ke, gH = _CPU
    m = gga.GGAModel(net)
    with threadpool_limits(1):
        sol = adj.solve_polished(m, d[a:b], rh[a:b], ke[a:b], m.r_hw, accuracy=1e-6,
                                 polish_steps=3)
        adj.cpu_adjoint(m, sol, gH[a:b], np.zeros((b - a, net.L)), np.zeros((b - a, net.N)))
    return b - a


_CPU = None


def cpu_baseline(n_single, B):
    """The oracle (float64 numpy/scipy restatement of the reference's CPU
    path: dense GGA + SuperLU polish + full-Jacobian adjoint) on the same
    workload: (i) one thread on the first n_single scenarios, the
    reference's fair per-scenario unit; (ii) all B scenarios split over the
    host's cores (the box's CPU share: OMP_NUM_THREADS, at most 16), one
    single-threaded worker process per core.  Runs before this process
    touches the GPU, so the workers are plain forks of a CPU-only process."""
    import multiprocessing as mp
    global _CPU
    _CPU = _cpu_inputs(B)
    t0 = time.perf_counter()
    _cpu_chunk((0, n_single))
    t1 = time.perf_counter() - t0
    cores = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
                       or (os.cpu_count() or 1)))
    step = (B + cores - 1) // cores
    jobs = [(a, min(B, a + step)) for a in range(0, B, step)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        done = sum(pool.map(_cpu_chunk, jobs))
    t2 = time.perf_counter() - t0
    return {"value": round(done / t2, 3), "unit": "scenarios/s", "cores": cores,
            "kind": "port",
            "sample": "all %d scenarios of the B=%d workload (forward + polish + adjoint) "
                      "over %d single-threaded worker processes" % (done, B, cores),
            "seconds": round(t2, 2),
            "single_thread": {"value": round(n_single / t1, 3), "unit": "scenarios/s",
                              "cores": 1, "sample": "%d scenarios" % n_single,
                              "seconds": round(t1, 2)},
            "host_cpus": os.cpu_count()}


RAGGED_METRIC = "fwd+adjoint solves/s, ragged 64-network batch in one launch per kernel mode"


def bench_leak(args):
    """The Adam leak-inversion training loop (the reference's training-loop
    benchmark, data/mainline_evidence.md §5.4): B scenarios of the
    782-junction planar network, emitter coefficients of 60 candidate
    junctions as softplus parameters, 3 true leaks of C = 1.5, 33 sensors,
    loss = head MSE + 1e-4 sum C, Adam(lr = 0.1).  A step is one forward +
    implicit adjoint + Adam update; value = B / seconds per step.  The warmup
    steps and the timed steps are consecutive steps of one optimisation."""
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    from wdsgpu_amd.workloads import LeakInversion
    dev = torch.device("cuda:0")
    B = args.B
    log("building the leak-inversion workload B=%d" % B)
    wl = LeakInversion(B=B, device=dev)
    s = wl.solver
    losses = []
    for _ in range(args.warmup):
        losses.append(wl.step())
    torch.cuda.synchronize()
    s.event_log = []
    log("timing %d Adam steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        losses.append(wl.step())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ev = s.event_log
    s.event_log = None
    fw = [a.elapsed_time(b) for (n, a, b, _, _) in ev if n == "k_solve"]
    bw = [a.elapsed_time(b) for (n, a, b, _, _) in ev if n == "k_adjoint"]
    ms = (t1 - t0) * 1e3 / args.steps
    L = [float(x) for x in losses]
    rank = wl.ranking()
    out = {"metric": "Adam leak-inversion training steps (forward + implicit adjoint + "
                     "update), scenarios/s",
           "value": round(B * args.steps / (t1 - t0), 1), "unit": "scenarios/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded planar network, lognormal demand scenarios, seeded "
                   "true leaks)",
           "config": {"workload": "leak inversion: Adam(lr=0.1) on softplus emitter C of 60 "
                                  "candidate junctions, 3 true leaks (C=1.5), 33 sensors, "
                                  "loss = head MSE + 1e-4 sum C",
                      "B": B, "Nj": s.Nj, "L": s.L, "accuracy": wl.accuracy,
                      "polish_steps": wl.polish_steps},
           "kernel_ms": {"k_solve": round(float(np.mean(fw)), 4),
                         "k_adjoint": round(float(np.mean(bw)), 4)},
           "loss": {"first": L[0], "last": L[-1], "steps_taken": len(L)},
           "true_leaks": [int(x) for x in wl.true_leaks],
           "top3_candidates": [int(x) for x in rank[:3]]}
    print(json.dumps(out), flush=True)


def bench_ragged(args):
    """The ragged batch (64 seeded networks, 50-5000 junctions, 4114 solves):
    a step is the forward (accuracy 1e-6, polish 3) plus the implicit adjoint
    of every solve through the multi-network entry points, LDS budget picked
    by the autotuner.  value = solves per second over the whole batch."""
    from wdsgpu_amd.generators import ragged_batch
    batch = ragged_batch(0)
    cpu = None if args.no_cpu_baseline else cpu_baseline_ragged(batch)
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    from wdsgpu_amd.roofline import adjoint_bytes, forward_bytes
    from wdsgpu_amd.workloads import Ragged
    dev = torch.device("cuda:0")
    log("building the ragged workload and autotuning")
    wl = Ragged(device=dev, batch=batch)
    ms = wl.multi
    n_lds, n_arena, lds_bytes = ms.info()
    log("autotune (budget: ms) %s, picked %d" % (
        {k: round(v, 2) for k, v in wl.tuning.items()}, ms.lds_budget))
    for _ in range(args.warmup):
        wl.step()
    torch.cuda.synchronize()
    wl.check()
    ms.event_log = []
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs, grads = wl.step()
    torch.cuda.synchronize()
    wl.check()
    t1 = time.perf_counter()
    n = wl.n_solves
    value = n * args.steps / (t1 - t0)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    ev = ms.event_log
    ms.event_log = None
    fw_ms = float(np.mean([a.elapsed_time(b) for (nm, a, b) in ev if nm == "wds_multi_solve"]))
    bw_ms = float(np.mean([a.elapsed_time(b) for (nm, a, b) in ev
                           if nm == "wds_multi_adjoint"]))
    fbytes = abytes = 0
    iters_total = 0
    for s, o, (net, B) in zip(wl.solvers, outs, wl.batch):
        st = s.stats
        ne = int(np.count_nonzero(net.node_ke > 0))
        it = int(o["iters"].sum().item())
        iters_total += it
        fbytes += forward_bytes(st, it, wl.polish_steps, B, ne)
        abytes += adjoint_bytes(st["L"], st["Nj"], st["nnz_lower"], st["nnzL"], ne,
                                st["N"] - st["Nj"]) * B
    achieved = fbytes / (fw_ms * 1e-3) / 1e9
    kern = "k_solve_multi_arena" if n_lds == 0 else "k_solve_multi + k_solve_multi_arena"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kern, "kernel_ms": round(fw_ms, 4), "algorithmic_bytes": int(fbytes),
                "adjoint_kernel_ms": round(bw_ms, 4),
                "adjoint_achieved_GBps": round(abytes / (bw_ms * 1e-3) / 1e9, 1),
                "mean_newton_iters": round(iters_total / n, 3)}
    # the same work as one forward and one adjoint call per network
    for _ in range(2):
        wl.step(per_network=True)
    torch.cuda.synchronize()
    wl.check()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        wl.step(per_network=True)
    torch.cuda.synchronize()
    wl.check()
    per_net = n * args.steps / (time.perf_counter() - t2)
    nets = [net for net, _ in wl.batch]
    out = {"metric": RAGGED_METRIC, "value": round(value, 1), "unit": "solves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (64 seeded networks, lognormal demand scenarios)",
           "config": {"workload": "ragged batch: 64 networks, forward (polish 3) + implicit "
                                  "adjoint of sensor-head functionals",
                      "networks": len(nets), "solves": n,
                      "Nj_range": [min(x.Nj for x in nets), max(x.Nj for x in nets)],
                      "D-W_networks": sum(x.meta["headloss"] == "D-W" for x in nets),
                      "pumps": sum(x.meta["pumps"] for x in nets),
                      "prvs": sum(x.meta["prvs"] for x in nets),
                      "tcvs": sum(x.meta["tcvs"] for x in nets),
                      "accuracy": wl.accuracy, "polish_steps": wl.polish_steps,
                      "lds_budget": ms.lds_budget, "lds_solves": n_lds,
                      "arena_solves": n_arena,
                      "autotune_ms": {str(k): round(v, 2) for k, v in wl.tuning.items()}},
           "roofline": roofline,
           "per_network_calls": {"value": round(per_net, 1), "unit": "solves/s",
                                 "launches_per_step": 2 * len(nets)}}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


EPS_METRIC = "differentiable EPS fwd+adjoint scenario-steps/s (24 h, checkpointed adjoint)"


def bench_eps(args):
    """The extended-period calibration workload: the seeded 20000-junction
    geometric network (4 reservoirs, 6 tanks, 8 pumps with level controls),
    B = 256 scenarios (lognormal demand factors) over 24 hourly steps; loss
    = squared misfit of 30 sensor heads and the 6 tank levels at every step
    against a run with perturbed roughness (C + N(0, 5)); gradients with
    respect to every demand, pipe C, reservoir head, initial tank level and
    leak C.  A step of the bench is one forward + backward of the whole
    24-hour simulation; value = B * 24 / seconds per step."""
    from wdsgpu_amd.workloads import EpsCalibration
    B = 256 if args.B == 1024 else args.B
    cpu = None if args.no_cpu_baseline else cpu_baseline_eps()
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    from wdsgpu_amd.roofline import adjoint_bytes, forward_bytes
    dev = torch.device("cuda:0")
    log("building the EPS workload B=%d" % B)
    wl = EpsCalibration(B=B, device=dev, checkpoint_every=args.checkpoint_every)
    m, s = wl.model, wl.model.s
    st = s.stats
    for _ in range(args.warmup):
        wl.step()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    base = torch.cuda.memory_allocated(dev)
    s.event_log = []
    solves0 = m.solves
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, grads = wl.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    peak = torch.cuda.max_memory_allocated(dev)
    ev = s.event_log
    s.event_log = None
    fw = [(a.elapsed_time(b), it) for (n, a, b, _, it) in ev if n == "k_solve"]
    bw = [a.elapsed_time(b) for (n, a, b, _, _) in ev if n == "k_adjoint"]
    fw_ms = float(np.mean([x[0] for x in fw]))
    bw_ms = float(np.mean(bw))
    iters = float(np.mean([int(x[1].sum()) for x in fw]))
    ne = int(np.count_nonzero(wl.net.node_ke > 0))
    fbytes = forward_bytes(st, iters, wl.polish_steps, B, ne)
    abytes = adjoint_bytes(st["L"], st["Nj"], st["nnz_lower"], st["nnzL"], ne,
                           st["N"] - st["Nj"]) * B
    achieved = fbytes / (fw_ms * 1e-3) / 1e9
    sec = (t1 - t0) / args.steps
    out = {"metric": EPS_METRIC, "value": round(B * m.T / sec, 1), "unit": "scenario-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded geometric network, lognormal demand scenarios, "
                   "diurnal pattern)",
           "config": {"workload": "EPS calibration: 24 hourly steps, tank-level and "
                                  "sensor-head misfit, checkpointed discrete adjoint",
                      "B": B, "steps_per_simulation": m.T, "Nj": st["Nj"], "L": st["L"],
                      "nnzL": st["nnzL"], "tanks": m.nt, "pumps": len(wl.net.pump_link),
                      "controls": len(wl.net.ctl_link),
                      "checkpoint_every": m.k, "accuracy": m.accuracy,
                      "polish_steps": m.polish_steps,
                      "solves_per_step": (m.solves - solves0) // args.steps,
                      "arena_bytes_per_scenario": st["arena_bytes"],
                      "peak_device_MiB": round(peak / 2 ** 20, 1),
                      "peak_above_inputs_MiB": round((peak - base) / 2 ** 20, 1)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None, "kernel": "k_solve_arena", "kernel_ms": round(fw_ms, 4),
                        "algorithmic_bytes": int(fbytes), "adjoint_kernel_ms": round(bw_ms, 4),
                        "adjoint_achieved_GBps": round(abytes / (bw_ms * 1e-3) / 1e9, 1),
                        "mean_newton_iters": round(iters / B, 3)}}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def cpu_baseline_eps(B=1, T=2):
    """The oracle's fixed-step EPS (sparse SuperLU route) and its discrete
    adjoint on the same 20000-junction network: one scenario, the first T
    steps, one thread; extrapolated per scenario-step.  Runs before this
    process touches the GPU."""
    from threadpoolctl import threadpool_limits
    from oracle import eps_adjoint as ea
    from wdsgpu_amd.generators import geometric_network
    net = geometric_network(0)
    dem = net.demand_cfs[None, :]
    W = np.zeros((T, B, net.N))
    W[:, :, net.tank_node] = 1.0
    t0 = time.perf_counter()
    with threadpool_limits(1):
        fw = ea.forward(net, dem, np.nan_to_num(net.reservoir_head_ft()), net.node_ke,
                        net.r_hw, net.tank_h0[None, :], T, accuracy=1e-6, polish_steps=3)
        ea.adjoint(net, fw, W)
    t = time.perf_counter() - t0
    return {"value": round(B * T / t, 4), "unit": "scenario-steps/s", "cores": 1, "kind": "port",
            "sample": "%d scenario x %d steps of the 20000-junction EPS (forward + polish + "
                      "discrete adjoint, sparse route), one thread" % (B, T),
            "seconds": round(t, 2), "host_cpus": os.cpu_count()}


def _ragged_chunk(job):
    """Forward + polish + adjoint of scenarios [a, b) of ragged network i,
    one thread; returns (i, b - a, seconds)."""
    i, a, b = job
    from threadpoolctl import threadpool_limits
    from oracle import adjoint as adj
    from oracle import gga
    from wdsgpu_amd.generators import scenario_batch, sensor_nodes
    net, B = _CPU[i]
    m = gga.GGAModel(net)
    d, rh = scenario_batch(net, B, seed=100 + i)
    ke = np.broadcast_to(net.node_ke, (B, net.N)).copy()
    W = np.zeros((b - a, net.N))
    W[:, sensor_nodes(net, min(30, max(1, net.Nj // 10)), 1000 + i)] = 1.0
    t0 = time.perf_counter()
    with threadpool_limits(1):
        sol = adj.solve_polished(m, d[a:b], rh[a:b], ke[a:b], m.r_hw, accuracy=1e-6,
                                 polish_steps=3, linsolve="sparse")
        adj.cpu_adjoint(m, sol, W, np.zeros((b - a, net.L)), np.zeros((b - a, net.N)))
    return i, b - a, time.perf_counter() - t0


def cpu_baseline_ragged(batch, per_net=2, chunk=32):
    """The oracle (sparse SuperLU route) on the ragged batch: (i) one thread
    on the first per_net scenarios of every network, extrapolated to the
    batch by each network's scenario count; (ii) the whole batch in chunks of
    at most ``chunk`` scenarios over the host's cores (at most 16), one
    single-threaded worker process each.  Runs before this process touches
    the GPU."""
    import multiprocessing as mp
    global _CPU
    _CPU = batch
    t_single = 0.0
    for i, (net, B) in enumerate(batch):
        _, k, t = _ragged_chunk((i, 0, per_net))
        t_single += t / k * B
    n = sum(B for _, B in batch)
    cores = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
                       or (os.cpu_count() or 1)))
    jobs = [(i, a, min(B, a + chunk)) for i, (_, B) in enumerate(batch)
            for a in range(0, B, chunk)]
    jobs.sort(key=lambda j: -batch[j[0]][0].Nj * (j[2] - j[1]))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        done = sum(r[1] for r in pool.imap_unordered(_ragged_chunk, jobs))
    t2 = time.perf_counter() - t0
    return {"value": round(done / t2, 3), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": "all %d solves of the ragged batch (forward + polish + adjoint, sparse "
                      "route) in chunks of <= %d over %d single-threaded worker processes"
                      % (done, chunk, cores),
            "seconds": round(t2, 2),
            "single_thread": {"value": round(n / t_single, 3), "unit": "solves/s",
                              "cores": 1,
                              "sample": "first %d scenarios of each of the %d networks, "
                                        "extrapolated by scenario count" % (per_net, len(batch))},
            "host_cpus": os.cpu_count()}


if __name__ == "__main__":
    main()
