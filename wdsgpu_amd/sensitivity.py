# -*- coding: utf-8 -*-
"""Sensor-pressure sensitivities to pipe roughness on the HIP path
(dgga/sensitivity.py).

``sensitivity_matrix`` returns S = dp_sensor/dtheta (theta = C or r) for m
sensors over T frames.  The reference factorises each frame's full Jacobian
once with SuperLU and solves the m sensor right-hand sides against it
(dgga/sensitivity.py:59-160).  Here one HIP forward (status machine, polish,
terminal factor) solves every frame, and one HIP adjoint launch solves all
T m right-hand sides, each on its frame's terminal factor: no
factorisation beyond the forward's.  p = H - El, so dp/dtheta = dH/dtheta.

``fisher_information``, ``svd_spectrum`` and ``identifiability_mask`` are the
reference's numpy post-processing of S (dgga/sensitivity.py:166-200).
"""
from __future__ import annotations

from time import perf_counter

import numpy as np
import torch

from . import autodiff
from .calib import _PIPE_TYPES, dr_dC  # noqa: F401

__all__ = ["sensitivity_matrix", "fisher_information", "svd_spectrum", "dr_dC",
           "identifiability_mask"]


def _sensor_index(s, sensor_nodes):
    idx = np.asarray(sensor_nodes, dtype=np.int64).ravel()
    if idx.size == 0:
        raise ValueError("no sensor nodes")
    if idx.min() < 0 or idx.max() >= s.N:
        raise ValueError("sensor node index out of range")
    junc = np.zeros(s.N, dtype=bool)
    junc[s.junc_nodes] = True
    bad = idx[~junc[idx]]
    if bad.size:
        raise ValueError("sensor nodes %s are not junctions (a fixed head is an input, "
                         "dp/dC = 0 there)" % bad.tolist())
    return idx


def sensitivity_matrix(solver, demand, res_head, r_hw, sensor_nodes, wrt="C", frames=None,
                       ke=None, accuracy=1e-12, max_iter=200, polish_steps=3, speed=None,
                       status=None, pump_h0=None, pump_r=None, return_info=False):
    """S [T*m, L] = dp_sensor/dtheta with theta = C (wrt="C", H-W pipes) or
    the resistance r (wrt="r"), frame-major (frame 0's m rows, then frame
    1's, ...), as numpy (dgga/sensitivity.py:59-160).  demand/res_head [N]
    or [B,N]; r_hw [L] (None: the solver's); frames: batch indices (None:
    all).  return_info adds dict(resid_inf, frames, sensors,
    n_factorizations, n_rhs, t_forward, t_adjoint, solution)."""
    s = solver
    if wrt not in ("C", "r"):
        raise ValueError("wrt must be 'C' or 'r'")
    sensors = _sensor_index(s, sensor_nodes)
    m = sensors.size
    if r_hw is not None and np.ndim(r_hw.detach().cpu() if isinstance(r_hw, torch.Tensor)
                                    else r_hw) != 1:
        raise ValueError("r_hw must be [L] (batch-shared), as in the reference")
    t0 = perf_counter()
    st = autodiff.forward_state(s, demand, res_head, ke, r_hw, accuracy, max_iter,
                                polish_steps, speed, pump_h0, pump_r, status)
    torch.cuda.current_stream(s.device).synchronize()
    t_fwd = perf_counter() - t0
    B = st.B
    fr = list(range(B)) if frames is None else [int(b) for b in np.asarray(frames).ravel()]
    if not fr or min(fr) < 0 or max(fr) >= B:
        raise ValueError("frames must index the %d scenarios" % B)
    dev = s.device
    rows = torch.as_tensor(fr, dtype=torch.int64, device=dev).repeat_interleave(m)
    R = rows.numel()
    gH = torch.zeros(R, s.N, dtype=s.dtype, device=dev)
    cols = torch.as_tensor(sensors, dtype=torch.int64, device=dev).repeat(len(fr))
    gH[torch.arange(R, device=dev), cols] = 1.0
    t1 = perf_counter()
    _, _, _, gr = autodiff.adjoint_rows(st, rows, gH=gH)
    S = gr.cpu().numpy()
    t_adj = perf_counter() - t1
    r_used = st.keep[3].detach().cpu().numpy()
    if wrt == "C":
        S = S * dr_dC(s, r_used)[None, :]
    if not return_info:
        return S
    info = dict(resid_inf=st.out["resid"].cpu().numpy(), frames=fr, sensors=sensors,
                n_factorizations=len(fr), n_rhs=int(R), t_forward=t_fwd, t_adjoint=t_adj)
    q = st.out["flow"].cpu().numpy()
    sol = dict(q=q, r_hw=r_used, head=st.out["head"].cpu().numpy(),
               status=st.out["status"].cpu().numpy())
    info["solution"] = sol
    return S, info


def fisher_information(S, sigma):
    """FIM = S^T diag(sigma^-2) S ([L, L]) for a scalar or per-row sigma."""
    S = np.asarray(S, dtype=np.float64)
    sig = np.asarray(sigma, dtype=np.float64)
    if sig.ndim == 0:
        sig = np.full(S.shape[0], float(sig))
    if sig.shape != (S.shape[0],):
        raise ValueError("sigma has shape %s, S has %d rows" % (sig.shape, S.shape[0]))
    if np.any(sig <= 0):
        raise ValueError("sigma must be positive")
    W = S / sig[:, None]
    return W.T @ W


def svd_spectrum(S, rtol=None, full=False):
    """Singular values of S (descending), numerical rank (sv > rtol sv[0],
    rtol default max(shape) eps), effective condition number and cumulative
    energy; with full, the right singular vectors and the null directions."""
    S = np.asarray(S, dtype=np.float64)
    U, sv, Vt = np.linalg.svd(S, full_matrices=False)
    if rtol is None:
        rtol = max(S.shape) * np.finfo(np.float64).eps
    thr = rtol * (sv[0] if sv.size and sv[0] > 0 else 0.0)
    rank = int(np.sum(sv > thr))
    tot = float(np.sum(sv ** 2))
    out = dict(sv=sv, rank=rank, cond=float(sv[0] / sv[rank - 1]) if rank > 0 else np.inf,
               energy=np.cumsum(sv ** 2) / tot if tot > 0 else np.zeros_like(sv))
    if full:
        out.update(Vt=Vt, null_dirs=Vt[rank:], U=U)
    return out


def identifiability_mask(S, atol=0.0):
    """[L] bool: max_i |S[i, k]| > atol."""
    return np.max(np.abs(np.asarray(S, dtype=np.float64)), axis=0) > atol
