# -*- coding: utf-8 -*-
"""Algorithmic HBM bytes of the GGA stages (FP64, 8 B per value).

Compulsory input/output of each stage when every stage streams its operands
through HBM in the batched layout; static topology is excluded (amortised
over the batch).  The fused kernels keep most of this traffic in LDS and L2,
so achieved = algorithmic bytes / kernel time is a figure of merit against
the 8 TB/s HBM roofline, not a traffic measurement (rocprofv3 FETCH_SIZE /
WRITE_SIZE give the traffic).

Per scenario (L links, Nj junctions, nnzA_lo = lower pattern of A incl.
diagonal, nnzL = entries of L, ne = emitters, r = refinement steps):
  K1 link coefficients   8 * 3L
  K2 assembly            8 * (3L + 2Nj + nnzA_lo)
  K3 factor              8 * (nnzA_lo + nnzL)
  K4 fwd+bwd solve pair  8 * (2 nnzL + 2Nj)            x (1 + r)
  K4b refinement         8 * (nnzA_lo + 3Nj)            x r
  K5 update + relerr     8 * (4L + Nj) + 8 * 4 ne
  adjoint                K1 + RHS 8(L + 2Nj) + K4 x (1 + r) + K4b x r
                         + scatter 8(2L + 2Nj + n_fixed + ne)
"""
from __future__ import annotations


def iteration_bytes(L, Nj, nnzA_lo, nnzL, ne, n_refine=2):
    k1 = 8 * 3 * L
    k2 = 8 * (3 * L + 2 * Nj + nnzA_lo)
    k3 = 8 * (nnzA_lo + nnzL)
    k4 = 8 * (2 * nnzL + 2 * Nj) * (1 + n_refine)
    k4b = 8 * (nnzA_lo + 3 * Nj) * n_refine
    k5 = 8 * (4 * L + Nj) + 8 * 4 * ne
    return k1 + k2 + k3 + k4 + k4b + k5


def factor_solve_bytes(Nj, nnzA_lo, nnzL, n_refine=2):
    """K3 + K4 (+ K4b) of one linear solve."""
    return (8 * (nnzA_lo + nnzL) + 8 * (2 * nnzL + 2 * Nj) * (1 + n_refine)
            + 8 * (nnzA_lo + 3 * Nj) * n_refine)


def adjoint_bytes(L, Nj, nnzA_lo, nnzL, ne, n_fixed, n_refine=2):
    return (8 * 3 * L + 8 * (L + 2 * Nj) + 8 * (2 * nnzL + 2 * Nj) * (1 + n_refine)
            + 8 * (nnzA_lo + 3 * Nj) * n_refine + 8 * (2 * L + 2 * Nj + n_fixed + ne))


def forward_bytes(stats, iters_sum, polish_steps, B, ne, n_refine=2):
    """Algorithmic bytes of one fused forward launch over B scenarios whose
    Newton iteration counts sum to iters_sum; each polish step and the
    terminal factorisation count as one more iteration."""
    per = iteration_bytes(stats["L"], stats["Nj"], stats["nnz_lower"], stats["nnzL"], ne,
                          n_refine)
    return per * (iters_sum + B * (polish_steps + 1))
