// Kernel bodies of the batched GGA forward solve (see kernels.hip).
#pragma once
#include "device_common.h"

namespace wds {

constexpr int64_t kMaxLds = kMaxLdsBytes;
// dynamic LDS after the layouts of the multi-network forward (its item word)
constexpr int64_t kItemSlotBytes = 16;

// per-scenario global workspace, in doubles (wp_tot path nodes, np paths)
__host__ __device__ inline int64_t ws_doubles(int32_t L, int32_t Nj, int32_t na,
                                              int32_t wp_tot, int32_t np) {
  int64_t n = 2 * (int64_t)L + na + 12 * (int64_t)Nj + 2 * (int64_t)wp_tot +
              (int64_t)np * np + ((int64_t)L + 7) / 8 + 512;
  return (n + 31) / 32 * 32;
}

struct SolveParams {
  const double* d;
  const double* rh;
  const double* ke;
  int32_t ke_stride;
  const double* r;
  int32_t r_stride;
  const double* q0;
  const double* e0;
  int32_t max_iter;
  double hacc;
  int32_t sm;
  int32_t n_refine;
  int32_t polish_steps;
  double* head;
  double* flow;
  double* emitter;
  int32_t* iters;
  double* relerr;
  int8_t* status;
  double* diag_ratio;
  int32_t* info;
  double* resid;
  double* factor;
  double* ws;
  const int8_t* status0;
  uint8_t* sched_flags;
  int32_t* sched_events;
  int32_t* sched_count;
  int32_t sched_cap;
  int32_t rh_stride;  // N, or 0 for one vector shared by every scenario
};

inline SolveParams make_params(const wds_solve_args* a, int32_t N) {
  SolveParams s;
  s.d = a->demand;
  s.rh = a->res_head;
  s.ke = a->ke;
  s.ke_stride = a->ke_stride;
  s.r = a->r_hw;
  s.r_stride = a->r_stride;
  s.q0 = a->q0;
  s.e0 = a->e0;
  s.max_iter = a->max_iter;
  s.hacc = a->accuracy;
  s.sm = a->status_machine;
  s.n_refine = a->n_refine;
  s.polish_steps = a->polish_steps;
  s.head = a->head;
  s.flow = a->flow;
  s.emitter = a->emitter;
  s.iters = a->iters;
  s.relerr = a->relerr;
  s.status = a->status;
  s.diag_ratio = a->diag_ratio;
  s.info = a->info;
  s.resid = a->resid;
  s.factor = a->factor;
  s.ws = (double*)a->workspace;
  s.status0 = a->status0;
  s.sched_flags = a->sched_flags;
  s.sched_events = a->sched_events;
  s.sched_count = a->sched_count;
  s.sched_cap = a->sched_cap;
  s.rh_stride = a->res_head_shared ? 0 : N;
  return s;
}

struct AdjParams {
  const double* d;
  const double* ke;
  int32_t ke_stride;
  const double* r;
  int32_t r_stride;
  const double* head;
  const double* flow;
  const double* emitter;
  const int8_t* status;
  const double* factor;
  const double* gH;
  const double* gQ;
  const double* gE;
  int32_t n_refine;
  double* gd;
  double* grh;
  double* gke;
  double* gr;
  int32_t* info;
  double* ws;
  double* gpump;  // [B][3][L] or null
};

inline AdjParams make_adj_params(const wds_adjoint_args* a) {
  AdjParams s;
  s.d = a->demand;
  s.ke = a->ke;
  s.ke_stride = a->ke_stride;
  s.r = a->r_hw;
  s.r_stride = a->r_stride;
  s.head = a->head;
  s.flow = a->flow;
  s.emitter = a->emitter;
  s.status = a->status;
  s.factor = a->factor;
  s.gH = a->gH;
  s.gQ = a->gQ;
  s.gE = a->gE;
  s.n_refine = a->n_refine;
  s.gd = a->g_demand;
  s.grh = a->g_res_head;
  s.gke = a->g_ke;
  s.gr = a->g_r;
  s.info = a->info;
  s.ws = (double*)a->workspace;
  s.gpump = a->g_pump;
  return s;
}

// ---------------------------------------------------------------------------
// Per-scenario context: LDS arrays and global workspace slices.
struct Scn {
  // LDS
  double* V;    // [nnzL + 2] factor (V[nnzL] == 0)
  double* Hn;   // [Nj + Nf] heads by slot
  double* R;    // [Nj] work vector
  double* X2;   // [Nj] second vector (solves of the polish / adjoint)
  double* sh;   // [2*kMaxPrv] small shared scratch
  // global workspace
  double* P;    // [L]
  double* Y;    // [L]
  double* A;    // [na] assembled values in the row-block layout of rs_ell
  double* F;    // [Nj]
  double* e;    // [Nj] emitter flow by slot
  double* hle;  // [Nj]
  double* hge;  // [Nj]
  double* kej;  // [Nj] emitter coefficient by slot
  double* dj;   // [Nj] demand by slot
  double* v1;   // [Nj] extra vectors (polish / adjoint)
  double* v2;   // [Nj]
  double* v3;   // [Nj]
  double* v4;   // [Nj]
  double* v5;   // [Nj]
  double* Yp;   // LDS [wp_tot] y_u = L^-1 e_u on the valve-end paths
  double* Zp;   // LDS [wp_tot] z_u = D^-1 y_u
  double* Cm;   // LDS [np*np] e_u^T A-hat^-1 e_v between valve ends
  double* sM;   // LDS [np^2] small dense (capacitance) system, np = 2p
  double* sw;   // LDS [np] its right-hand side / solution
  double* sx0;  // LDS [np] probes e_u^T A-hat^-1 r
  double* sprv; // LDS [p] valve residuals (polish) / valve cotangents (adjoint)
  int32_t* emj; // LDS [kEmPass * 64 + 1] emitter rows by lane (load_inputs)
  bool em_list; // emj holds this scenario's list (set once load_inputs ran)
  int32_t* wpn; // LDS [wp_tot] slots of the valve-end path nodes (wood_paths)
  int8_t* st0;  // [L]
  double* misc; // [512] small dense systems (Woodbury)
  // per-scenario inputs / state
  double* q;    // [L] flows (output array)
  int8_t* S;    // [L] statuses (output array)
  const double* r;  // [L] resistance
};

// doubles of one scenario's workspace slice: the workspace proper, then (arena
// mode) the arena holding the LDS layout
__host__ __device__ inline int64_t ws_stride(const DevPlan& pl) {
  return ws_doubles(pl.L, pl.Nj, pl.na, pl.wp_tot, pl.np) + pl.arena;
}

// the scenario's LDS layout: the workgroup's dynamic LDS, or (arena mode) the
// arena at the end of the scenario's workspace slice
#ifndef WDS_GLOBAL_ARENA
#define WDS_SCENARIO_MEM(ws, b) extern __shared__ double lds[]
#else
#define WDS_SCENARIO_MEM(ws, b)                                               \
  double* const lds = (ws) + (int64_t)(b) * ws_stride(pl) +                  \
                      ws_doubles(pl.L, pl.Nj, pl.na, pl.wp_tot, pl.np)
#endif

__device__ inline Scn make_scn(const DevPlan& pl, double* lds, double* ws_all, int b,
                               double* flow_all, int8_t* status_all, const double* r_all,
                               int r_stride) {
  Scn c;
  const int Nj = pl.Nj, L = pl.L;
  c.V = lds;
  c.Hn = lds + pl.nnzL + 2;  // V[nnzL] = 0 (zero slot), V[nnzL + 1] spare
  c.R = c.Hn + Nj + pl.Nf;
  c.X2 = c.R + Nj;
  c.sh = c.X2 + Nj;
  double* w = ws_all + (int64_t)b * ws_stride(pl);
  c.P = w;
  w += L;
  c.Y = w;
  w += L;
  c.A = w;
  w += pl.na;
  c.F = w;
  w += Nj;
  c.e = w;
  w += Nj;
  c.hle = w;
  w += Nj;
  c.hge = w;
  w += Nj;
  c.kej = w;
  w += Nj;
  c.dj = w;
  w += Nj;
  c.v1 = w;
  w += Nj;
  c.v2 = w;
  w += Nj;
  c.v3 = w;
  w += Nj;
  c.v4 = w;
  w += Nj;
  c.v5 = w;
  w += Nj;
  w += 2 * (int64_t)pl.wp_tot + (int64_t)pl.np * pl.np;  // unused (kept for layout)
  double* sm = lds + lds_sink_index(pl.nnzL, Nj, pl.Nf) + 64;
  c.Yp = sm;
  sm += pl.wp_tot;
  c.Zp = sm;
  sm += pl.wp_tot;
  c.Cm = sm;
  sm += pl.np * pl.np;
  c.sM = sm;
  sm += pl.np * pl.np;
  c.sw = sm;
  sm += pl.np;
  c.sx0 = sm;
  sm += pl.np;
  c.sprv = sm;
  sm += pl.np / 2;
  c.emj = reinterpret_cast<int32_t*>(sm);
  c.em_list = false;
  sm += kEmDoubles;
  c.wpn = reinterpret_cast<int32_t*>(sm);
  c.st0 = reinterpret_cast<int8_t*>(w);
  w += ((int64_t)L + 7) / 8;
  c.misc = w;
  c.q = flow_all + (int64_t)b * L;
  c.S = status_all + (int64_t)b * L;
  c.r = r_all + (int64_t)b * r_stride;
  return c;
}

// ---------------------------------------------------------------------------
// K1: link and emitter coefficients at the current state
// emitter hloss / hgrad of every junction with an emitter, rows batched
__device__ inline void emitter_coeffs(const DevPlan& pl, Scn& c, int w = 0, int nw = 1) {
  const int lane = lane_id();
  if (c.em_list) {
    const int npass = c.emj[kEmPass * 64];
    if (npass <= kEmPass) {
      // the rows with an emitter from load_inputs' list, pass u on
      // wavefront u % nw: one round trip instead of a walk over every row
      for (int u = w; u < npass; u += nw) {
        const int j = c.emj[u * 64 + lane];
        const int jj = max(j, 0);
        const double kv = c.kej[jj], ev = c.e[jj];
        if (j >= 0) {
          double hl, hg;
          emitter_hl(pl, ev, kv, hl, hg);
          c.hle[j] = hl;
          c.hge[j] = hg;
        }
      }
      return;
    }
  }
  for (int j0 = lane + 64 * kRowG8 * w; j0 < pl.Nj; j0 += 64 * kRowG8 * nw) {
    double kv[kRowG8], ev[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = min(j0 + 64 * u, pl.Nj - 1);
      kv[u] = c.kej[j];
      ev[u] = c.e[j];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = j0 + 64 * u;
      if (j < pl.Nj && kv[u] > 0.0) {
        double hl, hg;
        emitter_hl(pl, ev[u], kv[u], hl, hg);
        c.hle[j] = hl;
        c.hge[j] = hg;
      }
    }
  }
}

__device__ inline void k1_coeffs(const DevPlan& pl, Scn& c, int w = 0, int nw = 1) {
  link_coeffs_all(pl, c.q, c.r, c.S, c.P, c.Y, w, nw);
  emitter_coeffs(pl, c, w, nw);
  mem_sync();
}

// K2: gather assembly in the reference's order (dgga/solver.py:2958-3007).
// Writes A values (global and LDS V), F (global) and Xflow (LDS R).
// polish: pcv PRV links carry their prvcoeff P (adjoint/polish matrix,
// dgga/autodiff.py:928-962) instead of being skipped.
// K2 gathers: entries of a row (and off-diagonal positions) are fetched in
// batches of kGather with every load issued before any is consumed, and the
// accumulation is branch-free.  Sums keep the reference's order exactly.

constexpr int kAsE = 4;   // assembly entries per row loaded per round
constexpr int kOffG = 8;  // off-diagonal entries per lane loaded per round

template <bool kF>
__device__ inline void k2_assemble(const DevPlan& pl, Scn& c, int w = 0, int nw = 1) {
  const int lane = lane_id();
  const int nb = (pl.Nj + 63) >> 6;
  const lds_f64* Hn = (const lds_f64*)c.Hn;
  lds_f64* Vl = (lds_f64*)c.V;
  lds_f64* Rl = (lds_f64*)c.R;
  // the row blocks in nw contiguous shares of about equal entry counts
  const int bw0 = block_split(pl.as_blk, nb, pl.as_blk[nb], w, nw);
  const int bw1 = block_split(pl.as_blk, nb, pl.as_blk[nb], w + 1, nw);
  for (int b0 = bw0; b0 < bw1; b0 += 2) {
    int e0[2], K[2], jr[2], dp[2], da[2];
    double kv[2], hg[2], hl[2], ev[2], dj[2], el[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bk = min(b0 + t, bw1 - 1);
      e0[t] = pl.as_blk[bk];
      K[t] = b0 + t < bw1 ? pl.as_blk[bk + 1] - e0[t] : 0;
      jr[t] = bk * 64 + lane;
      const int j = min(jr[t], pl.Nj - 1);
      kv[t] = c.kej[j];
      hg[t] = c.hge[j];
      hl[t] = c.hle[j];
      ev[t] = c.e[j];
      dj[t] = c.dj[j];
      el[t] = pl.el_slot[j];
      dp[t] = pl.diag_pos[j];
      da[t] = pl.diag_ae[j];
    }
    double dg[2] = {0.0, 0.0}, F[2] = {0.0, 0.0}, X[2] = {0.0, 0.0};
    const int Km = max(K[0], K[1]);
    for (int e = 0; e < Km; e += kAsE) {
      int2 en[2][kAsE];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u)
          en[t][u] = pl.as_ell[(e0[t] + min(e + u, max(K[t] - 1, 0))) * 64 + lane];
      __builtin_amdgcn_sched_barrier(0);
      double Pv[2][kAsE], Yv[2][kAsE], qv[2][kAsE], Hv[2][kAsE];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u) Pv[t][u] = c.P[en[t][u].x & 0x7fffffff];
      if (kF) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int u = 0; u < kAsE; ++u) {
            const int k = en[t][u].x & 0x7fffffff;
            Yv[t][u] = c.Y[k];
            qv[t][u] = c.q[k];
          }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int u = 0; u < kAsE; ++u) Hv[t][u] = Hn[max(en[t][u].y, 0)];
        __builtin_amdgcn_sched_group_barrier(0x0020, 6 * kAsE, 0);
        __builtin_amdgcn_sched_group_barrier(0x0100, 2 * kAsE, 0);
      } else {
        __builtin_amdgcn_sched_group_barrier(0x0020, 2 * kAsE, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u) {
          const bool in_row = e + u < K[t];
          const int y = en[t][u].y;
          const double P = Pv[t][u];
          // links with P == 0 (pcv PRVs) take no part (dgga/solver.py:2939-2956)
          const bool use = in_row && y == -1 && P != 0.0;
          dg[t] = use ? dg[t] + P : dg[t];
          if (kF) {
            const bool n1side = en[t][u].x >= 0;
            F[t] = use ? F[t] + (n1side ? Yv[t][u] : -Yv[t][u]) : F[t];
            X[t] = use ? X[t] + (n1side ? -qv[t][u] : qv[t][u]) : X[t];
            // fixed-head terms: links with a fixed n1, then with a fixed n2
            const bool fix = in_row && y >= 0 && P != 0.0;
            F[t] = fix ? F[t] + P * Hv[t][u] : F[t];
          }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (kv[t] > 0.0) {
        dg[t] = dg[t] + 1.0 / hg[t];
        if (kF) {
          F[t] = F[t] + (hl[t] + el[t]) / hg[t];
          X[t] = X[t] - ev[t];
        }
      }
      if (b0 + t < bw1 && jr[t] < pl.Nj) {
        c.A[da[t]] = dg[t];
        Vl[dp[t]] = dg[t];
        if (kF) {
          X[t] = X[t] - dj[t];
          c.F[jr[t]] = F[t] + X[t];
          Rl[jr[t]] = X[t];
        }
      }
    }
  }
  for (int o0 = lane + 64 * kOffG * w; o0 < pl.noff; o0 += 64 * kOffG * nw) {
    int4 w[kOffG];
    int2 wa[kOffG];
#pragma unroll
    for (int u = 0; u < kOffG; ++u) {
      w[u] = pl.off4[min(o0 + 64 * u, pl.noff - 1)];
      wa[u] = pl.off_ae[min(o0 + 64 * u, pl.noff - 1)];
    }
    __builtin_amdgcn_sched_barrier(0);
    double v[kOffG];
#pragma unroll
    for (int u = 0; u < kOffG; ++u) v[u] = c.P[w[u].y];
    __builtin_amdgcn_sched_group_barrier(0x0020, kOffG, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kOffG; ++u) {
      const double P = v[u];
      v[u] = (w[u].w > 0 && P != 0.0) ? -P : 0.0;
      // parallel links between one junction pair (rare): in link order
      for (int e = w[u].z + 1; e < w[u].z + w[u].w; ++e) {
        const double Pe = c.P[pl.off_link[e]];
        if (Pe != 0.0) v[u] += -Pe;
      }
      if (o0 + 64 * u < pl.noff) {
        c.A[wa[u].x] = v[u];
        c.A[wa[u].y] = v[u];
        Vl[w[u].x] = v[u];
      }
    }
  }
  for (int f0 = lane + 64 * kRowG * w; f0 < pl.nfill; f0 += 64 * kRowG * nw) {
    int fp[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) fp[u] = pl.fill_pos[min(f0 + 64 * u, pl.nfill - 1)];
#pragma unroll
    for (int u = 0; u < kRowG; ++u)
      if (f0 + 64 * u < pl.nfill) Vl[fp[u]] = 0.0;
  }
  if (lane == 0 && w == 0) Vl[pl.nnzL] = 0.0;
  mem_sync();
  lds_sync();
}

// prvcoeff for every PRV with a setting, in valve order, by one lane
// (dgga/solver.py:850-912).  Must follow k2_assemble (reads Xflow[j2]).
// the coefficient terms of control valve i (A, V, F, P, Y; the Xflow R of
// an ACTIVE FCV): prvcoeff, psvcoeff, fcvcoeff (dgga/solver.py:850-912 for
// the batched PRV; the scalar valvecoeffs 1616-1688 for PSV and FCV).  A
// valve that is not ACTIVE takes valvecoeff's open (or closed) terms.
__device__ __forceinline__ void prv_apply(const DevPlan& pl, Scn& c, int i) {
  const int k = pl.prv_k[i], j1 = pl.prv_s1[i], j2 = pl.prv_s2[i];
  const int kind = pl.prv_kind[i];
  const double qk = c.q[k];
  const int8_t sk = c.S[k];
  const bool act = sk == ST_ACTIVE, cl = sk <= ST_CLOSED;
  double pk_o, yk_o;
  valve_open(pl, qk, pl.prv_km[i], pk_o, yk_o);
  const double pk_na = cl ? 1.0 / CBIG : pk_o;
  const double yk_na = cl ? qk : yk_o;
  const double mna = act ? 0.0 : 1.0, mact = act ? 1.0 : 0.0;
  const double xf1 = c.R[j1], xf2 = c.R[j2];
  const double hset = pl.prv_hset[i];
  const bool fcv_act = kind == 2 && act;
  // an ACTIVE FCV cuts its link: P = 1/CBIG, flow set to the setting
  const double pna = fcv_act ? 1.0 / CBIG : pk_na * mna;
  const int po = pl.prv_pos_off[i], d1 = pl.diag_pos[j1], d2 = pl.diag_pos[j2];
  const int4 ae = pl.prv_ae[i];
  // V holds the same values as A here (k2_assemble writes both); an ACTIVE
  // PRV pins its downstream row, an ACTIVE PSV its upstream row
  double v;
  v = c.V[po] + (-pna);
  c.A[ae.x] = v;
  c.A[ae.y] = v;
  c.V[po] = v;
  v = c.V[d1] + pna;
  if (kind == 1) v = v + CBIG * mact;
  c.A[ae.z] = v;
  c.V[d1] = v;
  v = c.V[d2] + pna;
  if (kind == 0) v = v + CBIG * mact;
  c.A[ae.w] = v;
  c.V[d2] = v;
  const double dy = (yk_na - qk) * mna;
  double f1, f2, y;
  if (kind == 0) {
    f1 = dy + ((act && xf2 < 0.0) ? xf2 : 0.0);
    f2 = -dy + hset * CBIG * mact;
    y = act ? qk + xf2 : yk_na;
  } else if (kind == 1) {
    f1 = dy + hset * CBIG * mact;
    f2 = -dy + ((act && xf1 > 0.0) ? xf1 : 0.0);
    y = act ? qk - xf1 : yk_na;
  } else {
    f1 = fcv_act ? -hset : (yk_na - qk);
    f2 = fcv_act ? hset : -(yk_na - qk);
    y = fcv_act ? qk - hset : yk_na;
    if (fcv_act) {
      // the setting leaves j1 and enters j2 as fixed flows for the valves
      // that follow (fcvcoeff's Xflow update)
      c.R[j1] = xf1 - hset;
      c.R[j2] = xf2 + hset;
    }
  }
  c.F[j1] = c.F[j1] + f1;
  c.F[j2] = c.F[j2] + f2;
  c.P[k] = pna;
  c.Y[k] = y;
}

__device__ inline void k2_prv(const DevPlan& pl, Scn& c) {
  if (pl.p == 0) return;
  const int lane = lane_id();
  if (pl.prv_disjoint) {
    // no two valves share an end node: their terms touch disjoint entries of
    // A, V and F, so every valve goes on its own lane at once
    if (lane < pl.p) prv_apply(pl, c, lane);
  } else if (lane == 0) {
    // valves in valve order: shared diagonal and right-hand-side entries are
    // summed in the reference's order
    for (int i = 0; i < pl.p; ++i) prv_apply(pl, c, i);
  }
  mem_sync();
  lds_sync();
}

// K3 + K4: factor V in place, then x = A^-1 b with n_refine refinement steps.
// b is read from global (bsrc, slot order), x is left in X (LDS).
// Solve programs on the solve vector S = c.R with the unit factor in V.
// Forward y = L^-1 S then the pivot scaling z = D^-1 y (in place):
__device__ inline void solve_fwd(const DevPlan& pl, Scn& c, i32x4 (&ring)[kSlots]) {
  int dp[kRowG];
  pivots_prefetch(pl, dp);
#ifdef WDS_PROFILE_STAGES
  run_dot(pl.fwd, c.V, ring, c.misc + 408 + 32);
#else
  run_dot(pl.fwd, c.V, ring);
#endif
  // the backward program's first steps load while the pivots scale
  prefetch_prog(pl.bwd, ring);
  scale_by_pivots(pl, c.V, c.R, dp);
  lds_sync();
}

// backward x = L^-T z in place in S; ring holds the backward program's
// first steps (left there by solve_fwd)
__device__ inline void solve_bwd(const DevPlan& pl, Scn& c, i32x4 (&ring)[kSlots]) {
#ifdef WDS_PROFILE_STAGES
  run_dot(pl.bwd, c.V, ring, c.misc + 408 + 64);
#else
  run_dot(pl.bwd, c.V, ring);
#endif
}

// Stage cycle counters for a profiling build (-DWDS_PROFILE_STAGES): lane 0
// accumulates clock64() deltas of each stage into misc[400 + stage].
#ifdef WDS_PROFILE_STAGES
#define WDS_T0() long long wds_t0_ = clock64()
#define WDS_TACC(c, i)                                              \
  do {                                                              \
    const long long t1_ = clock64();                                \
    if (threadIdx.x == 0) (c).misc[400 + (i)] += (double)(t1_ - wds_t0_); \
    wds_t0_ = t1_;                                                  \
  } while (0)
#else
#define WDS_T0() (void)0
#define WDS_TACC(c, i) (void)0
#endif

// X = A^-1 b (b global, slot order; X in LDS, not S) with n_refine steps of
// refinement on the double-double residual of the assembled A.  With nw > 1
// wavefronts the copies and residuals are shared and wavefront 0 runs the
// triangular solves (the others wait at the workgroup barriers).
__device__ inline void k4_solve(const DevPlan& pl, Scn& c, const double* bsrc, double* X,
                                int n_refine, int w = 0, int nw = 1) {
  const int lane = lane_id();
  double* const S = c.R;
  WDS_T0();
  // the forward program's first steps load under the copy (and under each
  // refinement residual)
  i32x4 ring[kSlots];
  if (w == 0) prefetch_prog(pl.fwd, ring);
  vec_copy(S, bsrc, nullptr, pl.Nj, w, nw);
  lds_sync();
  if (nw > 1) wg_sync();
  WDS_TACC(c, 28);
  if (w == 0) {
    solve_fwd(pl, c, ring);
    WDS_TACC(c, 29);
    solve_bwd(pl, c, ring);
    WDS_TACC(c, 30);
    for (int j = lane; j < pl.Nj; j += 64) X[j] = S[j];
    lds_sync();
  }
  for (int it = 0; it < n_refine; ++it) {
    if (nw > 1) wg_sync();
    WDS_TACC(c, 31);
    if (w == 0) prefetch_prog(pl.fwd, ring);
    residual_all(pl, c.A, X, bsrc, S, w, nw);
    lds_sync();
    if (nw > 1) wg_sync();
    WDS_TACC(c, 27);
    if (w == 0) {
      solve_fwd(pl, c, ring);
      WDS_TACC(c, 29);
      solve_bwd(pl, c, ring);
      WDS_TACC(c, 30);
      for (int j = lane; j < pl.Nj; j += 64) X[j] += S[j];
      lds_sync();
    }
  }
  WDS_TACC(c, 31);
}

// Numeric factorisation, then the unit-factor scaling l_ij = w_ij / D_j of
// every off-diagonal entry (the solves and the Woodbury paths use L).
__device__ inline bool k3_factor(const DevPlan& pl, Scn& c) {
  int bad = 0;
#ifdef WDS_PROFILE_STAGES
  run_factor(pl.factor, c.V, bad, c.misc + 408);
#else
  run_factor(pl.factor, c.V, bad);
#endif
  lds_f64* Vl = (lds_f64*)c.V;
  for (int p0 = lane_id(); p0 < pl.nnzL; p0 += 64 * kRowG) {
    int sd[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) sd[u] = pl.l_scale[min(p0 + 64 * u, pl.nnzL - 1)];
#pragma unroll
    for (int u = 0; u < kRowG; ++u)
      if (p0 + 64 * u < pl.nnzL && sd[u] >= 0) Vl[p0 + 64 * u] *= Vl[sd[u]];
  }
  lds_sync();
  return wave_any(bad != 0);
}

// K5: newflows and relerr (dgga/solver.py:3092-3110); returns relerr.
__device__ inline double k5_update(const DevPlan& pl, Scn& c, double hacc) {
  const int lane = lane_id();
  double dqs = 0.0, qs = 0.0, dqe = 0.0, qse = 0.0;
  for (int k0 = lane; k0 < pl.L; k0 += 64 * kGather) {
    int s1[kGather], s2[kGather], fl[kGather];
    double qv[kGather], Yv[kGather], Pv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = min(k0 + 64 * u, pl.L - 1);
      s1[u] = pl.l_s1[k];
      s2[u] = pl.l_s2[k];
      fl[u] = pl.l_flags[k];
      qv[u] = c.q[k];
      Yv[u] = c.Y[k];
      Pv[u] = c.P[k];
    }
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = k0 + 64 * u;
      if (k >= pl.L) break;
      const double dh = c.Hn[s1[u]] - c.Hn[s2[u]];
      const double q = qv[u];
      double dq = Yv[u] - Pv[u] * dh;
      if ((fl[u] & LF_PUMP_CHP) && dq > q) dq = q / 2.0;
      const double qn = q - dq;
      dqs += fabs(dq);
      qs += fabs(qn);
      c.q[k] = qn;
    }
  }
  // emitter rows in original junction order (row r = lane + 64 i), so the
  // reduction tree below is defined on file order: lane partials over links
  // k = lane + 64 i and rows r = lane + 64 i in ascending order, then an xor
  // butterfly (m = 32 .. 1).  The rows with an emitter come from the list
  // load_inputs made (same rows, same order per lane; one memory round
  // trip), unless some lane has more than kEmPass of them.
  const int npass = c.emj[kEmPass * 64];
  if (npass <= kEmPass) {
    int jv[kEmPass];
#pragma unroll
    for (int u = 0; u < kEmPass; ++u) jv[u] = c.emj[u * 64 + lane];
    double hl[kEmPass], hg[kEmPass], ev[kEmPass], hn[kEmPass], el[kEmPass];
#pragma unroll
    for (int u = 0; u < kEmPass; ++u) {
      const int j = max(jv[u], 0);
      hl[u] = c.hle[j];
      hg[u] = c.hge[j];
      ev[u] = c.e[j];
      hn[u] = c.Hn[j];
      el[u] = pl.el_slot[j];
    }
#pragma unroll
    for (int u = 0; u < kEmPass; ++u) {
      if (jv[u] < 0) continue;
      const double dh = hn[u] - el[u];
      const double dq = (hl[u] - dh) / hg[u];
      const double en = ev[u] - dq;
      dqe += fabs(dq);
      qse += fabs(en);
      c.e[jv[u]] = en;
    }
  } else {
    for (int r0 = lane; r0 < pl.Nj; r0 += 64 * kRowG8) {
      int jv[kRowG8];
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) jv[u] = pl.row_slot[min(r0 + 64 * u, pl.Nj - 1)];
      double kv[kRowG8];
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) kv[u] = c.kej[jv[u]];
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) {
        const int j = jv[u];
        if (r0 + 64 * u < pl.Nj && kv[u] > 0.0) {
          const double dh = c.Hn[j] - pl.el_slot[j];
          const double dq = (c.hle[j] - dh) / c.hge[j];
          const double en = c.e[j] - dq;
          dqe += fabs(dq);
          qse += fabs(en);
          c.e[j] = en;
        }
      }
    }
  }
  const double dqsum = wave_sum(dqs) + wave_sum(dqe);
  const double qsum = wave_sum(qs) + wave_sum(qse);
  mem_sync();
  return qsum > hacc ? dqsum / qsum : dqsum;
}

// Status-change log of record_schedule: every lane whose link changed
// status appends (it << 20) | (kind << 19) | link to its scenario's list;
// slots come from a ballot prefix count, nev (uniform) counts all changes.
struct SchedLog {
  int32_t* ev;  // this scenario's events, or null when not recording
  int cap, it, nev;
};

__device__ inline void sched_log(SchedLog& lg, int kind, bool chg, int link) {
  const uint64_t m = __ballot(chg);
  const int slot = lg.nev + (int)__popcll(m & __lanemask_lt());
  if (chg && slot < lg.cap) lg.ev[slot] = (lg.it << 20) | (kind << 19) | link;
  lg.nev += (int)__popcll(m);
}

// K6a: valvestatus every iteration: prvstatus (dgga/solver.py:914-947) for
// every PRV and psvstatus (the scalar replica's 1740-1767, hydstatus.c
// psvstatus) for every PSV with a setting; FCVs change in linkstatus
__device__ inline bool k6_prvstatus(const DevPlan& pl, Scn& c, SchedLog& lg) {
  const int i = lane_id();
  bool chg = false;
  if (i < pl.p && pl.prv_kind[i] != 2) {
    const int k = pl.prv_k[i];
    const double h1 = c.Hn[pl.prv_s1[i]], h2 = c.Hn[pl.prv_s2[i]];
    const double qk = c.q[k];
    const int8_t sk = c.S[k];
    const double hset = pl.prv_hset[i];
    const double hml = pl.prv_km[i] * (qk * qk);
    const double htol = pl.htol;
    const bool neg = qk < -pl.qtol;
    int8_t sn = sk;
    if (pl.prv_kind[i] == 0) {
      if (sk == ST_ACTIVE)
        sn = neg ? ST_CLOSED : (h1 - hml < hset - htol ? ST_OPEN : ST_ACTIVE);
      else if (sk == ST_OPEN)
        sn = neg ? ST_CLOSED : (h2 >= hset + htol ? ST_ACTIVE : ST_OPEN);
      else if (sk == ST_CLOSED)
        sn = (h1 >= hset + htol && h2 < hset - htol)
                 ? ST_ACTIVE
                 : ((h1 < hset - htol && h1 > h2 + htol) ? ST_OPEN : ST_CLOSED);
      else if (sk == ST_XPRESSURE)
        sn = neg ? ST_CLOSED : sk;
    } else {
      if (sk == ST_ACTIVE)
        sn = neg ? ST_CLOSED : (h2 + hml > hset + htol ? ST_OPEN : ST_ACTIVE);
      else if (sk == ST_OPEN)
        sn = neg ? ST_CLOSED : (h1 < hset - htol ? ST_ACTIVE : ST_OPEN);
      else if (sk == ST_CLOSED)
        sn = (h2 > hset + htol && h1 > h2 + htol)
                 ? ST_OPEN
                 : ((h1 >= hset + htol && h1 > h2 + htol) ? ST_ACTIVE : ST_CLOSED);
      else if (sk == ST_XPRESSURE)
        sn = neg ? ST_CLOSED : sk;
    }
    chg = sn != sk;
    c.S[k] = sn;
  }
  if (lg.ev) sched_log(lg, 0, chg, i < pl.p ? pl.prv_k[i] : 0);
  const bool any = wave_any(chg);
  mem_sync();
  return any;
}

__device__ inline int8_t cvstatus(double htol, double qtol, int8_t s, double dh, double q) {
  const bool big = fabs(dh) > htol;
  const bool rev = q < -qtol;
  if (big) return (dh < -htol || rev) ? ST_CLOSED : ST_OPEN;
  return rev ? ST_CLOSED : s;
}
__device__ inline bool cv_closed(double htol, double qtol, bool s_is_open, double dh, double q) {
  const bool big = fabs(dh) > htol;
  if (big) return (dh < -htol) || (q < -qtol);
  return s_is_open ? (q < -qtol) : true;
}

// K6b: linkstatus (dgga/solver.py:697-739); returns "changed"
__device__ inline bool k6_linkstatus(const DevPlan& pl, Scn& c, SchedLog& lg) {
  const int lane = lane_id();
  bool chg = false;
  if (!pl.has_cv && !pl.has_pump && pl.ntl == 0 && !pl.has_fcv) return false;
  for (int k = lane; k < pl.L; k += 64) {
    const int8_t st0 = c.S[k];
    c.st0[k] = st0;
    int8_t s = (st0 == ST_XHEAD || st0 == ST_TEMPCLOSED) ? ST_OPEN : st0;
    const int f = pl.l_flags[k];
    const double dh = c.Hn[pl.l_s1[k]] - c.Hn[pl.l_s2[k]];
    if (f & LF_CV) s = cvstatus(pl.htol, pl.qtol, s, dh, c.q[k]);
    if ((f & LF_PUMP_LIVE) && s >= ST_OPEN)
      s = (-dh > pl.pu_hmax[k] + pl.htol) ? ST_XHEAD : ST_OPEN;
    c.S[k] = s;
  }
  mem_sync();
  for (int t = lane; t < pl.ntl; t += 64) {
    const int k = pl.tl_k[t], bits = pl.tl_bits[t];
    const int8_t s = c.S[k];
    const bool live = s > ST_CLOSED;
    const double ht = c.Hn[pl.tl_tank[t]];
    const double dht = ht - c.Hn[pl.tl_other[t]];
    const double qk = c.q[k] * ((bits & 1) ? -1.0 : 1.0);
    const bool full = (ht >= pl.tl_hmax[t] - pl.htol) && !(bits & 2);
    const bool empty = ht <= pl.tl_hmin[t] + pl.htol;
    const bool is_pump = bits & 4;
    const bool cut_f = is_pump ? (bool)(bits & 8) : cv_closed(pl.htol, pl.qtol, true, dht, qk);
    const bool cut_e = is_pump ? (bool)(bits & 16) : !cv_closed(pl.htol, pl.qtol, false, dht, qk);
    if (live && ((full && cut_f) || (empty && cut_e))) c.S[k] = ST_TEMPCLOSED;
  }
  // FCVs with a setting: fcvstatus from the status before the reopening
  // (the scalar replica's _linkstatus, dgga/solver.py:1875-1877, and
  // _fcvstatus 1819-1831): XFCV on a reversed head or flow, back to ACTIVE
  // from XFCV once the flow reaches the setting
  if (pl.has_fcv && lane < pl.p && pl.prv_kind[lane] == 2) {
    const int i = lane, k = pl.prv_k[i];
    const int8_t st0 = c.st0[k];
    const double h1 = c.Hn[pl.prv_s1[i]], h2 = c.Hn[pl.prv_s2[i]], qk = c.q[k];
    int8_t s = st0;
    if (h1 - h2 < -pl.htol)
      s = ST_XFCV;
    else if (qk < -pl.qtol)
      s = ST_XFCV;
    else if (st0 == ST_XFCV && qk >= pl.prv_hset[i])
      s = ST_ACTIVE;
    c.S[k] = s;
  }
  mem_sync();
  if (lg.ev) {
    for (int k0 = 0; k0 < pl.L; k0 += 64) {
      const int k = k0 + lane;
      const bool ck = k < pl.L && c.S[k] != c.st0[k];
      sched_log(lg, 1, ck, k);
      chg |= ck;
    }
  } else {
    for (int k = lane; k < pl.L; k += 64) chg |= c.S[k] != c.st0[k];
  }
  return wave_any(chg);
}

// diag ratio proxy max|diag| / min|diag| (dgga/solver.py:3026-3032) of the
// matrix last assembled (with the PRV terms), read from its global copy A:
// the loop reports the last iteration's value, so it is taken once after it
__device__ inline double diag_ratio_now(const DevPlan& pl, Scn& c) {
  double mx = 0.0, mn = __builtin_huge_val();
  for (int j0 = lane_id(); j0 < pl.Nj; j0 += 64 * kRowG) {
    int da[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) da[u] = pl.diag_ae[min(j0 + 64 * u, pl.Nj - 1)];
    double v[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) v[u] = c.A[da[u]];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      if (j0 + 64 * u >= pl.Nj) break;
      mx = fmax(mx, fabs(v[u]));
      mn = fmin(mn, fabs(v[u]));
    }
  }
  mx = wave_max(mx);
  mn = wave_min(mn);
  return mx / fmax(mn, 1e-300);
}

// state load shared by the fused and the per-iteration kernels
__device__ inline void load_inputs(const DevPlan& pl, Scn& c, const SolveParams& sp, int b) {
  const int lane = lane_id();
  const double* d = sp.d + (int64_t)b * pl.N;
  const double* ke = sp.ke + (int64_t)b * sp.ke_stride;
  const double* rh = sp.rh + (int64_t)b * sp.rh_stride;
  for (int j0 = lane; j0 < pl.Nj; j0 += 64 * kRowG8) {
    int nd[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) nd[u] = pl.slot_node[min(j0 + 64 * u, pl.Nj - 1)];
    double dv[kRowG8], kv[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      dv[u] = d[nd[u]];
      kv[u] = ke[nd[u]];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = j0 + 64 * u;
      if (j >= pl.Nj) break;
      c.dj[j] = dv[u];
      c.kej[j] = kv[u];
    }
  }
  for (int f = lane; f < pl.Nf; f += 64) {
    const double h = rh[pl.slot_node[pl.Nj + f]];
    c.Hn[pl.Nj + f] = (h != h) ? 0.0 : h;  // NaN reads as 0 (torch.nan_to_num)
  }
  // the emitter rows of each lane (rows r = lane + 64 i in ascending r), by
  // slot, for k5_update; kej was stored above by this wavefront
  int n = 0;
  for (int r0 = lane; r0 < pl.Nj; r0 += 64 * kRowG) {
    int jv[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) jv[u] = pl.row_slot[min(r0 + 64 * u, pl.Nj - 1)];
    double kv[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) kv[u] = c.kej[jv[u]];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      if (r0 + 64 * u >= pl.Nj || !(kv[u] > 0.0)) continue;
      if (n < kEmPass) c.emj[n * 64 + lane] = jv[u];
      ++n;
    }
  }
  for (int u = n; u < kEmPass; ++u) c.emj[u * 64 + lane] = -1;
  const int nmax = (int)wave_max((double)n);
  if (lane == 0) c.emj[kEmPass * 64] = nmax;
}


// The scenario's LDS word that carries wavefront 0's decisions (a
// non-positive pivot, the end of the loop) to the other wavefronts: the
// spare slot after the factor's zero slot.
__device__ __forceinline__ lds_f64& wave_flag(const DevPlan& pl, Scn& c) {
  return ((lds_f64*)c.V)[pl.nnzL + 1];
}

// wavefront 0's v on every wavefront of the workgroup (two workgroup
// barriers: the flag slot is free again on return)
__device__ inline int wave0_int(const DevPlan& pl, Scn& c, int v, int w, int nw) {
  if (nw == 1) return v;
  if (w == 0 && lane_id() == 0) wave_flag(pl, c) = (double)v;
  wg_sync();
  const int r = (int)wave_flag(pl, c);
  wg_sync();
  return r;
}

// one Newton iteration; returns relerr (on wavefront 0), sets bad on a
// non-positive pivot (on every wavefront).  With nw > 1 wavefronts the link
// coefficients, the assembly, the copies and the refinement residuals are
// shared; wavefront 0 alone applies the valve terms, factors, runs the
// triangular solves and updates the flows.
__device__ inline double newton_iteration(const DevPlan& pl, Scn& c, const SolveParams& sp,
                                          bool& bad, int w = 0, int nw = 1) {
  WDS_T0();
  k1_coeffs(pl, c, w, nw);
  if (nw > 1) wg_sync();
  WDS_TACC(c, 0);
  k2_assemble<true>(pl, c, w, nw);
  if (nw > 1) wg_sync();
  WDS_TACC(c, 1);
  if (w == 0) {
    k2_prv(pl, c);
    WDS_TACC(c, 7);
    bad = k3_factor(pl, c);
    if (nw > 1 && lane_id() == 0) wave_flag(pl, c) = bad ? 1.0 : 0.0;
  }
  if (nw > 1) {
    wg_sync();
    bad = wave_flag(pl, c) != 0.0;
  }
  WDS_TACC(c, 2);
  if (bad) return 0.0;
  k4_solve(pl, c, c.F, c.Hn, sp.n_refine, w, nw);
  WDS_TACC(c, 3);
  double r = 0.0;
  if (w == 0) r = k5_update(pl, c, sp.hacc);
  WDS_TACC(c, 4);
  return r;
}

__device__ inline int maxtrials_of(const DevPlan& pl, const SolveParams& sp) {
  return sp.max_iter + ((sp.sm && pl.extra_iter > 0) ? pl.extra_iter : 0);
}

struct Ctrl {
  int active, iters, nextcheck, info;
  int nev;  // status changes logged so far (record_schedule)
};

// status-machine bookkeeping after an iteration (dgga/solver.py:3113-3179);
// returns true when the scenario is done
__device__ inline bool control_step(const DevPlan& pl, Scn& c, const SolveParams& sp, int b,
                                    int it, double relerr, Ctrl& st) {
  if (!sp.sm) return !(relerr > sp.hacc);
  SchedLog lg{sp.sched_events ? sp.sched_events + (int64_t)b * sp.sched_cap : nullptr,
              sp.sched_cap, it, st.nev};
  bool vchg = false;
  if (pl.p > 0) vchg = k6_prvstatus(pl, c, lg);
  const bool conv = relerr <= sp.hacc;
  const bool br_conv = conv && it <= sp.max_iter;
  const bool br_per = !conv && it <= pl.maxcheck && st.nextcheck == it;
  bool stat_change = false;
  if (br_conv || br_per) {
    const bool chg = k6_linkstatus(pl, c, lg);
    stat_change = chg && br_conv;
  }
  if (lg.ev) {
    st.nev = lg.nev;
    if (lane_id() == 0)
      sp.sched_flags[(int64_t)b * maxtrials_of(pl, sp) + (it - 1)] =
          (uint8_t)(1 | (br_conv ? 2 : 0) | (br_per ? 4 : 0));
  }
  if (pl.p > 0) stat_change = stat_change || (vchg && br_conv);
  if (br_conv || br_per) {
    if (br_conv && stat_change) st.nextcheck = it + pl.checkfreq;
    if (br_per) st.nextcheck = st.nextcheck + pl.checkfreq;
  }
  return conv && ((it > sp.max_iter) || !stat_change);
}

__device__ inline void write_outputs(const DevPlan& pl, Scn& c, const SolveParams& sp, int b) {
  const int lane = lane_id();
  double* H = sp.head + (int64_t)b * pl.N;
  double* E = sp.emitter + (int64_t)b * pl.N;
  for (int n0 = lane; n0 < pl.N; n0 += 64 * kRowG8) {
    int sl[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) sl[u] = pl.node_slot[min(n0 + 64 * u, pl.N - 1)];
    double kv[kRowG8], ev[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = min(sl[u], pl.Nj - 1);
      kv[u] = c.kej[j];
      ev[u] = c.e[j];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int n = n0 + 64 * u;
      if (n >= pl.N) break;
      H[n] = c.Hn[sl[u]];
      E[n] = (sl[u] < pl.Nj && kv[u] > 0.0) ? ev[u] : 0.0;
    }
  }
}

}  // namespace wds

#include "kernels_polish.h"

namespace wds {

// ---------------------------------------------------------------------------
// CSR values of scenario b into the factor slots (V) and the matrix copy
// used by the refinement residual (A); fill-in positions start at zero
__device__ inline void load_csr(const DevPlan& pl, Scn& c, const double* csr, int b,
                                bool to_v) {
  const int lane = lane_id();
  const double* a = csr + (int64_t)b * pl.nnz_csr;
  if (to_v)
    for (int f = lane; f < pl.nfill; f += 64) c.V[pl.fill_pos[f]] = 0.0;
  for (int e = lane; e < pl.nnz_csr; e += 64) {
    const int t = pl.csr_to_L[e];
    if (t >= 0 && to_v) c.V[t] = a[e];
    c.A[pl.csr_to_ell[e]] = a[e];
  }
}

__global__ WDS_SCENARIO_KERNEL void WDS_K(k_factor_solve)(DevPlan pl, const double* csr,
                                                     const double* F, double* H, int n_refine,
                                                     double* factor_out, int32_t* info,
                                                     double* ws) {
  const int b = blockIdx.x, lane = lane_id();
  WDS_SCENARIO_MEM(ws, b);
  Scn c = make_scn(pl, lds, ws, b, nullptr, nullptr, nullptr, 0);
  load_csr(pl, c, csr, b, true);
  const double* Fb = F + (int64_t)b * pl.Nj;
  for (int r = lane; r < pl.Nj; r += 64) c.F[pl.row_slot[r]] = Fb[r];
  if (lane == 0) c.V[pl.nnzL] = 0.0;
  mem_sync();
  lds_sync();
  const bool bad = k3_factor(pl, c);
  if (bad) {
    if (lane == 0) info[b] = 1;
    return;
  }
  if (factor_out)
    for (int t = lane; t < pl.nnzL; t += 64) factor_out[(int64_t)b * pl.nnzL + t] = c.V[t];
  k4_solve(pl, c, c.F, c.Hn, n_refine);
  double* Hb = H + (int64_t)b * pl.Nj;
  for (int r = lane; r < pl.Nj; r += 64) Hb[r] = c.Hn[pl.row_slot[r]];
  if (lane == 0) info[b] = 0;
}

// Numeric factorisation only (the reference's cuDSS factorize phase,
// dgga/solver.py:2630-2640): factor_out [B][nnzL], info [B].
__global__ WDS_SCENARIO_KERNEL void WDS_K(k_factor)(DevPlan pl, const double* csr,
                                               double* factor_out, int32_t* info, double* ws) {
  const int b = blockIdx.x, lane = lane_id();
  WDS_SCENARIO_MEM(ws, b);
  Scn c = make_scn(pl, lds, ws, b, nullptr, nullptr, nullptr, 0);
  load_csr(pl, c, csr, b, true);
  if (lane == 0) c.V[pl.nnzL] = 0.0;
  mem_sync();
  lds_sync();
  const bool bad = k3_factor(pl, c);
  for (int t = lane; t < pl.nnzL; t += 64) factor_out[(int64_t)b * pl.nnzL + t] = c.V[t];
  if (lane == 0) info[b] = bad ? 1 : 0;
}

// Triangular solves with refinement against a stored factor (the cuDSS
// solve phase and the refinement loop, dgga/solver.py:2641-2651,
// 3052-3064): H [B][Nj] from F [B][Nj], both in original junction order.
__global__ WDS_SCENARIO_KERNEL void WDS_K(k_trisolve)(DevPlan pl, const double* factor,
                                                 const double* csr, const double* F, double* H,
                                                 int n_refine, double* ws) {
  const int b = blockIdx.x, lane = lane_id();
  WDS_SCENARIO_MEM(ws, b);
  Scn c = make_scn(pl, lds, ws, b, nullptr, nullptr, nullptr, 0);
  load_csr(pl, c, csr, b, false);
  for (int t = lane; t < pl.nnzL; t += 64) c.V[t] = factor[(int64_t)b * pl.nnzL + t];
  const double* Fb = F + (int64_t)b * pl.Nj;
  for (int r = lane; r < pl.Nj; r += 64) c.F[pl.row_slot[r]] = Fb[r];
  if (lane == 0) c.V[pl.nnzL] = 0.0;
  mem_sync();
  lds_sync();
  k4_solve(pl, c, c.F, c.Hn, n_refine);
  double* Hb = H + (int64_t)b * pl.Nj;
  for (int r = lane; r < pl.Nj; r += 64) Hb[r] = c.Hn[pl.row_slot[r]];
}

// initial flows, statuses, emitter flows and junction heads of scenario b
// (warm starts when given), loads batched
__device__ inline void init_state(const DevPlan& pl, Scn& c, const SolveParams& sp, int b) {
  const int lane = lane_id();
  const double* q0 = sp.q0 ? sp.q0 + (int64_t)b * pl.L : pl.l_q0;
  const int8_t* s0 = sp.status0 ? sp.status0 + (int64_t)b * pl.L : pl.l_st0;
  for (int k0 = lane; k0 < pl.L; k0 += 64 * kRowG) {
    double qv[kRowG];
    int8_t sv[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      const int k = min(k0 + 64 * u, pl.L - 1);
      qv[u] = q0[k];
      sv[u] = s0[k];
    }
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      const int k = k0 + 64 * u;
      if (k >= pl.L) break;
      c.q[k] = qv[u];
      c.S[k] = sv[u];
    }
  }
  const double* e0 = sp.e0 ? sp.e0 + (int64_t)b * pl.N : nullptr;
  for (int j0 = lane; j0 < pl.Nj; j0 += 64 * kRowG8) {
    int nd[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) nd[u] = pl.slot_node[min(j0 + 64 * u, pl.Nj - 1)];
    double ev[kRowG8], kv[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      ev[u] = e0 ? e0[nd[u]] : 0.0;
      kv[u] = c.kej[min(j0 + 64 * u, pl.Nj - 1)];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = j0 + 64 * u;
      if (j >= pl.Nj) break;
      c.e[j] = e0 ? ev[u] : (kv[u] > 0.0 ? 1.0 : 0.0);
      c.Hn[j] = 0.0;
    }
  }
}

// Fused forward: the whole Newton loop of scenario b, its layout at lds.
// The workgroup's wavefronts (blockDim.x / 64) share the Newton iterations
// and the polish; the status machine and the outputs are wavefront 0's.
__device__ __forceinline__ void solve_scenario(const DevPlan& pl, const SolveParams& sp, int b,
                                      double* lds) {
  const int lane = lane_id();
  const int w = wave_idx(), nw = (int)(blockDim.x >> 6);
  Scn c = make_scn(pl, lds, sp.ws, b, sp.flow, sp.status, sp.r, sp.r_stride);
  if (w == 0) {
    load_inputs(pl, c, sp, b);
    init_state(pl, c, sp, b);
#ifdef WDS_PROFILE_STAGES
    for (int i = lane; i < 112; i += 64) c.misc[400 + i] = 0.0;
    // marks the counter block for the host that reads it back
    if (lane == 0) c.misc[511] = 12345.0;
#endif
    mem_sync();
    lds_sync();
  }
  if (nw > 1) wg_sync();
  c.em_list = true;  // load_inputs built the emitter row list
  const int maxtrials = maxtrials_of(pl, sp);
  Ctrl st{1, 0, pl.checkfreq, 0, 0};
  double relerr_out = 0.0, dratio = 0.0;
  for (int it = 1; it <= maxtrials; ++it) {
    bool bad = false;
    const double relerr = newton_iteration(pl, c, sp, bad, w, nw);
    if (bad) {
      st.info = it;
      break;
    }
    st.iters = it;
    relerr_out = relerr;
    WDS_T0();
    bool done = false;
    if (w == 0) done = control_step(pl, c, sp, b, it, relerr, st);
    if (nw > 1) {
      if (w == 0 && lane == 0) wave_flag(pl, c) = done ? 1.0 : 0.0;
      wg_sync();
      done = wave_flag(pl, c) != 0.0;
    }
    WDS_TACC(c, 5);
    if (done) {
      st.active = 0;
      break;
    }
  }
  // every wavefront has read the last flag before the polish writes it again
  if (nw > 1) wg_sync();
  if (st.active && st.info == 0) st.iters = maxtrials + 1;
  // the last assembled matrix is still in A (the polish assembles its own
  // after a workgroup barrier)
  if (w == 0 && pl.p > 0 && sp.diag_ratio) dratio = diag_ratio_now(pl, c);
  double resid = 0.0;
  // st.info is uniform: a bad pivot reached every wavefront
  if (st.info == 0 && sp.polish_steps >= 0) {
    WDS_T0();
    const int pinfo = polish_and_factor(pl, c, sp, b, resid, w, nw);
    WDS_TACC(c, 6);
    if (pinfo) st.info = pinfo;
  }
  // the outputs are wavefront 0's (no workgroup barrier follows)
  if (w != 0) return;
  write_outputs(pl, c, sp, b);
  if (lane == 0) {
    sp.iters[b] = st.iters;
    sp.relerr[b] = relerr_out;
    sp.info[b] = st.info;
    if (sp.diag_ratio) sp.diag_ratio[b] = dratio;
    if (sp.resid) sp.resid[b] = resid;
    if (sp.sched_count) sp.sched_count[b] = st.nev;
  }
}

// one scenario per workgroup (scenario blockIdx.x), kSolveWaves wavefronts
__global__ WDS_SOLVE_KERNEL void WDS_K(k_solve)(DevPlan pl, SolveParams sp) {
  const int b = blockIdx.x;
  WDS_SCENARIO_MEM(sp.ws, b);
  solve_scenario(pl, sp, b, lds);
}

// One Newton iteration per launch (the unfused path).  ctrl[b] = {active,
// iters, nextcheck, info}; heads/flows/emitters/statuses persist in the
// output arrays between launches (it == 1 initialises them).
__global__ WDS_SCENARIO_KERNEL void WDS_K(k_newton_iter)(DevPlan pl, SolveParams sp, int it,
                                                    int32_t* ctrl) {
  const int b = blockIdx.x, lane = lane_id();
  WDS_SCENARIO_MEM(sp.ws, b);
  int32_t* cb = ctrl + 4 * b;
  Scn c = make_scn(pl, lds, sp.ws, b, sp.flow, sp.status, sp.r, sp.r_stride);
  load_inputs(pl, c, sp, b);
  c.em_list = true;
  const double* ke = sp.ke + (int64_t)b * sp.ke_stride;
  double* H = sp.head + (int64_t)b * pl.N;
  double* E = sp.emitter + (int64_t)b * pl.N;
  if (it == 1) {
    for (int k = lane; k < pl.L; k += 64) {
      c.q[k] = sp.q0 ? sp.q0[(int64_t)b * pl.L + k] : pl.l_q0[k];
      c.S[k] = sp.status0 ? sp.status0[(int64_t)b * pl.L + k] : pl.l_st0[k];
    }
    for (int j = lane; j < pl.Nj; j += 64) {
      const int node = pl.slot_node[j];
      c.e[j] = sp.e0 ? sp.e0[(int64_t)b * pl.N + node] : (ke[node] > 0.0 ? 1.0 : 0.0);
      c.Hn[j] = 0.0;
    }
    if (lane == 0) {
      cb[0] = 1;
      cb[1] = 0;
      cb[2] = pl.checkfreq;
      cb[3] = 0;
    }
  } else {
    for (int j = lane; j < pl.Nj; j += 64) {
      const int node = pl.slot_node[j];
      c.Hn[j] = H[node];
      c.e[j] = E[node];
    }
  }
  mem_sync();
  lds_sync();
  Ctrl st{cb[0], cb[1], cb[2], cb[3], 0};
  if (sp.sched_count) st.nev = it == 1 ? 0 : sp.sched_count[b];
  if (!st.active || st.info) return;
  const int maxtrials = maxtrials_of(pl, sp);
  bool bad = false;
  double dratio = 0.0;
  const double relerr = newton_iteration(pl, c, sp, bad);
  if (pl.p > 0 && sp.diag_ratio) dratio = diag_ratio_now(pl, c);
  if (bad) {
    st.info = it;
  } else {
    st.iters = it;
    if (control_step(pl, c, sp, b, it, relerr, st)) st.active = 0;
    if (st.active && it == maxtrials) st.iters = maxtrials + 1;
  }
  write_outputs(pl, c, sp, b);
  if (lane == 0) {
    cb[0] = st.active;
    cb[1] = st.iters;
    cb[2] = st.nextcheck;
    cb[3] = st.info;
    if (!bad) sp.relerr[b] = relerr;
    sp.iters[b] = st.iters;
    sp.info[b] = st.info;
    if (sp.diag_ratio && pl.p > 0) sp.diag_ratio[b] = dratio;
    if (sp.sched_count) sp.sched_count[b] = st.nev;
  }
}

// ---------------------------------------------------------------------------
// Ragged batches: many networks, each with its own plan and arrays, in one
// persistent launch.  The work items (network, scenario) are in the order the
// host chose (longest first); every workgroup takes the next item from a
// global counter, solves it with the network's plan, and exits when the list
// is exhausted.  In LDS mode the dynamic LDS is the largest layout of the
// group and every item reuses it; in arena mode each item's layout is its
// arena.  The kSolveWaves wavefronts of a workgroup share each item; the item
// index reaches them through an LDS word: in LDS mode the dynamic LDS word at
// byte next_at, after every item's layout (the layouts address the dynamic
// LDS from 0, so the kernels declare no static LDS); in arena mode a static
// one.
__global__ WDS_SOLVE_KERNEL void WDS_K(k_solve_multi)(const DevPlan* plans,
                                                      const SolveParams* params,
                                                      const int2* items, int32_t nitems,
                                                      int32_t* counter, int32_t next_at) {
#ifndef WDS_GLOBAL_ARENA
  __attribute__((address_space(3))) int32_t* const next =
      (__attribute__((address_space(3))) int32_t*)(size_t)(unsigned)next_at;
#else
  __shared__ int32_t next_word;
  int32_t* const next = &next_word;
  (void)next_at;
#endif
  for (;;) {
    if (threadIdx.x == 0) *next = atomicAdd(counter, 1);
    wg_sync();
    const int i = *next;
    // every wavefront has read the index before the next item's is stored
    wg_sync();
    if (i >= nitems) break;
    const int2 w = items[i];
    const int net = __builtin_amdgcn_readfirstlane(w.x);
    const int b = __builtin_amdgcn_readfirstlane(w.y);
    const DevPlan& pl = plans[net];
    const SolveParams& sp = params[net];
    WDS_SCENARIO_MEM(sp.ws, b);
    solve_scenario(pl, sp, b, lds);
    mem_sync();
    lds_sync();
  }
}

__global__ WDS_SOLVE_KERNEL void WDS_K(k_adjoint_multi)(const DevPlan* plans,
                                                        const AdjParams* params,
                                                        const int2* items, int32_t nitems,
                                                        int32_t* counter, int32_t next_at) {
#ifndef WDS_GLOBAL_ARENA
  __attribute__((address_space(3))) int32_t* const next =
      (__attribute__((address_space(3))) int32_t*)(size_t)(unsigned)next_at;
#else
  __shared__ int32_t next_word;
  int32_t* const next = &next_word;
  (void)next_at;
#endif
  for (;;) {
    if (threadIdx.x == 0) *next = atomicAdd(counter, 1);
    wg_sync();
    const int i = *next;
    wg_sync();
    if (i >= nitems) break;
    const int2 w = items[i];
    const int net = __builtin_amdgcn_readfirstlane(w.x);
    const int b = __builtin_amdgcn_readfirstlane(w.y);
    const DevPlan& pl = plans[net];
    const AdjParams& ap = params[net];
    WDS_SCENARIO_MEM(ap.ws, b);
    adjoint_scenario(pl, ap, b, lds);
    mem_sync();
    lds_sync();
  }
}

}  // namespace wds
