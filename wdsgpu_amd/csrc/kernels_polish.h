// Reduced full-Newton polish, terminal factor and implicit adjoint.
//
// Mathematics of dgga/autodiff.py:735-1312 (the GPU batched reduced
// adjoint): eliminating the link and emitter rows of the full residual
// Jacobian leaves the converged GGA node matrix A; ACTIVE PRVs (big-M rows in
// A) are handled by Woodbury row replacement so the symmetric factor of A is
// reused for the primal polish system and for the adjoint node system.
#pragma once
#include "device_common.h"

namespace wds {

// Woodbury row replacement without column solves.  With A-hat = L D L^T,
// e_u^T A-hat^-1 e_v = y_u^T D^-1 y_v and e_u^T A-hat^-1 r = y_u^T z, where
// y_u = L^-1 e_u (nonzero only on the elimination-tree path from u to its
// root) and z = D^-1 L^-1 r is what the forward program leaves in X.  The
// capacitance matrix therefore costs O(height^2) per valve end, and the
// correction A-hat^-1 (r - sum_c w_c e_{u_c}) is applied to z between the
// forward and backward programs, so one Woodbury solve is one solve pair
// (the reference solves A-hat for each column, dgga/autodiff.py:965-1029,
// 1107-1154).

// Coefficients of the polish / adjoint at the current state
// (_adjgpu_coeffs, dgga/autodiff.py:786-875): link P/Y with pcv PRVs taking
// their prvcoeff values (ACTIVE -> P = 0), emitter hloss/hgrad.  Writes P, Y
// (global), hle/hge and the per-valve ACTIVE mask into sh[0..p).
// With nw > 1 wavefronts the link and emitter loops are shared and
// wavefront 0 writes the valve rows after a workgroup barrier (they
// overwrite what the link loop wrote for the valve links).
__device__ inline void polish_coeffs(const DevPlan& pl, Scn& c, int w = 0, int nw = 1) {
  const int lane = lane_id();
  link_coeffs_all(pl, c.q, c.r, c.S, c.P, c.Y, w, nw);
  emitter_coeffs(pl, c, w, nw);
  mem_sync();
  if (nw > 1) wg_sync();
  const int i = lane;
  if (w == 0 && i < pl.p) {
    const int k = pl.prv_k[i];
    const double qk = c.q[k];
    const int8_t sk = c.S[k];
    const bool act = sk == ST_ACTIVE, cl = sk <= ST_CLOSED;
    double pk_o, yk_o;
    valve_open(pl, qk, pl.prv_km[i], pk_o, yk_o);
    const double pk_na = cl ? 1.0 / CBIG : pk_o;
    const double yk_na = cl ? qk : yk_o;
    c.P[k] = pk_na * (act ? 0.0 : 1.0);
    c.Y[k] = act ? qk : yk_na;
    c.sh[i] = act ? 1.0 : 0.0;
  }
  mem_sync();
  lds_sync();
}

// Assemble A-hat = A(P) + diag(em/hge) + CBIG * ACTIVE at j2 into A and V
// (rows shared by the nw wavefronts; the ACTIVE pins by wavefront 0 after a
// workgroup barrier)
__device__ inline void polish_assemble(const DevPlan& pl, Scn& c, int w = 0, int nw = 1) {
  k2_assemble<false>(pl, c, w, nw);
  if (nw > 1) wg_sync();
  if (w == 0 && lane_id() == 0) {
    for (int i = 0; i < pl.p; ++i) {
      if (c.sh[i] != 0.0) {
        const int d2 = pl.diag_pos[pl.prv_s2[i]];
        const double v = c.V[d2] + CBIG;  // V holds A's values here
        c.A[pl.prv_ae[i].w] = v;
        c.V[d2] = v;
      }
    }
  }
  mem_sync();
  lds_sync();
}

// (A-hat x)_i for all rows into dst (global)
// dst (global) = A x (x in LDS)
__device__ inline void matvec(const DevPlan& pl, Scn& c, const double* x, double* dst) {
  spmv_all(pl, c.A, x, dst);
  mem_sync();
}

// one stage of the (largest value, lowest index) reduction within rows of 16
// lanes; the combination is commutative and associative, so after the four
// stages (xor 1, xor 2, half mirror, mirror) every lane of a row of 16 holds
// its row's result
template <int CTRL>
__device__ __forceinline__ void argmax_stage(double& v, int& idx) {
  const double ov = dpp_f64<CTRL>(v);
  const int oi = __builtin_amdgcn_update_dpp(0, idx, CTRL, 0xf, 0xf, false);
  if (ov > v || (ov == v && oi < idx)) {
    v = ov;
    idx = oi;
  }
}

// Solve the n x n system M y = w (row-major, LDS; n <= 2 kMaxPrv = 16) by
// Gaussian elimination with partial pivoting (largest magnitude, lowest row
// on ties), rows eliminated in parallel, one lane per row; y overwrites w.
// The pivot search runs within the first 16 lanes (DPP) and reaches every
// lane from lane 0.  Returns false if singular.
__device__ inline bool small_solve(double* M, double* w, int n) {
  const int lane = lane_id();
  for (int col = 0; col < n; ++col) {
    double v = (lane >= col && lane < n) ? fabs(M[lane * n + col]) : -1.0;
    int idx = lane;
    argmax_stage<0xB1>(v, idx);   // quad_perm [1, 0, 3, 2]
    argmax_stage<0x4E>(v, idx);   // quad_perm [2, 3, 0, 1]
    argmax_stage<0x141>(v, idx);  // row_half_mirror
    argmax_stage<0x140>(v, idx);  // row_mirror
    {
      const long long vb = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_readfirstlane((int)vb);
      const int hi = __builtin_amdgcn_readfirstlane((int)(vb >> 32));
      v = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
    if (!(v > 0.0)) return false;  // uniform: lane 0's maximum
    const int piv = __builtin_amdgcn_readfirstlane(idx);
    if (piv != col) {
      if (lane < n) {
        const double t = M[col * n + lane];
        M[col * n + lane] = M[piv * n + lane];
        M[piv * n + lane] = t;
      } else if (lane == n) {
        const double t = w[col];
        w[col] = w[piv];
        w[piv] = t;
      }
      lds_sync();
    }
    if (lane > col && lane < n) {
      const double f = M[lane * n + col] / M[col * n + col];
      if (f != 0.0) {
        for (int cc = col; cc < n; ++cc) M[lane * n + cc] -= f * M[col * n + cc];
        w[lane] -= f * w[col];
      }
    }
    lds_sync();
  }
  if (lane == 0)
    for (int r = n - 1; r >= 0; --r) {
      double s = w[r];
      for (int cc = r + 1; cc < n; ++cc) s -= M[r * n + cc] * w[cc];
      w[r] = s / M[r * n + r];
    }
  lds_sync();
  return true;
}

// Paths of the valve ends for the factor in V: y_u, z_u = D^-1 y_u (one
// lane per path, in path order) and C[u][v] = y_u . z_v over the shared
// ancestors of u and v.
__device__ inline void wood_paths(const DevPlan& pl, Scn& c) {
  const int lane = lane_id();
  // the path nodes' slots into LDS for the probes and corrections of the
  // solves that follow (their loads overlap the first path's)
  for (int i0 = lane; i0 < pl.wp_tot; i0 += 64 * kRowG) {
    int nd[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) nd[u] = pl.wp_node[min(i0 + 64 * u, pl.wp_tot - 1)];
#pragma unroll
    for (int u = 0; u < kRowG; ++u)
      if (i0 + 64 * u < pl.wp_tot) c.wpn[i0 + 64 * u] = nd[u];
  }
  if (lane < pl.np && pl.wq_ok) {
    // fixed-stride terms: the next node's words load while this node's
    // sum runs, so the chain waits on LDS only (same terms, same order)
    const int i0 = pl.wp_ptr[lane], len = pl.wp_ptr[lane + 1] - i0;
    const int2* q = pl.wq + lane * kPathMax * kTermMax;
    const int32_t* qn = pl.wq_n + lane * kPathMax;
    int2 tm[kTermMax];
#pragma unroll
    for (int m = 0; m < kTermMax; ++m) tm[m] = q[m];
    int n = qn[0], dg = pl.wp_diag[i0];
    for (int k = 0; k < len; ++k) {
      const int k1 = min(k + 1, len - 1);
      int2 tn[kTermMax];
#pragma unroll
      for (int m = 0; m < kTermMax; ++m) tn[m] = q[k1 * kTermMax + m];
      const int nn = qn[k1], dgn = pl.wp_diag[i0 + k1];
      double acc = (k == 0) ? 1.0 : 0.0;
#pragma unroll
      for (int m = 0; m < kTermMax; ++m)
        if (m < n) acc = fma(-c.V[tm[m].x], c.Yp[tm[m].y], acc);  // V: the unit factor L
      c.Yp[i0 + k] = acc;
      c.Zp[i0 + k] = acc * c.V[dg];
#pragma unroll
      for (int m = 0; m < kTermMax; ++m) tm[m] = tn[m];
      n = nn;
      dg = dgn;
    }
  } else if (lane < pl.np) {
    const int i0 = pl.wp_ptr[lane], i1 = pl.wp_ptr[lane + 1];
    for (int i = i0; i < i1; ++i) {
      double acc = (i == i0) ? 1.0 : 0.0;
      for (int t = pl.wt_ptr[i]; t < pl.wt_ptr[i + 1]; ++t) {
        const int2 e = pl.wt[t];
        acc = fma(-c.V[e.x], c.Yp[e.y], acc);  // V holds the unit factor L
      }
      c.Yp[i] = acc;
      c.Zp[i] = acc * c.V[pl.wp_diag[i]];
    }
  }
  lds_sync();
  for (int pr = lane; pr < pl.np * pl.np; pr += 64) {
    const int4 w = pl.wpair[pr];
    double s = 0.0;
    for (int t = 0; t < w.z; ++t) s = fma(c.Yp[w.x + t], c.Zp[w.y + t], s);
    c.Cm[pr] = s;
  }
  lds_sync();
}

// e_u^T A-hat^-1 r for every valve end u, from z = D^-1 L^-1 r in X (LDS)
// (one lane per path, its range [i0, i1) loaded by the caller; the path's
// node slots come from LDS kPathG at a time, the sum runs in path order)
constexpr int kPathG = 16;
__device__ inline void wood_probe(const DevPlan& pl, Scn& c, const double* X, double* x0,
                                  int i0, int i1) {
  const int lane = lane_id();
  if (lane < pl.np) {
    double s = 0.0;
    for (int b = i0; b < i1; b += kPathG) {
      int nd[kPathG];
#pragma unroll
      for (int u = 0; u < kPathG; ++u) nd[u] = c.wpn[min(b + u, i1 - 1)];
#pragma unroll
      for (int u = 0; u < kPathG; ++u)
        if (b + u < i1) s = fma(c.Yp[b + u], X[nd[u]], s);
    }
    x0[lane] = s;
  }
  lds_sync();
}

// z -= sum_c w[c] z_{u_c} for the first ncol paths (the forward image of
// r - sum_c w[c] e_{u_c}); paths of different ends may share nodes, so one
// path at a time
__device__ inline void wood_apply(const DevPlan& pl, Scn& c, double* X, const double* w,
                                  int ncol) {
  const int lane = lane_id();
  constexpr int kApG = 4;
  if (pl.wp_tot <= 64 * kApG) {
    // every path node's slot loads once, up front (entry i = lane + 64 u);
    // the columns then run in order from registers and LDS
    int nd[kApG];
#pragma unroll
    for (int u = 0; u < kApG; ++u) nd[u] = c.wpn[min(lane + 64 * u, pl.wp_tot - 1)];
    for (int col = 0; col < ncol; ++col) {
      const double wc = w[col];
      const int c0 = pl.wp_ptr[col], c1 = pl.wp_ptr[col + 1];
#pragma unroll
      for (int u = 0; u < kApG; ++u) {
        const int i = lane + 64 * u;
        if (i >= c0 && i < c1) X[nd[u]] -= wc * c.Zp[i];
      }
      lds_sync();
    }
    return;
  }
  for (int col = 0; col < ncol; ++col) {
    const double wc = w[col];
    for (int i = pl.wp_ptr[col] + lane; i < pl.wp_ptr[col + 1]; i += 64)
      X[c.wpn[i]] -= wc * c.Zp[i];
    lds_sync();
  }
}

// One solve of the row-replaced system for right-hand side rr (global); the
// result is left in the solve vector S.  adjoint = false: primal M_p (rows
// j2 -> dH[j2] = rr[j2], rows j1 merged; 2p columns e_j2, e_j1); adjoint =
// true: node system M (rows j2 -> lam[j2] - lam[j1] = rr[j2]; p columns e_j2).
// (in_S: the right-hand side is already in S, rr unused)
__device__ inline bool wood_solve(const DevPlan& pl, Scn& c, const double* rr, bool adjoint,
                                  bool in_S = false) {
  const int lane = lane_id();
  const int p = pl.p;
  double* M = c.sM;            // [n x n], n = 2p or p
  double* w = c.sw;            // [n]
  double* x0 = c.sx0;          // [2p]
  double* const S = c.R;
  WDS_T0();
  // plan words of the probes and the valve rows, loaded before the copy and
  // the forward solve so that their latency hides under them
  const int i0 = pl.wp_ptr[min(lane, pl.np - 1)], i1 = pl.wp_ptr[min(lane, pl.np - 1) + 1];
  const int s2v = pl.prv_s2[min(lane, p - 1)];
  i32x4 ring[kSlots];
  prefetch_prog(pl.fwd, ring);
  if (!in_S) vec_copy(S, rr, nullptr, pl.Nj);
  lds_sync();
  // rr at the valves' downstream rows: S holds rr until the forward solve
  const double r2v = S[s2v];
  WDS_TACC(c, 23);
  solve_fwd(pl, c, ring);
  WDS_TACC(c, 24);
  wood_probe(pl, c, S, x0, i0, i1);
  const int n = adjoint ? p : 2 * p;
  // capacitance rows, one entry per lane: row a < p is valve a's replaced
  // downstream row; rows p + a (primal only) its merged upstream row
  for (int e = lane; e < n * n; e += 64) {
    const int row = e / n, col = e % n;
    const int a = row < p ? row : row - p;
    const double act = c.sh[a];
    const double s2 = c.Cm[a * pl.np + col];          // e_j2a^T A^-1 e_col
    double v;
    if (adjoint) {
      const double s1 = c.Cm[(p + a) * pl.np + col];  // e_j1a^T A^-1 e_col
      v = (s2 - s1) * (act * c.sh[col]) + ((a == col) ? 1.0 - act : 0.0);
    } else if (row < p) {
      v = s2 * act + ((a == col) ? 1.0 - act : 0.0);
    } else {
      const double eye = (col == a || col == a + p) ? 1.0 : 0.0;
      v = (eye - CBIG * s2) * act + ((a + p == col) ? 1.0 - act : 0.0);
    }
    M[row * n + col] = v;
  }
  if (lane < p) {
    const int a = lane;
    const double act = c.sh[a];
    const double r2 = r2v;
    if (adjoint) {
      w[a] = (x0[a] - x0[p + a] - r2) * act;
    } else {
      w[a] = (x0[a] - r2) * act;
      w[a + p] = (r2 - CBIG * x0[a]) * act;
    }
  }
  lds_sync();
  if (!small_solve(M, w, n)) return false;
  wood_apply(pl, c, S, w, n);
  WDS_TACC(c, 25);
  solve_bwd(pl, c, ring);
  WDS_TACC(c, 26);
  return true;
}

// Primal reduced Newton system M_p dH = r_p (_adjgpu_primal_solve,
// dgga/autodiff.py:965-1029).  rhs_p in c.v1, r_prv in sprv; result in dst
// (LDS, not the solve vector).  Uses v3 (rp), v4 (work), v5 (residual).
__device__ inline bool primal_solve(const DevPlan& pl, Scn& c, double* dst) {
  const int lane = lane_id();
  const int p = pl.p;
  bool any_act = false;
  for (int i = 0; i < p; ++i) any_act |= c.sh[i] != 0.0;
  if (!any_act) {
    k4_solve(pl, c, c.v1, dst, 2);
    return true;
  }
  double* rp = c.v3;
  double* wk = c.v4;
  const double* rprv = c.sprv;
  vec_copy(rp, c.v1, nullptr, pl.Nj);
  mem_sync();
  if (lane == 0) {
    // rp[j1] += v1[j2] for the active valves in valve order, then
    // rp[j2] = -r_prv (active) or v1[j2]: every value loads first (one round
    // trip), the sums run in registers (an upstream node shared by two
    // valves takes both terms in order; no upstream node is a downstream one)
    int s1[kMaxPrv], s2[kMaxPrv];
    double a1[kMaxPrv], a2[kMaxPrv], r1[kMaxPrv];
#pragma unroll
    for (int i = 0; i < kMaxPrv; ++i) {
      const int k = min(i, p - 1);
      s1[i] = pl.prv_s1[k];
      s2[i] = pl.prv_s2[k];
    }
#pragma unroll
    for (int i = 0; i < kMaxPrv; ++i) {
      a1[i] = c.v1[s1[i]];
      a2[i] = c.v1[s2[i]];
    }
#pragma unroll
    for (int i = 0; i < kMaxPrv; ++i) {
      double cur = a1[i];
#pragma unroll
      for (int k = 0; k < i; ++k)
        if (c.sh[k] != 0.0 && s1[k] == s1[i]) cur = r1[k];
      r1[i] = cur + a2[i];
    }
#pragma unroll
    for (int i = 0; i < kMaxPrv; ++i)
      if (i < p && c.sh[i] != 0.0) rp[s1[i]] = r1[i];
#pragma unroll
    for (int i = 0; i < kMaxPrv; ++i)
      if (i < p) rp[s2[i]] = (c.sh[i] != 0.0) ? -rprv[i] : a2[i];
  }
  mem_sync();
  WDS_T0();
  wood_paths(pl, c);
  WDS_TACC(c, 32);
  if (!wood_solve(pl, c, rp, false)) return false;
  WDS_TACC(c, 33);
  for (int j = lane; j < pl.Nj; j += 64) dst[j] = c.R[j];
  lds_sync();
  // two refinement steps on the residual of the true replaced system
  for (int ref = 0; ref < 2; ++ref) {
    matvec(pl, c, dst, wk);  // base = A-hat dH
    // the residual goes straight into the solve vector S
    double* res = c.R;
    vec_copy(res, rp, wk, pl.Nj);
    lds_sync();
    if (lane == 0) {
      // the replaced rows of the active valves, stored in valve order (all
      // loads first: none reads res)
      int s1[kMaxPrv], s2[kMaxPrv];
      double w1[kMaxPrv], w2[kMaxPrv], p1[kMaxPrv], p2[kMaxPrv];
#pragma unroll
      for (int i = 0; i < kMaxPrv; ++i) {
        const int k = min(i, p - 1);
        s1[i] = pl.prv_s1[k];
        s2[i] = pl.prv_s2[k];
      }
#pragma unroll
      for (int i = 0; i < kMaxPrv; ++i) {
        w1[i] = wk[s1[i]];
        w2[i] = wk[s2[i]];
        p1[i] = rp[s1[i]];
        p2[i] = rp[s2[i]];
      }
#pragma unroll
      for (int i = 0; i < kMaxPrv; ++i) {
        if (i >= p || c.sh[i] == 0.0) continue;
        const double d2 = dst[s2[i]];
        const double row1 = w1[i] + w2[i] - CBIG * d2;
        res[s1[i]] = p1[i] - row1;
        res[s2[i]] = p2[i] - d2;
      }
    }
    lds_sync();
    if (!wood_solve(pl, c, nullptr, false, true)) return false;
    for (int j = lane; j < pl.Nj; j += 64) dst[j] += c.R[j];
    lds_sync();
  }
  return true;
}

// link residuals r_l (into Y), PRV residuals (sprv), emitter residual
// r_e (into v4 when want_re) and node mass m -> rhs_p in v1
// (_adjgpu_residuals + the right-hand side of _adjgpu_polish,
// dgga/autodiff.py:1032-1050, 1066-1075).  Returns the inf-norm residual
// (on wavefront 0).  With nw > 1 wavefronts the link loop and the mass row
// blocks are shared (a workgroup barrier between them: the mass rows read
// every link's residual) and the maxima meet in LDS.
__device__ inline double polish_residuals(const DevPlan& pl, Scn& c, bool build_rhs, int w = 0,
                                          int nw = 1) {
  const int lane = lane_id();
  double rmax = 0.0;
  for (int k0 = lane + 64 * kGather * w; k0 < pl.L; k0 += 64 * kGather * nw) {
    double Pv[kGather], Yv[kGather];
    int s1[kGather], s2[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = min(k0 + 64 * u, pl.L - 1);
      Pv[u] = c.P[k];
      Yv[u] = c.Y[k];
      s1[u] = pl.l_s1[k];
      s2[u] = pl.l_s2[k];
    }
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = k0 + 64 * u;
      if (k >= pl.L) break;
      const double P = Pv[u];
      const double hl = (P != 0.0) ? Yv[u] / P : 0.0;
      const double rl = hl - (c.Hn[s1[u]] - c.Hn[s2[u]]);
      c.Y[k] = rl;
      if (P != 0.0) rmax = fmax(rmax, fabs(rl));
    }
  }
  double* rprv = c.sprv;
  if (w == 0 && lane < pl.p) {
    const double rp = (c.Hn[pl.prv_s2[lane]] - pl.prv_hset[lane]) * c.sh[lane];
    rprv[lane] = rp;
    rmax = fmax(rmax, fabs(rp));
  }
  mem_sync();
  lds_sync();
  if (nw > 1) wg_sync();
  // mass rows over the polish row layout: a row's links with n2 == j first
  // (reference scatter order), then those with n1 == j negated; two 64-row
  // blocks per pass, every load of a pass issued before use; the blocks in
  // nw contiguous shares of about equal entry counts
  const int nb = (pl.Nj + 63) >> 6;
  const int bw0 = block_split(pl.ps_blk, nb, pl.ps_blk[nb], w, nw);
  const int bw1 = block_split(pl.ps_blk, nb, pl.ps_blk[nb], w + 1, nw);
  const lds_f64* Hl = (const lds_f64*)c.Hn;
  for (int b0 = bw0; b0 < bw1; b0 += 2) {
    int e0[2], K[2], len[2], jr[2];
    double kv[2], hlv[2], hgv[2], ev[2], djv[2], elv[2], hn[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bk = min(b0 + t, bw1 - 1);
      e0[t] = pl.ps_blk[bk];
      K[t] = b0 + t < bw1 ? pl.ps_blk[bk + 1] - e0[t] : 0;
      jr[t] = bk * 64 + lane;
      const int j = min(jr[t], pl.Nj - 1);
      len[t] = pl.ps_len[j];
      kv[t] = c.kej[j];
      hlv[t] = c.hle[j];
      hgv[t] = c.hge[j];
      ev[t] = c.e[j];
      djv[t] = c.dj[j];
      elv[t] = pl.el_slot[j];
      hn[t] = Hl[j];
    }
    double m[2] = {0.0, 0.0}, bt[2] = {0.0, 0.0};
    const int Km = max(K[0], K[1]);
    for (int e = 0; e < Km; e += kAsE) {
      int en[2][kAsE];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u)
          en[t][u] = pl.ps_ell[(e0[t] + min(e + u, max(K[t] - 1, 0))) * 64 + lane];
      __builtin_amdgcn_sched_barrier(0);
      double qv[2][kAsE], Pv[2][kAsE], Yv[2][kAsE];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u) {
          const int k = en[t][u] & 0x7fffffff;
          qv[t][u] = c.q[k];
          Pv[t][u] = c.P[k];
          Yv[t][u] = c.Y[k];
        }
      __builtin_amdgcn_sched_group_barrier(0x0020, 6 * kAsE, 0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < kAsE; ++u) {
          if (e + u >= len[t]) continue;
          const bool n2side = en[t][u] < 0;
          const double pw = Pv[t][u] * Yv[t][u];
          m[t] += n2side ? qv[t][u] : -qv[t][u];
          bt[t] += n2side ? pw : -pw;
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (!(b0 + t < bw1 && jr[t] < pl.Nj)) continue;
      double mm = m[t], re = 0.0, emh = 0.0;
      if (kv[t] > 0.0) {
        re = hlv[t] - (hn[t] - elv[t]);
        emh = 1.0 / hgv[t];
        mm = mm - ev[t];
        rmax = fmax(rmax, fabs(re));
      }
      mm = mm - djv[t];
      rmax = fmax(rmax, fabs(mm));
      if (build_rhs) {
        c.v1[jr[t]] = mm - bt[t] + emh * re;
        c.v4[jr[t]] = re;
      }
    }
  }
  mem_sync();
  double m = wave_max(rmax);
  if (nw > 1) {
    // the maxima of the other wavefronts, one at a time through the flag
    // slot (a maximum is exact in any order)
    for (int v = 1; v < nw; ++v) {
      if (w == v && lane == 0) wave_flag(pl, c) = m;
      wg_sync();
      if (w == 0) m = fmax(m, wave_flag(pl, c));
      wg_sync();
    }
  }
  return m;
}

// A CONST_HP pump on its CBIG / RQtol clamp branch: the assembled P is
// +CBIG / +RQtol while the residual's derivative is -CBIG / -RQtol, so the
// reduced system is not the Jacobian's Schur complement there and the
// reference refuses the adjoint (dgga/autodiff.py:808-823).
constexpr int kInfoChpClamp = -300;

__device__ inline bool chp_on_clamp(const DevPlan& pl, const Scn& c) {
  if (!pl.has_pump) return false;
  bool bad = false;
  for (int k = lane_id(); k < pl.L; k += 64) {
    const int f = pl.l_flags[k];
    if (!(f & LF_PUMP_CHP) || c.S[k] <= ST_CLOSED || (f & LF_ZERO_SPEED)) continue;
    const double qa = fmax(fabs(c.q[k]), 1e-30);
    const double spc = fmax(pl.pu_speed[k], 1e-12);
    const double re = pl.pu_r[k] * det_pow(spc, 2.0 - pl.pu_n[k]);
    const double raw = -re / qa / qa;
    bad = bad || raw > CBIG || raw < pl.rqtol;
  }
  return wave_any(bad);
}

// Polish steps + terminal factor (_adjgpu_polish, dgga/autodiff.py:1053-1104).
// Returns 0, or a negative info code on failure (on every wavefront).  With
// nw > 1 wavefronts the coefficient, assembly, residual and update loops are
// shared; wavefront 0 alone factors, runs the Woodbury primal solve and
// applies the valve rows, and its outcome reaches the others through the
// flag slot, so every wavefront returns at the same point.
__device__ inline int polish_and_factor(const DevPlan& pl, Scn& c, const SolveParams& sp, int b,
                                        double& resid, int w = 0, int nw = 1) {
  const int lane = lane_id();
  double* dH = c.X2;  // LDS
  for (int step = 0; step < sp.polish_steps; ++step) {
    WDS_T0();
    int fail = 0;
    if (w == 0 && chp_on_clamp(pl, c)) fail = kInfoChpClamp;
    fail = wave0_int(pl, c, fail, w, nw);
    if (fail) return fail;
    polish_coeffs(pl, c, w, nw);
    if (nw > 1) wg_sync();
    polish_assemble(pl, c, w, nw);
    WDS_TACC(c, 104);
    if (w == 0 && k3_factor(pl, c)) fail = -(step + 1);
    fail = wave0_int(pl, c, fail, w, nw);
    if (fail) return fail;
    WDS_TACC(c, 105);
    polish_residuals(pl, c, true, w, nw);
    // v4 holds r_e; keep it in F (free during the polish) before primal_solve
    vec_copy(c.F, c.v4, nullptr, pl.Nj, w, nw);
    mem_sync();
    if (nw > 1) wg_sync();
    WDS_TACC(c, 106);
    if (w == 0 && !primal_solve(pl, c, dH)) fail = -100 - step;
    fail = wave0_int(pl, c, fail, w, nw);
    if (fail) return fail;
    WDS_TACC(c, 107);
    // dq = -P (r_l + dH[j2] - dH[j1]); ACTIVE PRVs from the downstream row
    for (int k0 = lane + 64 * kRowG8 * w; k0 < pl.L; k0 += 64 * kRowG8 * nw) {
      int fv[kRowG8], s1[kRowG8], s2[kRowG8];
      double Pv[kRowG8], Yv[kRowG8];
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) {
        const int k = min(k0 + 64 * u, pl.L - 1);
        fv[u] = pl.l_flags[k];
        s1[u] = pl.l_s1[k];
        s2[u] = pl.l_s2[k];
        Pv[u] = c.P[k];
        Yv[u] = c.Y[k];
      }
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) {
        const int k = k0 + 64 * u;
        if (k >= pl.L) break;
        const double b1 = (fv[u] & LF_J1) ? dH[s1[u]] : 0.0;
        const double b2 = (fv[u] & LF_J2) ? dH[s2[u]] : 0.0;
        c.P[k] = -Pv[u] * (Yv[u] + (b2 - b1));  // P now holds dq
      }
    }
    mem_sync();
    bool any_act = false;
    for (int i = 0; i < pl.p; ++i) any_act |= c.sh[i] != 0.0;
    if (any_act) {
      // An ACTIVE valve's flow closes the linearised mass balance of its
      // downstream junction j2:  m + dq_k - sum_{l != k} dq_l - de = 0, which
      // with dq_l and de from dH gives dq_k = (Abar dH)[j2] - rhs_p[j2]
      // (Abar = the node matrix without the CBIG pin).  dgga/autodiff.py:1084
      // applies rhs_p - Abar dH, the opposite sign: that doubles the j2 mass
      // residual every step instead of removing it; this follows the full
      // Newton step of the CPU polish (dgga/autodiff.py:473-491).
      // (A dH)[j2] of each ACTIVE valve's downstream row only, one lane per
      // valve, its entries in the order of the full product
      if (nw > 1) wg_sync();
      if (w == 0 && lane < pl.p && c.sh[lane] != 0.0) {
        const int4 rt = pl.prv_rt[lane];
        const int* cw = (const int*)pl.rs_ell;
        double sum = 0.0;
        for (int e = 0; e < rt.y; ++e) {
          const int slot = (rt.x + e) * 64 + rt.z;
          sum = fma(c.A[slot], dH[cw[2 * slot + 1] & kColMask], sum);
        }
        const int j2 = pl.prv_s2[lane];
        const double abar2 = sum - CBIG * dH[j2];
        c.P[pl.prv_k[lane]] = abar2 - c.v1[j2];
      }
      mem_sync();
    }
    if (nw > 1) wg_sync();
    for (int k0 = lane + 64 * kRowG * w; k0 < pl.L; k0 += 64 * kRowG * nw) {
      double qv[kRowG], dv[kRowG];
#pragma unroll
      for (int u = 0; u < kRowG; ++u) {
        const int k = min(k0 + 64 * u, pl.L - 1);
        qv[u] = c.q[k];
        dv[u] = c.P[k];
      }
#pragma unroll
      for (int u = 0; u < kRowG; ++u)
        if (k0 + 64 * u < pl.L) c.q[k0 + 64 * u] = qv[u] + dv[u];
    }
    for (int j0 = lane + 64 * kRowG8 * w; j0 < pl.Nj; j0 += 64 * kRowG8 * nw) {
      double kv[kRowG8], ev[kRowG8], fv[kRowG8], hgv[kRowG8];
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) {
        const int j = min(j0 + 64 * u, pl.Nj - 1);
        kv[u] = c.kej[j];
        ev[u] = c.e[j];
        fv[u] = c.F[j];
        hgv[u] = c.hge[j];
      }
#pragma unroll
      for (int u = 0; u < kRowG8; ++u) {
        const int j = j0 + 64 * u;
        if (j >= pl.Nj) break;
        if (kv[u] > 0.0) c.e[j] = ev[u] + (-fv[u] + dH[j]) / hgv[u];
        c.Hn[j] = c.Hn[j] + dH[j];
      }
    }
    mem_sync();
    lds_sync();
    if (nw > 1) wg_sync();
    WDS_TACC(c, 108);
  }
  // terminal state: coefficients, factor kept for the adjoint, residual
  WDS_T0();
  int fail = 0;
  if (w == 0 && chp_on_clamp(pl, c)) fail = kInfoChpClamp;
  fail = wave0_int(pl, c, fail, w, nw);
  if (fail) return fail;
  polish_coeffs(pl, c, w, nw);
  if (nw > 1) wg_sync();
  polish_assemble(pl, c, w, nw);
  if (w == 0 && k3_factor(pl, c)) fail = -99;
  fail = wave0_int(pl, c, fail, w, nw);
  if (fail) return fail;
  if (sp.factor) {
    double* fo = sp.factor + (int64_t)b * pl.nnzL;
    for (int t = lane + 64 * w; t < pl.nnzL; t += 64 * nw) fo[t] = c.V[t];
  }
  resid = polish_residuals(pl, c, false, w, nw);
  WDS_TACC(c, 109);
  return 0;
}

// Node adjoint M lam = r (_adjgpu_node_solve, dgga/autodiff.py:1107-1154):
// r in v1, gQ of the valves in sprv; result in dst (LDS, not the solve
// vector).
// With nw > 1 wavefronts the plain solve shares its copies and residuals;
// the Woodbury solve is wavefront 0's (the caller broadcasts its outcome).
__device__ inline bool node_solve(const DevPlan& pl, Scn& c, double* dst, int n_ref,
                                  int w = 0, int nw = 1) {
  const int lane = lane_id();
  const int p = pl.p;
  bool any_act = false;
  for (int i = 0; i < p; ++i) any_act |= c.sh[i] != 0.0;
  if (!any_act) {
    k4_solve(pl, c, c.v1, dst, n_ref, w, nw);
    return true;
  }
  if (w != 0) return true;
  double* r2 = c.v3;
  double* wk = c.v4;
  const double* gqp = c.sprv;
  vec_copy(r2, c.v1, nullptr, pl.Nj);
  mem_sync();
  // one lane per valve: the downstream junctions are distinct
  if (lane < p && c.sh[lane] != 0.0) r2[pl.prv_s2[lane]] = gqp[lane];
  mem_sync();
  wood_paths(pl, c);
  if (!wood_solve(pl, c, r2, true)) return false;
  for (int j = lane; j < pl.Nj; j += 64) dst[j] = c.R[j];
  lds_sync();
  for (int ref = 0; ref < 2; ++ref) {
    matvec(pl, c, dst, wk);
    // the residual goes straight into the solve vector S
    double* res = c.R;
    vec_copy(res, r2, wk, pl.Nj);
    lds_sync();
    if (lane < p && c.sh[lane] != 0.0) {
      const int j1 = pl.prv_s1[lane], j2 = pl.prv_s2[lane];
      res[j2] = r2[j2] - (dst[j2] - dst[j1]);
    }
    lds_sync();
    if (!wood_solve(pl, c, nullptr, true, true)) return false;
    for (int j = lane; j < pl.Nj; j += 64) dst[j] += c.R[j];
    lds_sync();
  }
  return true;
}

// Backward of the implicit solve (ImplicitGGASolveGPU.backward,
// dgga/autodiff.py:1245-1312), one workgroup per scenario: its wavefronts
// (blockDim.x / 64) share the state loads, the coefficients and assembly,
// the factor copy, the right-hand side, the plain solve's residuals and the
// contractions; the Woodbury solve and the fixed-node sums are wavefront
// 0's.
__device__ __forceinline__ void adjoint_scenario(const DevPlan& pl, const AdjParams& ap, int b,
                                        double* lds) {
  const int lane = lane_id();
  const int w = wave_idx(), nw = (int)(blockDim.x >> 6);
  Scn c = make_scn(pl, lds, ap.ws, b, const_cast<double*>(ap.flow),
                   const_cast<int8_t*>(ap.status), ap.r, ap.r_stride);
  const double* ke = ap.ke + (int64_t)b * ap.ke_stride;
  const double* H = ap.head + (int64_t)b * pl.N;
  const double* E = ap.emitter + (int64_t)b * pl.N;
  // null cotangents are zero (outputs that do not reach the loss)
  const double* gH = ap.gH ? ap.gH + (int64_t)b * pl.N : nullptr;
  const double* gE = ap.gE ? ap.gE + (int64_t)b * pl.N : nullptr;
  const double* gQ = ap.gQ ? ap.gQ + (int64_t)b * pl.L : nullptr;
  auto GH = [&](int n) { return gH ? gH[n] : 0.0; };
  auto GE = [&](int n) { return gE ? gE[n] : 0.0; };
  auto GQ = [&](int k) { return gQ ? gQ[k] : 0.0; };
  const int ns = pl.Nj + pl.Nf;
  for (int s0 = lane + 64 * kRowG8 * w; s0 < ns; s0 += 64 * kRowG8 * nw) {
    int nd[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) nd[u] = pl.slot_node[min(s0 + 64 * u, ns - 1)];
    double hv[kRowG8], kv[kRowG8], ev[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      hv[u] = H[nd[u]];
      kv[u] = ke[nd[u]];
      ev[u] = E[nd[u]];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int sl = s0 + 64 * u;
      if (sl >= ns) break;
      c.Hn[sl] = hv[u];
      if (sl < pl.Nj) {
        c.kej[sl] = kv[u];
        c.e[sl] = ev[u];
      }
    }
  }
  mem_sync();
  lds_sync();
  if (nw > 1) wg_sync();
  // coefficients at the converged state; assemble A-hat (values for the
  // refinement mat-vec) and load the forward's terminal factor into V
  polish_coeffs(pl, c, w, nw);
  if (nw > 1) wg_sync();
  polish_assemble(pl, c, w, nw);
  if (nw > 1) wg_sync();
  const double* fac = ap.factor + (int64_t)b * pl.nnzL;
  for (int t = lane + 64 * w; t < pl.nnzL; t += 64 * nw) c.V[t] = fac[t];
  if (w == 0 && lane == 0) c.V[pl.nnzL] = 0.0;
  lds_sync();
  // r = B^T P gQ - em gE / hge - gH (n2 side first, reference scatter order);
  // rows in groups of kRowG8 per lane with their loads issued together
  for (int j0 = lane + 64 * kRowG8 * w; j0 < pl.Nj; j0 += 64 * kRowG8 * nw) {
    int nd[kRowG8];
    double kv[kRowG8], hgv[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = min(j0 + 64 * u, pl.Nj - 1);
      nd[u] = pl.slot_node[j];
      kv[u] = c.kej[j];
      hgv[u] = c.hge[j];
    }
    double gev[kRowG8], ghv[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      gev[u] = GE(nd[u]);
      ghv[u] = GH(nd[u]);
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = j0 + 64 * u;
      if (j >= pl.Nj) break;
      double r = 0.0, unused = 0.0;
      if (gQ) {
        const int e0 = pl.jadj_ptr[j], e1 = pl.jadj_ptr[j + 1], em = pl.jadj_mid[j];
        row_gather(pl, em, e1, 1.0, c.P, c.P, gQ, unused, r);
        row_gather(pl, e0, em, -1.0, c.P, c.P, gQ, unused, r);
      } else {
        // a zero cotangent still adds (+-0.0) terms: keep r exactly 0.0
      }
      const double emh = kv[u] > 0.0 ? 1.0 / hgv[u] : 0.0;
      r = r - emh * gev[u];
      r = r - ghv[u];
      c.v1[j] = r;
    }
  }
  if (w == 0 && lane < pl.p) c.sprv[lane] = GQ(pl.prv_k[lane]);
  mem_sync();
  lds_sync();
  if (nw > 1) wg_sync();
  double* lam = c.X2;  // LDS
  const bool solved = node_solve(pl, c, lam, ap.n_refine, w, nw);
  if (!wave0_int(pl, c, solved ? 1 : 0, w, nw)) {
    if (w == 0 && lane == 0) ap.info[b] = -1;
    return;
  }
  // back-substitution and theta contraction
  double* gd = ap.gd + (int64_t)b * pl.N;
  double* gke = ap.gke + (int64_t)b * pl.N;
  double* grh = ap.grh + (int64_t)b * pl.N;
  double* gr = ap.gr + (int64_t)b * pl.L;
  double* laml = c.Y;  // lambda_l per link
  for (int k0 = lane + 64 * kRowG8 * w; k0 < pl.L; k0 += 64 * kRowG8 * nw) {
    int fv[kRowG8], s1[kRowG8], s2[kRowG8];
    int8_t sv[kRowG8];
    double Pv[kRowG8], qv[kRowG8], rv[kRowG8], gq[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int k = min(k0 + 64 * u, pl.L - 1);
      fv[u] = pl.l_flags[k];
      s1[u] = pl.l_s1[k];
      s2[u] = pl.l_s2[k];
      sv[u] = c.S[k];
      Pv[u] = c.P[k];
      qv[u] = c.q[k];
      rv[u] = c.r[k];
      gq[u] = GQ(k);
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int k = k0 + 64 * u;
      if (k >= pl.L) break;
      const int f = fv[u];
      const double l1 = (f & LF_J1) ? lam[s1[u]] : 0.0;
      const double l2 = (f & LF_J2) ? lam[s2[u]] : 0.0;
      const double ll = Pv[u] * (gq[u] - (l2 - l1));
      laml[k] = ll;
      // dr = d hloss / d r (open pipe; H-W unclamped)
      double dr = 0.0;
      const int8_t st = sv[u];
      const bool closed_now = (st <= ST_CLOSED || (f & LF_ZERO_SPEED)) && !(f & LF_PCV);
      if ((f & LF_PIPE) && !closed_now && pl.headloss != 0) {
        dr = dw_dr(pl, qv[u], pl.l_diam[k], pl.l_kc[k]);
      } else if ((f & LF_PIPE) && !closed_now) {
        const double q = qv[u];
        const double sgn = q < 0.0 ? -1.0 : 1.0;
        const double qa = q * sgn;
        const double hgr = pl.hexp * rv[u] * det_pow(fmax(qa, 1e-30), pl.hexp - 1.0);
        if (hgr >= pl.rqtol) dr = det_pow(qa, pl.hexp) * sgn;
      }
      gr[k] = -(ll * dr);
      if (ap.gpump) {
        double dw = 0.0, dh0 = 0.0, drp = 0.0;
        if ((f & LF_PUMP) && !closed_now) pump_dtheta(pl, k, f, qv[u], dw, dh0, drp);
        double* gp = ap.gpump + (int64_t)b * 3 * pl.L + k;
        gp[0] = -(ll * dw);
        gp[pl.L] = -(ll * dh0);
        gp[2 * pl.L] = -(ll * drp);
      }
    }
  }
  for (int j0 = lane + 64 * kRowG8 * w; j0 < pl.Nj; j0 += 64 * kRowG8 * nw) {
    int nd[kRowG8];
    double kv[kRowG8], hgv[kRowG8], ev[kRowG8];
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = min(j0 + 64 * u, pl.Nj - 1);
      nd[u] = pl.slot_node[j];
      kv[u] = c.kej[j];
      hgv[u] = c.hge[j];
      ev[u] = c.e[j];
    }
#pragma unroll
    for (int u = 0; u < kRowG8; ++u) {
      const int j = j0 + 64 * u;
      if (j >= pl.Nj) break;
      const int node = nd[u];
      gd[node] = lam[j];
      double g = 0.0;
      if (kv[u] > 0.0) {
        const double lame = (GE(node) + lam[j]) / hgv[u];
        const double ea = fabs(ev[u]);
        const double pw = det_pow(fmax(ea, 1e-30), pl.qexp - 1.0);
        const double hraw = pl.qexp * fmax(kv[u], CSMALL) * pw;
        const double dke = (hraw < pl.rqtol || kv[u] < CSMALL) ? 0.0 : pw * ev[u];
        g = -(lame * dke);
      }
      gke[node] = g;
      grh[node] = 0.0;
    }
  }
  mem_sync();
  // the fixed-node sums read every link's lambda
  if (nw > 1) wg_sync();
  if (w != 0) return;
  for (int f = lane; f < pl.Nf; f += 64) {
    double g = 0.0;
    for (int e = pl.fadj_ptr[f]; e < pl.fadj_ptr[f + 1]; ++e) {
      const int ent = pl.fadj[e];
      if (ent >= 0) g += laml[ent];
    }
    for (int e = pl.fadj_ptr[f]; e < pl.fadj_ptr[f + 1]; ++e) {
      const int ent = pl.fadj[e];
      if (ent < 0) g += -laml[ent & 0x7fffffff];
    }
    const int node = pl.slot_node[pl.Nj + f];
    grh[node] = g + GH(node);
    gd[node] = 0.0;
    gke[node] = 0.0;
  }
  if (lane == 0) ap.info[b] = 0;
}

// one scenario per workgroup, kSolveWaves wavefronts
__global__ WDS_SOLVE_KERNEL void WDS_K(k_adjoint)(DevPlan pl, AdjParams ap) {
  const int b = blockIdx.x;
  WDS_SCENARIO_MEM(ap.ws, b);
  adjoint_scenario(pl, ap, b, lds);
}

}  // namespace wds
