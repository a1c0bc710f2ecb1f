// Device-side view of a plan: plain pointers into one uploaded blob plus the
// scalars every kernel needs.  Shared by the host builder and the kernels.
#pragma once
#include <cstdint>

namespace wds {

constexpr int kMaxPrv = 8;     // pressure-reducing valves with a setting
constexpr int64_t kMaxLdsBytes = 160 * 1024;  // LDS of one CU: a scenario's layout
                                              // above this runs in arena mode

// per-link integer attributes (bit field in link_flags)
enum LinkFlag : int32_t {
  LF_PIPE = 1 << 0,        // CVPIPE or PIPE
  LF_CV = 1 << 1,
  LF_TCV = 1 << 2,
  LF_PUMP = 1 << 3,
  LF_PRV_FIXED = 1 << 4,   // PRV, PSV or FCV whose setting is MISSING (fixed open/closed)
  LF_PCV = 1 << 5,         // PRV, PSV or FCV with a setting: coefficients from the
                           // valve stage (prvcoeff, psvcoeff, fcvcoeff)
  LF_PUMP_CHP = 1 << 6,    // constant-horsepower pump
  LF_PUMP_LIN = 1 << 7,    // power-function pump with N == 1
  LF_PUMP_PF = 1 << 8,     // power-function pump
  LF_PUMP_NOCURVE = 1 << 9,
  LF_ZERO_SPEED = 1 << 10, // pump with speed 0 (always closed)
  LF_PUMP_LIVE = 1 << 11,  // pump with speed > 0 (pumpstatus applies)
  LF_J1 = 1 << 12,         // n1 is a junction
  LF_J2 = 1 << 13,         // n2 is a junction
  LF_PUMP_CUSTOM = 1 << 14,// multi-point pump curve
};

// LDS layout of a scenario (doubles): V [nnzL + 2] (factor, zero slot, spare),
// heads [Nj + Nf], solve vector S [Nj], second vector [Nj], shared scratch
// [2 kMaxPrv], then one sink slot per lane: the programs' stores that are not
// a task's last go to the lane's own sink, so they never collide in a bank.
#ifdef __HIP__
#define WDS_HD __host__ __device__
#else
#define WDS_HD
#endif

WDS_HD inline int64_t lds_sink_index(int64_t nnzL, int64_t Nj, int64_t Nf) {
  return nnzL + 2 + (Nj + Nf) + 2 * Nj + 2 * kMaxPrv;
}
// Emitter rows per lane that the Newton update reads from a per-scenario list
// (rows r = lane + 64 i with an emitter, in ascending r); a scenario with more
// in some lane walks every row instead.
constexpr int kEmPass = 2;
// fixed-stride valve-end path terms: nodes per path, terms per node
constexpr int kPathMax = 16;
constexpr int kTermMax = 8;
// the list: int32 [kEmPass * 64] slots (-1 past a lane's rows), then the
// wavefront's largest count, padded to whole doubles
constexpr int kEmDoubles = 32 * kEmPass + 2;

// after the sinks: the Woodbury block sized by the plan's np = 2p valve ends
// (valve-end paths y, z [wp_tot] each, capacitance [np^2], the small dense
// system [np^2], its right-hand side [np], probes [np], valve residuals [p]),
// then the emitter row list, then the path nodes' slots (int32 [wp_tot])
WDS_HD inline int64_t lds_small_doubles(int64_t wp_tot, int64_t np) {
  return 2 * wp_tot + 2 * np * np + 2 * np + np / 2 + kEmDoubles + (wp_tot + 1) / 2;
}

// solve programs: the store word of a level's last step carries the level
// end and log2 of the lane's group width above the LDS byte address
constexpr int32_t kRowEndFlag = 1 << 30;   // column word of a row's last entry
constexpr int32_t kColMask = kRowEndFlag - 1;
constexpr int32_t kDotAddrMask = (1 << 27) - 1;   // byte offsets below 128 MiB
constexpr int32_t kDotLevelEnd = 1 << 30;
constexpr int kDotWidthShift = 27;

struct ProgDev {
  const int4* instr;        // step-major instructions, nsteps * 64
  int32_t nsteps;           // steps the executor runs (NOP padded to whole
                            // turns of its prefetch ring; the array holds
                            // that ring's depth in further NOP steps)
  int32_t nlevels;
  int32_t dummy;            // first per-lane sink slot: factor, index from V;
                            // solves, LDS byte address
};

struct DevPlan {
  // sizes
  int32_t N, L, Nj, Nf, nnzL, p, ntl;
  // arena mode: doubles of the per-scenario arena that holds the LDS layout
  // in global memory (at the end of each scenario's workspace slice); 0 in
  // LDS mode
  int64_t arena;
  int32_t headloss;         // 0 Hazen-Williams, 1 Darcy-Weisbach
  double hexp, qexp, rqtol, viscos, htol, qtol;
  int32_t checkfreq, maxcheck, extra_iter;
  int32_t has_cv, has_pump, has_fcv;
  // nodes
  const int32_t* node_slot;   // [N] original node -> slot (junction slots
                              //     0..Nj-1 in elimination order, fixed Nj..)
  const int32_t* slot_node;   // [Nj+Nf]
  const double* el_slot;      // [Nj] junction elevation by slot
  // links
  const int32_t* l_s1;        // [L] slot of n1
  const int32_t* l_s2;        // [L] slot of n2
  const int32_t* l_flags;     // [L]
  const int8_t* l_st0;        // [L] initial internal status
  const double* l_diam;       // [L]
  const double* l_km;         // [L] pipe minor loss 0.02517 K / D^4
  const double* l_km_tcv;     // [L] TCV loss coefficient
  const double* l_km_valve;   // [L] valve minor loss (fixed PRVs)
  const double* l_kc;         // [L] D-W roughness
  const double* l_q0;         // [L] initial flow
  const double* pu_h0;        // [L] pump shutoff-head coefficient
  const double* pu_r;         // [L]
  const double* pu_n;         // [L]
  const double* pu_speed;     // [L]
  const double* pu_hmax;      // [L] pumpstatus cut-off head
  const int32_t* pu_cptr;     // [L+1] CUSTOM curve points of link k: [pu_cptr[k], pu_cptr[k+1])
  const double* pu_cx;        // curve flows (user units, strictly increasing)
  const double* pu_cy;        // curve heads (user units)
  double ucf_flow, ucf_head;
  // junction gather lists, reference assembly order: links with n1 == j
  // (link order) then links with n2 == j; entry = link | (side << 31)
  const int32_t* jadj_ptr;    // [Nj+1]
  const int32_t* jadj;        // [..]
  const int32_t* jadj_mid;    // [Nj] first n2-side entry of each row
  // fixed-head terms of each junction row: (link, fixed-node slot), links
  // with a fixed n1 first, then links with a fixed n2
  const int32_t* jg_ptr;      // [Nj+1]
  const int2* jg;             // [..]
  // fixed-head node incidence (slot Nj+f): links with n1 == f then n2 == f
  const int32_t* fadj_ptr;    // [Nf+1]
  const int32_t* fadj;        // [..] link | (side << 31)
  // off-diagonal A entries: L position and contributing links (segment order)
  int32_t noff;
  const int32_t* off_pos;     // [noff]
  const int32_t* off_ptr;     // [noff+1]
  const int32_t* off_link;    // [..]
  int32_t nfill;
  const int32_t* fill_pos;    // [nfill] positions of L not in the pattern of A
  // symmetric SpMV rows over A (slot order): (L position, column slot)
  const int32_t* spmv_ptr;    // [Nj+1]
  const int2* spmv_ent;       // [..]
  // the symmetric rows in lane sequences (lane l: rows l, l + 64, ..., each
  // row's entries in column order, packed; slot t of lane l is
  // rs_ell[t * 64 + l]; a row's last entry has kRowEndFlag in its column
  // word; pads (0, -1)); rs_share[c * 64 + l] = (t0, t1, first row, rows)
  // of lane l's pass c (0: all its rows; 1, 2: the halves of a pair of
  // wavefronts), rs_len[c] the pass's longest slot range
  const int4* rs_share;       // [3 * 64]
  const int2* rs_ell;
  int32_t rs_len[3];
  // assembly rows in the same blocks: (link | side bit, -1) for the row's
  // links in jadj order, then (link, fixed slot) for its fixed-head terms in
  // jg order; pads (0, -2)
  const int32_t* as_blk;
  const int2* as_ell;
  const int4* off4;           // [noff] (L position, first link, first index into off_link, count)
  // the polish's mass rows: links with n2 == row, then n1 == row (jadj
  // entries) per 64-row block; ps_len[row] entries per row
  const int32_t* ps_blk;
  const int32_t* ps_ell;
  const int32_t* ps_len;
  // A (global, per scenario) lives in the row-block layout of rs_ell: na
  // doubles, off-diagonal values stored twice.  Slots of the diagonals, of
  // each off-diagonal entry (pair), of each PRV's positions (off pair, d1,
  // d2) and of each entry of the reference CSR pattern
  int32_t na;
  const int32_t* diag_ae;     // [Nj]
  const int2* off_ae;         // [noff]
  const int4* prv_ae;         // [p]
  const int4* prv_rt;         // [p] downstream row in the lane layout (t0, entries, lane)
  const int32_t* csr_to_ell;  // [nnz_csr]
  // control valves with a setting (PRV, PSV, FCV; file order)
  const int32_t* prv_k;       // [p]
  const int32_t* prv_kind;    // [p] 0 PRV, 1 PSV, 2 FCV
  const int32_t* prv_s1;      // [p] slot of n1
  const int32_t* prv_s2;      // [p]
  const int32_t* prv_pos_off; // [p] L position of (j1,j2) lower
  const double* prv_hset;     // [p] PRV: El[n2] + setting, PSV: El[n1] + setting,
                              //     FCV: the flow setting (cfs)
  const double* prv_km;       // [p]
  int32_t prv_disjoint;       // 1: no two valves share an end node
  // elimination-tree paths of the valve ends (path a < p: n2 of valve a,
  // path p + a: n1 of valve a).  y_u = L^-1 e_u is nonzero only on the path
  // from u to its root, so unit solves with the factor cost O(height^2).
  int32_t np, wp_tot;
  const int32_t* wp_ptr;      // [np+1] path -> range of path nodes
  const int32_t* wp_node;     // [wp_tot] slot (elimination index) of each path node
  const int32_t* wp_diag;     // [wp_tot] V position of the node's reciprocal pivot
  const int32_t* wt_ptr;      // [wp_tot+1] path node -> range of its terms
  const int2* wt;             // [..] (V position of w(node, m), path-node index of m)
  const int4* wpair;          // [np*np] common suffix of paths u, v: (start u, start v, length)
  // the terms at a fixed stride (wq_ok: every path has at most kPathMax
  // nodes and every node at most kTermMax terms): entry k of path u at
  // wq[(u * kPathMax + k) * kTermMax + m], wq_n[u * kPathMax + k] terms
  int32_t wq_ok;
  const int2* wq;
  const int32_t* wq_n;
  // tank links (tankstatus)
  const int32_t* tl_k;        // [ntl]
  const int32_t* tl_tank;     // [ntl] slot of the tank
  const int32_t* tl_other;    // [ntl] slot of the other end
  const int32_t* tl_bits;     // [ntl] bit0 use_n2 (qsign -1), bit1 ovf, bit2 is_pump, bit3 pump_in, bit4 pump_out
  const double* tl_hmax;      // [ntl]
  const double* tl_hmin;      // [ntl]
  // reference CSR (full symmetric pattern, original junction order) -> L
  int32_t nnz_csr;
  const int32_t* csr_to_L;    // [nnz_csr] L position for lower-in-elimination-order entries, else -1
  const int32_t* l_scale;     // [nnzL] L position of the column's pivot, -1 on diagonals
  const int32_t* diag_pos;    // [Nj] L position of slot j's diagonal (colptr)
  const int32_t* row_slot;    // [Nj] original junction row -> slot
  // programs
  ProgDev factor, fwd, bwd;
};

}  // namespace wds
