// Host-side plan object behind the opaque wds_plan handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wds.h"
#include "devplan.h"
#include "symbolic.h"

struct wds_plan {
  wds::Symbolic sym;
  int32_t N = 0, L = 0, Nj = 0, Nf = 0, p = 0;
  // host copies of everything that goes to the device
  std::vector<int32_t> node_slot, slot_node, l_s1, l_s2, l_flags;
  std::vector<int8_t> l_st0;
  std::vector<double> el_slot, l_diam, l_km, l_km_tcv, l_km_valve, l_kc, l_q0,
      pu_h0, pu_r, pu_n, pu_speed, pu_hmax, pu_cx, pu_cy;
  std::vector<int32_t> pu_cptr;
  double ucf_flow = 1.0, ucf_head = 1.0;
  std::vector<int32_t> jadj_ptr, jadj, fadj_ptr, fadj, off_pos, off_ptr, off_link, fill_pos,
      spmv_ptr, prv_k, prv_s1, prv_s2, prv_pos_off, prv_kind, tl_k, tl_tank, tl_other,
      tl_bits, csr_indptr, csr_indices, csr_to_L, diag_pos, row_slot;
  std::vector<int32_t> spmv_ent;  // pairs (pos, col)
  // the same rows in lane sequences: lane l walks rows l, l + 64, ... in
  // order, each row's entries in column order, packed; slot t of lane l is
  // pair t * 64 + l, the column word of a row's last entry carries
  // kRowEndFlag, pads are (0, -1).  rs_share: per lane, the slot range and
  // first row of a pass over all its rows (entries 0..63) and of the two
  // halves a pair of wavefronts takes (64..127, 128..191): int4 (t0, t1,
  // first row, rows); rs_len: the longest range of each of those three
  // passes.  rs_t0: the first slot of each row (host only).
  std::vector<int32_t> rs_share, rs_ell, rs_t0;
  int32_t rs_len[3] = {0, 0, 0};
  // assembly rows in the same blocks: a row's links in jadj order (link |
  // side bit, -1), then its fixed-head terms in jg order (link, fixed slot);
  // pads (0, -2).  off4: per off-diagonal entry (L position, first link,
  // first index into off_link, number of links)
  std::vector<int32_t> as_blk, as_ell, off4;
  // the polish's mass rows in the same blocks: a row's links with n2 == row
  // first, then those with n1 == row (jadj entries, side bit kept); pads 0
  // with ps_len giving each row's entry count
  std::vector<int32_t> ps_blk, ps_ell, ps_len;
  // the assembled matrix A is stored in the row-block layout of rs_ell (each
  // off-diagonal value twice): ae_pos[t] the one or two slots of L position
  // t (-1 none), diag_ae / off_ae / prv_ae those of the diagonals, the
  // off-diagonal entries and the PRV positions (off pair, d1, d2), csr_to_ell
  // the slot of each entry of the reference CSR pattern
  std::vector<int32_t> ae_pos, diag_ae, off_ae, prv_ae, csr_to_ell;
  std::vector<int32_t> prv_rt;  // per valve: its downstream row (t0, entries, lane)
  std::vector<int32_t> wp_ptr, wp_node, wp_diag, wt_ptr, wt, wpair;
  // the same terms at a fixed stride: entry k of path u holds its terms at
  // pairs ((u * kPathMax + k) * kTermMax + m), wq_n[u * kPathMax + k] of them;
  // empty unless every path fits (wq_ok)
  std::vector<int32_t> wq, wq_n;
  int32_t wq_ok = 0;
  std::vector<int32_t> jg_ptr, jg;  // jg: pairs (link, fixed slot)
  std::vector<int32_t> l_scale, jadj_mid;
  std::vector<double> prv_hset, prv_km, tl_hmax, tl_hmin;
  // scalars
  int32_t headloss = 0, checkfreq = 2, maxcheck = 10, extra_iter = -1;
  int32_t has_cv = 0, has_pump = 0;
  int32_t prv_disjoint = 0;  // no two valves with a setting share an end node
  int32_t has_psv_fcv = 0;   // a PSV or FCV with a setting (no polish / adjoint)
  int32_t has_fcv = 0;       // an FCV with a setting (fcvstatus in linkstatus)
  double hexp = 1.852, qexp = 2.0, rqtol = 1e-7, viscos = 1.1e-5, htol = 5e-4, qtol = 1e-4;
  // arena mode (doubles per scenario of the global-memory copy of the LDS
  // layout; 0 in LDS mode)
  int64_t arena = 0;
  // device state
  wds::DevPlan dev{};
  void* d_blob = nullptr;
  int device = -1;
  wds_counters cnt{};
  wds_plan_stats stats{};
};

namespace wds {
void set_error(const std::string& msg);
}
