// Plan construction: junction numbering, elimination order, gather lists in
// the reference's assembly order, PRV/tank tables and the reference CSR map.
//
// Assembly order (must match the reference's scatter_add source order so the
// assembled values are reproducible, dgga/solver.py:283-334, 2958-3017):
//   off-diagonal (lo,hi): links whose off1 slot is (lo,hi) in link order, then
//     links whose off2 slot is (lo,hi); pairs of >=3 links in mixed directions
//     put every link in the first segment (the (min,max) rule, 316-325);
//   diagonal of junction j: links with n1 == j, then links with n2 == j.
#include <algorithm>
#include <cstring>
#include <map>
#include <stdexcept>

#include "plan.h"

namespace wds {

thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static constexpr double kMissing = -1.0e10;

static void build(const wds_network_desc* d, int ordering, wds_plan* P) {
  const int N = d->N, L = d->L;
  if (N <= 0 || L <= 0) throw std::invalid_argument("empty network");
  P->N = N;
  P->L = L;
  std::vector<int32_t> junc_row(N, -1), fixed_idx(N, -1);
  int Nj = 0, Nf = 0;
  for (int i = 0; i < N; ++i) {
    if (d->node_type[i] == 0)
      junc_row[i] = Nj++;
    else
      fixed_idx[i] = Nf++;
  }
  if (Nj == 0) throw std::invalid_argument("network has no junctions");
  P->Nj = Nj;
  P->Nf = Nf;
  // supported link types: CVPIPE 0, PIPE 1, PUMP 2, PRV 3, PSV 4, FCV 6, TCV 7
  for (int k = 0; k < L; ++k) {
    const int t = d->link_type[k];
    if (!(t == 0 || t == 1 || t == 2 || t == 3 || t == 4 || t == 6 || t == 7))
      throw std::domain_error("link type " + std::to_string(t) + " is not on the GGA dense path");
    if (d->link_n1[k] == d->link_n2[k]) throw std::invalid_argument("self-loop link");
    if (d->link_n1[k] < 0 || d->link_n1[k] >= N || d->link_n2[k] < 0 || d->link_n2[k] >= N)
      throw std::invalid_argument("link end out of range");
  }
  // ---- symbolic analysis on the junction graph ----
  std::vector<std::vector<int32_t>> adj(Nj);
  for (int k = 0; k < L; ++k) {
    const int a = junc_row[d->link_n1[k]], b = junc_row[d->link_n2[k]];
    if (a >= 0 && b >= 0) adj[a].push_back(b);
  }
  analyse(Nj, adj, ordering & 0xff, P->sym);
  const Symbolic& S = P->sym;
  auto slot_of_row = [&](int r) { return S.iperm[r]; };
  P->node_slot.assign(N, -1);
  P->slot_node.assign(Nj + Nf, -1);
  for (int i = 0; i < N; ++i) {
    const int s = junc_row[i] >= 0 ? slot_of_row(junc_row[i]) : Nj + fixed_idx[i];
    P->node_slot[i] = s;
    P->slot_node[s] = i;
  }
  P->el_slot.assign(Nj, 0.0);
  for (int i = 0; i < N; ++i)
    if (junc_row[i] >= 0) P->el_slot[P->node_slot[i]] = d->elev[i];
  // ---- per-link tables ----
  P->l_s1.resize(L);
  P->l_s2.resize(L);
  P->l_flags.assign(L, 0);
  P->l_st0.resize(L);
  P->l_diam.assign(d->diam, d->diam + L);
  P->l_km.assign(d->km, d->km + L);
  P->l_km_tcv.assign(d->km_tcv, d->km_tcv + L);
  P->l_km_valve.assign(d->km_valve, d->km_valve + L);
  P->l_kc.assign(d->kc, d->kc + L);
  P->l_q0.assign(d->q0, d->q0 + L);
  P->pu_h0.assign(L, 0.0);
  P->pu_r.assign(L, 0.0);
  P->pu_n.assign(L, 1.0);
  P->pu_speed.assign(L, 0.0);
  P->pu_hmax.assign(L, 0.0);
  P->pu_cptr.assign(L + 1, 0);
  P->pu_cx.clear();
  P->pu_cy.clear();
  P->ucf_flow = d->ucf_flow;
  P->ucf_head = d->ucf_head;
  for (int k = 0; k < L; ++k) {
    const int n1 = d->link_n1[k], n2 = d->link_n2[k], t = d->link_type[k];
    P->l_s1[k] = P->node_slot[n1];
    P->l_s2[k] = P->node_slot[n2];
    P->l_st0[k] = (int8_t)d->init_status[k];
    int32_t f = 0;
    if (t <= 1) f |= LF_PIPE;
    if (t == 0) {
      f |= LF_CV;
      P->has_cv = 1;
    }
    if (t == 7) f |= LF_TCV;
    if (t == 3 || t == 4 || t == 6) f |= (d->setting[k] == kMissing) ? LF_PRV_FIXED : LF_PCV;
    if (t == 2) {
      P->has_pump = 1;
      f |= LF_PUMP;
      const int kind = d->pump_kind[k];
      const double sp = d->setting[k];
      if (kind == 0) f |= LF_PUMP_CHP;
      else if (kind == 1) f |= LF_PUMP_PF;
      else if (kind == 2) f |= LF_PUMP_CUSTOM;
      else if (kind == 3) f |= LF_PUMP_NOCURVE;
      else throw std::domain_error("pump curve kind " + std::to_string(kind) + " not supported");
      if (kind == 2) {
        // the segment search of the batched path equals curvecoeff's linear
        // scan only for strictly increasing flows (dgga/solver.py:989-996)
        if (!d->pump_curve_ptr || !d->pump_curve_x || !d->pump_curve_y)
          throw std::invalid_argument("CUSTOM pump without curve points");
        const int a = d->pump_curve_ptr[k], b = d->pump_curve_ptr[k + 1];
        if (b - a < 2) throw std::invalid_argument("CUSTOM pump curve with fewer than 2 points");
        if (!(d->ucf_flow > 0.0) || !(d->ucf_head > 0.0))
          throw std::invalid_argument("CUSTOM pump curves need ucf_flow and ucf_head > 0");
        for (int i = a + 1; i < b; ++i)
          if (!(d->pump_curve_x[i] > d->pump_curve_x[i - 1]))
            throw std::domain_error("CUSTOM pump curve flows are not strictly increasing");
        for (int i = a; i < b; ++i) {
          P->pu_cx.push_back(d->pump_curve_x[i]);
          P->pu_cy.push_back(d->pump_curve_y[i]);
        }
      }
      if (kind == 1 && d->pump_n[k] == 1.0) f |= LF_PUMP_LIN;
      if (sp == 0.0) f |= LF_ZERO_SPEED;
      if (sp > 0.0) f |= LF_PUMP_LIVE;
      P->pu_h0[k] = d->pump_h0[k];
      P->pu_r[k] = d->pump_r[k];
      P->pu_n[k] = d->pump_n[k];
      P->pu_speed[k] = sp;
      P->pu_hmax[k] = d->pump_hmax[k];
    }
    P->pu_cptr[k + 1] = (int32_t)P->pu_cx.size();
    if (junc_row[n1] >= 0) f |= LF_J1;
    if (junc_row[n2] >= 0) f |= LF_J2;
    P->l_flags[k] = f;
  }
  // ---- junction gather lists ----
  std::vector<std::vector<int32_t>> ja(Nj);
  for (int k = 0; k < L; ++k)
    if (junc_row[d->link_n1[k]] >= 0) ja[P->node_slot[d->link_n1[k]]].push_back(k);
  for (int k = 0; k < L; ++k)
    if (junc_row[d->link_n2[k]] >= 0)
      ja[P->node_slot[d->link_n2[k]]].push_back(k | (int32_t)(1u << 31));
  P->jadj_ptr.assign(Nj + 1, 0);
  for (int s = 0; s < Nj; ++s) P->jadj_ptr[s + 1] = P->jadj_ptr[s] + (int32_t)ja[s].size();
  P->jadj.clear();
  for (int s = 0; s < Nj; ++s) P->jadj.insert(P->jadj.end(), ja[s].begin(), ja[s].end());
  // first n2-side entry of each row (n1-side entries come first)
  P->jadj_mid.assign(Nj, 0);
  for (int s = 0; s < Nj; ++s) {
    int32_t e = P->jadj_ptr[s];
    while (e < P->jadj_ptr[s + 1] && P->jadj[e] >= 0) ++e;
    P->jadj_mid[s] = e;
  }
  // fixed-head terms of each row's right-hand side, in the order the
  // reference adds them: links whose n1 is fixed (n2 == j), then links whose
  // n2 is fixed (n1 == j); entries (link, fixed-node slot)
  P->jg_ptr.assign(Nj + 1, 0);
  P->jg.clear();
  for (int s = 0; s < Nj; ++s) {
    for (int32_t ent : ja[s])
      if (ent < 0) {
        const int k = ent & 0x7fffffff;
        if (junc_row[d->link_n1[k]] < 0) {
          P->jg.push_back(k);
          P->jg.push_back(P->node_slot[d->link_n1[k]]);
        }
      }
    for (int32_t ent : ja[s])
      if (ent >= 0 && junc_row[d->link_n2[ent]] < 0) {
        P->jg.push_back(ent);
        P->jg.push_back(P->node_slot[d->link_n2[ent]]);
      }
    P->jg_ptr[s + 1] = (int32_t)(P->jg.size() / 2);
  }
  std::vector<std::vector<int32_t>> fa(Nf);
  for (int k = 0; k < L; ++k)
    if (fixed_idx[d->link_n1[k]] >= 0) fa[fixed_idx[d->link_n1[k]]].push_back(k);
  for (int k = 0; k < L; ++k)
    if (fixed_idx[d->link_n2[k]] >= 0)
      fa[fixed_idx[d->link_n2[k]]].push_back(k | (int32_t)(1u << 31));
  P->fadj_ptr.assign(Nf + 1, 0);
  for (int f = 0; f < Nf; ++f) P->fadj_ptr[f + 1] = P->fadj_ptr[f] + (int32_t)fa[f].size();
  P->fadj.clear();
  for (int f = 0; f < Nf; ++f) P->fadj.insert(P->fadj.end(), fa[f].begin(), fa[f].end());
  // ---- off-diagonal entries: unordered junction pairs ----
  std::map<std::pair<int, int>, std::vector<int>> pairs;  // (lo,hi) rows -> links
  for (int k = 0; k < L; ++k) {
    const int a = junc_row[d->link_n1[k]], b = junc_row[d->link_n2[k]];
    if (a < 0 || b < 0) continue;
    pairs[{std::min(a, b), std::max(a, b)}].push_back(k);
  }
  P->row_slot.resize(Nj);
  for (int r = 0; r < Nj; ++r) P->row_slot[r] = slot_of_row(r);
  P->diag_pos.resize(Nj);
  for (int s = 0; s < Nj; ++s) P->diag_pos[s] = S.colptr[s];
  std::vector<char> is_a(S.nnzL, 0);
  for (int s = 0; s < Nj; ++s) is_a[S.colptr[s]] = 1;
  P->off_pos.clear();
  P->off_ptr.assign(1, 0);
  P->off_link.clear();
  std::vector<std::vector<std::pair<int32_t, int32_t>>> rows(Nj);
  for (auto& kv : pairs) {
    const int lo = kv.first.first, hi = kv.first.second;
    const auto& ks = kv.second;
    int nfwd = 0;
    for (int k : ks) nfwd += (junc_row[d->link_n1[k]] == lo);
    const bool trig = ks.size() >= 3 && nfwd != 0 && nfwd != (int)ks.size();
    std::vector<int32_t> seq;
    for (int k : ks)
      if (trig || junc_row[d->link_n1[k]] == lo) seq.push_back(k);
    if (!trig)
      for (int k : ks)
        if (junc_row[d->link_n1[k]] == hi) seq.push_back(k);
    const int sa = slot_of_row(lo), sb = slot_of_row(hi);
    const int64_t pos = S.pos(std::max(sa, sb), std::min(sa, sb));
    if (pos < 0) throw std::logic_error("pattern of A not contained in L");
    P->off_pos.push_back((int32_t)pos);
    P->off_link.insert(P->off_link.end(), seq.begin(), seq.end());
    P->off_ptr.push_back((int32_t)P->off_link.size());
    is_a[pos] = 1;
    rows[sa].push_back({(int32_t)pos, sb});
    rows[sb].push_back({(int32_t)pos, sa});
  }
  P->fill_pos.clear();
  for (int64_t t = 0; t < S.nnzL; ++t)
    if (!is_a[t]) P->fill_pos.push_back((int32_t)t);
  P->spmv_ptr.assign(Nj + 1, 0);
  P->spmv_ent.clear();
  for (int s = 0; s < Nj; ++s) {
    auto& r = rows[s];
    r.push_back({S.colptr[s], s});
    std::sort(r.begin(), r.end(), [](const std::pair<int32_t, int32_t>& x,
                                     const std::pair<int32_t, int32_t>& y) {
      return x.second < y.second;
    });
    for (auto& e : r) {
      P->spmv_ent.push_back(e.first);
      P->spmv_ent.push_back(e.second);
    }
    P->spmv_ptr[s + 1] = P->spmv_ptr[s] + (int32_t)r.size();
  }
  {
    // lane sequences: lane l's rows l, l + 64, ... packed in order
    std::vector<int32_t> tl(kLanes, 0);
    P->rs_t0.assign(Nj, 0);
    for (int s = 0; s < Nj; ++s) {
      P->rs_t0[s] = tl[s % kLanes];
      tl[s % kLanes] += P->spmv_ptr[s + 1] - P->spmv_ptr[s];
    }
    const int32_t T = *std::max_element(tl.begin(), tl.end());
    P->rs_ell.assign((size_t)std::max<int32_t>(T, 1) * kLanes * 2, 0);
    for (size_t i = 1; i < P->rs_ell.size(); i += 2) P->rs_ell[i] = -1;
    for (int s = 0; s < Nj; ++s) {
      const int lane = s % kLanes;
      const int32_t len = P->spmv_ptr[s + 1] - P->spmv_ptr[s];
      for (int32_t e = 0; e < len; ++e) {
        const size_t x = ((size_t)(P->rs_t0[s] + e) * kLanes + lane) * 2;
        const int32_t col = P->spmv_ent[2 * (size_t)(P->spmv_ptr[s] + e) + 1];
        P->rs_ell[x] = P->spmv_ent[2 * (size_t)(P->spmv_ptr[s] + e)];
        P->rs_ell[x + 1] = col | (e == len - 1 ? kRowEndFlag : 0);
      }
    }
    // passes: all of a lane's rows, and two halves split at the row boundary
    // nearest the middle of the lane's slots
    P->rs_share.assign(3 * kLanes * 4, 0);
    for (int i = 0; i < 3; ++i) P->rs_len[i] = 0;
    auto put = [&](int c, int lane, int32_t t0, int32_t t1, int32_t j0, int32_t rows) {
      int32_t* q = &P->rs_share[4 * ((size_t)c * kLanes + lane)];
      q[0] = t0;
      q[1] = t1;
      q[2] = j0;
      q[3] = rows;
      P->rs_len[c] = std::max(P->rs_len[c], t1 - t0);
    };
    for (int lane = 0; lane < kLanes; ++lane) {
      const int32_t nrows = lane < Nj ? (Nj - 1 - lane) / kLanes + 1 : 0;
      put(0, lane, 0, tl[lane], lane, nrows);
      int32_t k = 0, best = tl[lane];  // rows in the first half, its imbalance
      for (int32_t r = 0; r <= nrows; ++r) {
        const int32_t t = r < nrows ? P->rs_t0[lane + kLanes * r] : tl[lane];
        const int32_t imb = std::abs(2 * t - tl[lane]);
        if (imb < best) {
          best = imb;
          k = r;
        }
      }
      const int32_t tk = k < nrows ? P->rs_t0[lane + kLanes * k] : tl[lane];
      put(1, lane, 0, tk, lane, k);
      put(2, lane, tk, tl[lane], lane + kLanes * k, nrows - k);
    }
  }
  {
    const int nb = (Nj + kLanes - 1) / kLanes;
    auto row_len = [&](int s) {
      return (P->jadj_ptr[s + 1] - P->jadj_ptr[s]) + (P->jg_ptr[s + 1] - P->jg_ptr[s]);
    };
    P->as_blk.assign(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
      int32_t k = 0;
      for (int s = b * kLanes; s < std::min(Nj, (b + 1) * kLanes); ++s) k = std::max(k, row_len(s));
      P->as_blk[b + 1] = P->as_blk[b] + k;
    }
    P->as_ell.assign((size_t)P->as_blk[nb] * kLanes * 2, 0);
    for (size_t i = 1; i < P->as_ell.size(); i += 2) P->as_ell[i] = -2;
    for (int s = 0; s < Nj; ++s) {
      const int b = s / kLanes, lane = s % kLanes;
      int32_t e = 0;
      auto put = [&](int32_t x, int32_t y) {
        const size_t o = ((size_t)(P->as_blk[b] + e) * kLanes + lane) * 2;
        P->as_ell[o] = x;
        P->as_ell[o + 1] = y;
        ++e;
      };
      for (int32_t i = P->jadj_ptr[s]; i < P->jadj_ptr[s + 1]; ++i) put(P->jadj[i], -1);
      for (int32_t g = P->jg_ptr[s]; g < P->jg_ptr[s + 1]; ++g) put(P->jg[2 * g], P->jg[2 * g + 1]);
    }
    P->ps_blk.assign(nb + 1, 0);
    P->ps_len.assign(Nj, 0);
    for (int s = 0; s < Nj; ++s) P->ps_len[s] = P->jadj_ptr[s + 1] - P->jadj_ptr[s];
    for (int b = 0; b < nb; ++b) {
      int32_t k = 0;
      for (int s = b * kLanes; s < std::min(Nj, (b + 1) * kLanes); ++s) k = std::max(k, P->ps_len[s]);
      P->ps_blk[b + 1] = P->ps_blk[b] + k;
    }
    P->ps_ell.assign((size_t)P->ps_blk[nb] * kLanes, 0);
    for (int s = 0; s < Nj; ++s) {
      const int b = s / kLanes, lane = s % kLanes;
      int32_t e = 0;
      auto put = [&](int32_t x) {
        P->ps_ell[(size_t)(P->ps_blk[b] + e) * kLanes + lane] = x;
        ++e;
      };
      for (int32_t i = P->jadj_mid[s]; i < P->jadj_ptr[s + 1]; ++i) put(P->jadj[i]);
      for (int32_t i = P->jadj_ptr[s]; i < P->jadj_mid[s]; ++i) put(P->jadj[i]);
    }
    const size_t noff = P->off_pos.size();
    P->off4.assign(noff * 4, 0);
    for (size_t o = 0; o < noff; ++o) {
      const int32_t b0 = P->off_ptr[o], n = P->off_ptr[o + 1] - b0;
      P->off4[4 * o] = P->off_pos[o];
      P->off4[4 * o + 1] = n > 0 ? P->off_link[b0] : 0;
      P->off4[4 * o + 2] = b0;
      P->off4[4 * o + 3] = n;
    }
  }

  // ---- control valves with a setting (PRV, PSV, FCV), in valve order ----
  P->prv_k.clear();
  for (int k = 0; k < L; ++k) {
    if (!(P->l_flags[k] & LF_PCV)) continue;
    const int n1 = d->link_n1[k], n2 = d->link_n2[k];
    if (junc_row[n1] < 0 || junc_row[n2] < 0)
      throw std::domain_error(
          "control valve with a fixed-head end (the batched valve path needs junction ends)");
    const int s1 = P->node_slot[n1], s2 = P->node_slot[n2];
    const int t = d->link_type[k];
    P->prv_k.push_back(k);
    P->prv_kind.push_back(t == 3 ? 0 : (t == 4 ? 1 : 2));
    P->prv_s1.push_back(s1);
    P->prv_s2.push_back(s2);
    P->prv_pos_off.push_back((int32_t)S.pos(std::max(s1, s2), std::min(s1, s2)));
    // PRV: downstream head El[n2] + setting; PSV: upstream head El[n1] +
    // setting (prvcoeff, psvcoeff); FCV: the flow setting (fcvcoeff)
    P->prv_hset.push_back(t == 3 ? d->elev[n2] + d->setting[k]
                                 : (t == 4 ? d->elev[n1] + d->setting[k] : d->setting[k]));
    P->prv_km.push_back(d->km_valve[k]);
    if (t != 3) P->has_psv_fcv = 1;
    if (t == 6) P->has_fcv = 1;
  }
  P->p = (int32_t)P->prv_k.size();
  {
    std::vector<int32_t> ends(P->prv_s1);
    ends.insert(ends.end(), P->prv_s2.begin(), P->prv_s2.end());
    std::sort(ends.begin(), ends.end());
    P->prv_disjoint = std::adjacent_find(ends.begin(), ends.end()) == ends.end() ? 1 : 0;
  }
  if (P->p > kMaxPrv)
    throw std::domain_error("more than " + std::to_string(kMaxPrv) + " PRVs with a setting");
  // ---- slots of A in the row-block layout ----
  {
    const size_t nell = P->rs_ell.size() / 2;
    P->ae_pos.assign(2 * (size_t)S.nnzL, -1);
    for (size_t x = 0; x < nell; ++x) {
      const int32_t pos = P->rs_ell[2 * x], col = P->rs_ell[2 * x + 1];
      if (col < 0) continue;
      int32_t* a = &P->ae_pos[2 * (size_t)pos];
      if (a[0] < 0) a[0] = (int32_t)x;
      else if (a[1] < 0) a[1] = (int32_t)x;
      else throw std::logic_error("an entry of A in more than two rows");
    }
    P->diag_ae.resize(Nj);
    for (int s = 0; s < Nj; ++s) P->diag_ae[s] = P->ae_pos[2 * (size_t)P->diag_pos[s]];
    P->off_ae.resize(2 * P->off_pos.size());
    for (size_t o = 0; o < P->off_pos.size(); ++o) {
      P->off_ae[2 * o] = P->ae_pos[2 * (size_t)P->off_pos[o]];
      P->off_ae[2 * o + 1] = P->ae_pos[2 * (size_t)P->off_pos[o] + 1];
    }
    P->prv_ae.assign(4 * (size_t)P->p, -1);
    for (int i = 0; i < P->p; ++i) {
      P->prv_ae[4 * i] = P->ae_pos[2 * (size_t)P->prv_pos_off[i]];
      P->prv_ae[4 * i + 1] = P->ae_pos[2 * (size_t)P->prv_pos_off[i] + 1];
      P->prv_ae[4 * i + 2] = P->diag_ae[P->prv_s1[i]];
      P->prv_ae[4 * i + 3] = P->diag_ae[P->prv_s2[i]];
    }
    // the downstream row of each valve in the lane layout: (first slot,
    // entries, lane)
    P->prv_rt.assign(4 * (size_t)P->p, 0);
    for (int i = 0; i < P->p; ++i) {
      const int32_t j2 = P->prv_s2[i];
      P->prv_rt[4 * i] = P->rs_t0[j2];
      P->prv_rt[4 * i + 1] = P->spmv_ptr[j2 + 1] - P->spmv_ptr[j2];
      P->prv_rt[4 * i + 2] = j2 % kLanes;
    }
  }
  // ---- elimination-tree paths of the valve ends ----
  // The Woodbury corrections of the polish and adjoint need A-hat^-1 e_u only
  // through y_u = L^-1 e_u, whose nonzeros lie on the etree path from u to its
  // root: store each path, the pattern entries L(k_i, k_m) between its nodes,
  // and the common suffix (shared ancestors) of every pair of paths.
  {
    std::vector<int32_t> ends;
    for (int a = 0; a < P->p; ++a) ends.push_back(P->prv_s2[a]);
    for (int a = 0; a < P->p; ++a) ends.push_back(P->prv_s1[a]);
    P->wp_ptr.assign(1, 0);
    P->wp_node.clear();
    P->wp_diag.clear();
    P->wt_ptr.assign(1, 0);
    P->wt.clear();
    std::vector<std::vector<int32_t>> paths;
    for (int32_t u : ends) {
      std::vector<int32_t> path;
      for (int32_t k = u; k >= 0; k = S.parent[k]) path.push_back(k);
      const int32_t base = (int32_t)P->wp_node.size();
      for (size_t i = 0; i < path.size(); ++i) {
        P->wp_node.push_back(path[i]);
        P->wp_diag.push_back(S.colptr[path[i]]);
        for (size_t m = 0; m < i; ++m) {
          const int64_t ps = S.pos(path[i], path[m]);
          if (ps < 0) continue;
          P->wt.push_back((int32_t)ps);
          P->wt.push_back(base + (int32_t)m);
        }
        P->wt_ptr.push_back((int32_t)(P->wt.size() / 2));
      }
      P->wp_ptr.push_back((int32_t)P->wp_node.size());
      paths.push_back(std::move(path));
    }
    const int np = (int)ends.size();
    // the fixed-stride copy of the terms, when every path fits it
    P->wq_ok = 1;
    for (int u = 0; u < np; ++u) {
      if (P->wp_ptr[u + 1] - P->wp_ptr[u] > kPathMax) P->wq_ok = 0;
      for (int32_t i = P->wp_ptr[u]; i < P->wp_ptr[u + 1]; ++i)
        if (P->wt_ptr[i + 1] - P->wt_ptr[i] > kTermMax) P->wq_ok = 0;
    }
    P->wq.assign(P->wq_ok ? (size_t)np * kPathMax * kTermMax * 2 : 0, 0);
    P->wq_n.assign(P->wq_ok ? (size_t)np * kPathMax : 0, 0);
    if (P->wq_ok)
      for (int u = 0; u < np; ++u)
        for (int32_t i = P->wp_ptr[u]; i < P->wp_ptr[u + 1]; ++i) {
          const int32_t k = i - P->wp_ptr[u];
          P->wq_n[(size_t)u * kPathMax + k] = P->wt_ptr[i + 1] - P->wt_ptr[i];
          for (int32_t t = P->wt_ptr[i]; t < P->wt_ptr[i + 1]; ++t) {
            const size_t x = (((size_t)u * kPathMax + k) * kTermMax + (t - P->wt_ptr[i])) * 2;
            P->wq[x] = P->wt[2 * (size_t)t];
            P->wq[x + 1] = P->wt[2 * (size_t)t + 1];
          }
        }
    P->wpair.clear();
    for (int u = 0; u < np; ++u)
      for (int v = 0; v < np; ++v) {
        const auto& a = paths[u];
        const auto& b = paths[v];
        size_t n = 0;
        while (n < a.size() && n < b.size() && a[a.size() - 1 - n] == b[b.size() - 1 - n]) ++n;
        P->wpair.push_back(P->wp_ptr[u] + (int32_t)(a.size() - n));
        P->wpair.push_back(P->wp_ptr[v] + (int32_t)(b.size() - n));
        P->wpair.push_back((int32_t)n);
        P->wpair.push_back(0);
      }
  }
  // ---- tank links (dgga/solver.py:536-568) ----
  std::vector<char> is_tank(N, 0), ovf(N, 0);
  std::vector<double> thmin(N, 0.0), thmax(N, 0.0);
  for (int t = 0; t < d->ntank; ++t) {
    const int n = d->tank_node[t];
    is_tank[n] = 1;
    thmin[n] = d->tank_hmin[t];
    thmax[n] = d->tank_hmax[t];
    ovf[n] = d->tank_overflow[t] != 0;
  }
  for (int k = 0; k < L; ++k) {
    const int n1 = d->link_n1[k], n2 = d->link_n2[k];
    const bool f1 = d->node_type[n1] != 0, f2 = d->node_type[n2] != 0;
    const bool use_n2 = !f1 && f2;
    const int cand = f1 ? n1 : (f2 ? n2 : -1);
    if (cand < 0 || !is_tank[cand]) continue;
    const bool is_pump = d->link_type[k] == 2;
    int32_t bits = (use_n2 ? 1 : 0) | (ovf[cand] ? 2 : 0) | (is_pump ? 4 : 0) |
                   ((is_pump && n2 == cand) ? 8 : 0) | ((is_pump && n1 == cand) ? 16 : 0);
    P->tl_k.push_back(k);
    P->tl_tank.push_back(P->node_slot[cand]);
    P->tl_other.push_back(P->node_slot[use_n2 ? n1 : n2]);
    P->tl_bits.push_back(bits);
    P->tl_hmax.push_back(thmax[cand]);
    P->tl_hmin.push_back(thmin[cand]);
  }
  // ---- reference CSR pattern (dgga/solver.py:336-369) and its map to L ----
  std::vector<int64_t> flat;
  for (auto& kv : pairs) {
    flat.push_back((int64_t)kv.first.first * Nj + kv.first.second);
    flat.push_back((int64_t)kv.first.second * Nj + kv.first.first);
  }
  for (int r = 0; r < Nj; ++r) flat.push_back((int64_t)r * Nj + r);
  std::sort(flat.begin(), flat.end());
  flat.erase(std::unique(flat.begin(), flat.end()), flat.end());
  P->csr_indptr.assign(Nj + 1, 0);
  P->csr_indices.resize(flat.size());
  // pivot position of each factor entry's column (the unit-factor scaling)
  P->l_scale.assign(S.nnzL, -1);
  for (int j = 0; j < Nj; ++j)
    for (int32_t q = S.colptr[j] + 1; q < S.colptr[j + 1]; ++q) P->l_scale[q] = S.colptr[j];
  P->csr_to_L.resize(flat.size());
  for (size_t e = 0; e < flat.size(); ++e) {
    const int x = (int)(flat[e] / Nj), y = (int)(flat[e] % Nj);
    P->csr_indptr[x + 1]++;
    P->csr_indices[e] = y;
    const int sx = slot_of_row(x), sy = slot_of_row(y);
    P->csr_to_L[e] = (x == y) ? S.colptr[sx] : (sx > sy ? (int32_t)S.pos(sx, sy) : -1);
  }
  for (int r = 0; r < Nj; ++r) P->csr_indptr[r + 1] += P->csr_indptr[r];
  // slot of each CSR entry in the lane layout of A
  P->csr_to_ell.assign(P->csr_indices.size(), -1);
  for (int r = 0; r < Nj; ++r) {
    const int sr = P->row_slot[r], lane = sr % kLanes;
    const int32_t t0 = P->rs_t0[sr], t1 = t0 + P->spmv_ptr[sr + 1] - P->spmv_ptr[sr];
    for (int32_t e = P->csr_indptr[r]; e < P->csr_indptr[r + 1]; ++e) {
      const int sc = P->row_slot[P->csr_indices[e]];
      for (int32_t k = t0; k < t1; ++k) {
        const size_t x = (size_t)k * kLanes + lane;
        if ((P->rs_ell[2 * x + 1] & kColMask) == sc) {
          P->csr_to_ell[e] = (int32_t)x;
          break;
        }
      }
      if (P->csr_to_ell[e] < 0) throw std::logic_error("CSR entry outside the pattern of A");
    }
  }
  // ---- scalars ----
  P->headloss = d->headloss;
  P->hexp = d->hexp;
  P->qexp = d->qexp;
  P->rqtol = d->rqtol;
  P->viscos = d->viscos;
  P->htol = d->htol;
  P->qtol = d->qtol;
  P->checkfreq = d->checkfreq;
  P->maxcheck = d->maxcheck;
  P->extra_iter = d->extra_iter;
  // ---- stats ----
  wds_plan_stats& st = P->stats;
  st.N = N;
  st.L = L;
  st.Nj = Nj;
  st.nnzL = (int32_t)S.nnzL;
  st.nlevels = S.nlevels;
  st.npcv = P->p;
  st.nnz_csr = (int64_t)flat.size();
  st.nnz_lower = Nj + (int64_t)P->off_pos.size();
  st.factor_updates = S.factor_updates;
  st.factor_steps = S.factor.nsteps();
  st.fwd_steps = S.fwd.nsteps();
  st.bwd_steps = S.bwd.nsteps();
  st.factor_instr = (int64_t)S.factor.instr.size();
  st.fwd_instr = (int64_t)S.fwd.instr.size();
  st.bwd_instr = (int64_t)S.bwd.instr.size();
  st.lds_bytes = 8 * (lds_sink_index(S.nnzL, Nj, Nf) + kLanes +
                      lds_small_doubles((int64_t)P->wp_node.size(), 2 * (int64_t)P->p));
  // the program words address the layout by byte offset below 2^27
  if (st.lds_bytes >= ((int64_t)1 << 27))
    throw std::domain_error("network too large: " + std::to_string(st.lds_bytes) +
                            " bytes per scenario exceed the 128 MiB the programs address");
  // arena mode when the layout exceeds a CU's LDS (or when asked for)
  P->arena = ((ordering & 0x100) || st.lds_bytes > kMaxLdsBytes)
                 ? (st.lds_bytes / 8 + 31) / 32 * 32
                 : 0;
  st.arena_bytes = 8 * P->arena;
  st.ordering = ordering & 0xff;
  st.iteration_cost = (int32_t)std::min<int64_t>(iteration_cost(S), INT32_MAX);
}

// Scenarios of the plan one CU holds at a time in LDS mode (one wavefront per
// SIMD at most); arena mode is not LDS-limited.
static int lds_occupancy(const wds_plan* P) {
  if (P->arena > 0) return 4;
  return (int)std::min<int64_t>(4, kMaxLdsBytes / std::max<int64_t>(1, P->stats.lds_bytes));
}

// ordering 2: build the plan under each candidate ordering and keep the one
// with the best LDS occupancy, then the lowest scheduled iteration cost
// (minimum degree on ties)
static void build_auto(const wds_network_desc* d, int ordering, wds_plan* P) {
  const int flags = ordering & ~0xff;
  build(d, 0 | flags, P);
  for (int cap = 3; cap <= 6; ++cap) {
    wds_plan C;
    build(d, (kOrderRounds + cap) | flags, &C);
    const int oc = lds_occupancy(&C), op = lds_occupancy(P);
    if (oc > op || (oc == op && C.stats.iteration_cost < P->stats.iteration_cost))
      *P = std::move(C);
  }
}

}  // namespace wds

extern "C" int wds_plan_create(const wds_network_desc* net, int32_t ordering, wds_plan** out) {
  if (!net || !out) {
    wds::set_error("wds_plan_create: null argument");
    return WDS_EINVAL;
  }
  wds_plan* P = new wds_plan();
  try {
    if ((ordering & 0xff) == 2)
      wds::build_auto(net, ordering, P);
    else
      wds::build(net, ordering, P);
  } catch (const std::domain_error& e) {
    wds::set_error(e.what());
    delete P;
    return WDS_EUNSUPPORTED;
  } catch (const std::exception& e) {
    wds::set_error(e.what());
    delete P;
    return WDS_EINVAL;
  }
  *out = P;
  return WDS_OK;
}

extern "C" int wds_plan_stats_get(const wds_plan* plan, wds_plan_stats* out) {
  if (!plan || !out) return WDS_EINVAL;
  *out = plan->stats;
  return WDS_OK;
}

static int64_t copy_out(const std::vector<int32_t>& v, int32_t* dst, int64_t cap) {
  if (dst && cap > 0) std::memcpy(dst, v.data(), sizeof(int32_t) * std::min<int64_t>(cap, v.size()));
  return (int64_t)v.size();
}

extern "C" int64_t wds_plan_export(const wds_plan* plan, int32_t what, int32_t* dst, int64_t cap) {
  if (!plan) return WDS_EINVAL;
  const wds::Symbolic& S = plan->sym;
  switch (what) {
    case 0: return copy_out(S.perm, dst, cap);
    case 1: return copy_out(S.colptr, dst, cap);
    case 2: return copy_out(S.rowidx, dst, cap);
    case 3: return copy_out(S.parent, dst, cap);
    case 4: return copy_out(plan->node_slot, dst, cap);
    case 5: return copy_out(plan->csr_indptr, dst, cap);
    case 6: return copy_out(plan->csr_indices, dst, cap);
    case 7: return copy_out(plan->csr_to_L, dst, cap);
    case 17: return copy_out(plan->rs_share, dst, cap);
    case 18: return copy_out(plan->rs_ell, dst, cap);
    case 19: return copy_out(plan->row_slot, dst, cap);
    default: break;
  }
  if (what >= 8 && what < 17) {
    const int p = (what - 8) / 3, part = (what - 8) % 3;
    const wds::Program& pr = p == 0 ? S.factor : (p == 1 ? S.fwd : S.bwd);
    if (part == 0) {
      std::vector<int32_t> v(pr.instr.size() * 4);
      std::memcpy(v.data(), pr.instr.data(), v.size() * sizeof(int32_t));
      return copy_out(v, dst, cap);
    }
    if (part == 1) return 0;  // reserved (no per-step table in this layout)
    return copy_out(pr.level_end, dst, cap);
  }
  return WDS_EINVAL;
}

extern "C" const char* wds_last_error(void) { return wds::g_err.c_str(); }
