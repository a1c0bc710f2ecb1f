// MI355X (gfx950) kernels of the batched GGA forward solve.
//
// Mapping: one 64-lane wavefront (one workgroup) per scenario.  The scenario's
// numeric factor, its heads and one work vector live in LDS for the whole
// Newton loop; per-link coefficient arrays and the assembled matrix values
// sit in a per-scenario global workspace that stays L2-resident.  Every
// Newton iteration (dgga/solver.py:2922-3179) runs inside the one launch:
//   K1 link coefficients  -> K2 gather assembly (reference order)
//   -> K3 LDL^T factor    -> K4 solve + refinement
//   -> K5 flow update + relerr (wave reduction) -> K6 status machine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <mutex>
#include <vector>

#include "plan.h"
#include "arena.h"
#include "device_common.h"
#include "kernels_impl.h"

using namespace wds;

// ---------------------------------------------------------------------------
// plan upload
namespace {

struct BlobBuilder {
  std::vector<char> bytes;
  template <class T>
  size_t add(const std::vector<T>& v) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + std::max<size_t>(v.size() * sizeof(T), 8));
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return WDS_EDEVICE;
}

// steps the executor runs: the program padded to whole turns of the ring
int32_t exec_steps(const Program& pr) {
  return (pr.nsteps() + kSlots - 1) / kSlots * kSlots;
}

template <class T>
const T* at(void* base, size_t off) {
  return reinterpret_cast<const T*>(static_cast<char*>(base) + off);
}

}  // namespace

extern "C" int wds_plan_upload(wds_plan* P) {
  if (!P) return WDS_EINVAL;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (P->d_blob && P->device == dev) return WDS_OK;
  if (P->d_blob) {
    set_error("plan already uploaded to another device");
    return WDS_EINVAL;
  }
  const Symbolic& S = P->sym;
  BlobBuilder bb;
  auto prog_vectors = [&](const Program& pr, size_t& oi, size_t& os, size_t& ol) {
    // trailing NOP steps: the prefetch ring reads up to kRing steps past the
    // end
    std::vector<Instr> padded(pr.instr);
    padded.resize((size_t)(exec_steps(pr) + kRing) * kLanes,
                  nop_instr((int32_t)S.nnzL));
    oi = bb.add(padded);
    os = ol = 0;
  };
  // solve programs: index instructions -> LDS byte addresses (V at byte 0,
  // the solve vector S after V, the heads and the fixed heads) and per-step
  // uniform flags
  const int32_t sink_i = (int32_t)lds_sink_index(S.nnzL, P->Nj, P->Nf);
  const int32_t zero_b = 8 * (int32_t)S.nnzL, sink_b = 8 * sink_i;
  const int32_t s_b = 8 * ((int32_t)S.nnzL + 2 + P->Nj + P->Nf);
  auto dot_vectors = [&](const Program& pr, size_t& oi, size_t& os) {
    const int32_t ns = exec_steps(pr);
    std::vector<Instr> w((size_t)(ns + kRing) * kLanes);
    for (size_t i = 0; i < w.size(); ++i) {
      const int32_t own = sink_b + 8 * (int32_t)(i % kLanes);  // the lane's sink
      w[i] = Instr{zero_b, s_b, own, own};
    }
    for (size_t i = 0; i < pr.instr.size(); ++i) {
      const Instr& in = pr.instr[i];
      const int32_t fl = in.t;
      const int32_t own = sink_b + 8 * (int32_t)(i % kLanes);
      if (!(fl & kNop)) {
        const int32_t t = fl & kTMask;
        const int32_t init = (fl & kFirst) ? ((fl & kZeroInit) ? zero_b : s_b + 8 * t) : own;
        const int32_t store = (fl & kLast) ? s_b + 8 * t : own;
        w[i] = Instr{8 * in.a, s_b + 8 * in.b, init, store};
      }
      // a level end (every lane) and the lane's group width ride in the
      // high bits of the store word
      if (fl & kLevelEnd)
        w[i].t |= kDotLevelEnd | (((fl >> kWidthShift) & 7) << kDotWidthShift);
    }
    oi = bb.add(w);
    os = 0;
  };
  size_t o_ns = bb.add(P->node_slot), o_sn = bb.add(P->slot_node), o_el = bb.add(P->el_slot);
  size_t o_s1 = bb.add(P->l_s1), o_s2 = bb.add(P->l_s2), o_fl = bb.add(P->l_flags),
         o_st0 = bb.add(P->l_st0);
  size_t o_di = bb.add(P->l_diam), o_km = bb.add(P->l_km), o_kt = bb.add(P->l_km_tcv),
         o_kv = bb.add(P->l_km_valve), o_kc = bb.add(P->l_kc), o_q0 = bb.add(P->l_q0);
  size_t o_ph = bb.add(P->pu_h0), o_pr = bb.add(P->pu_r), o_pn = bb.add(P->pu_n),
         o_ps = bb.add(P->pu_speed), o_pm = bb.add(P->pu_hmax);
  size_t o_cp = bb.add(P->pu_cptr), o_cx = bb.add(P->pu_cx), o_cy = bb.add(P->pu_cy);
  size_t o_jp = bb.add(P->jadj_ptr), o_ja = bb.add(P->jadj);
  size_t o_fap = bb.add(P->fadj_ptr), o_fa = bb.add(P->fadj);
  size_t o_op = bb.add(P->off_pos), o_optr = bb.add(P->off_ptr), o_ol = bb.add(P->off_link);
  size_t o_fp = bb.add(P->fill_pos), o_sp = bb.add(P->spmv_ptr), o_se = bb.add(P->spmv_ent);
  size_t o_rb = bb.add(P->rs_share), o_re = bb.add(P->rs_ell);
  size_t o_ab = bb.add(P->as_blk), o_ae = bb.add(P->as_ell), o_o4 = bb.add(P->off4);
  size_t o_pb = bb.add(P->ps_blk), o_pe = bb.add(P->ps_ell), o_pln = bb.add(P->ps_len);
  size_t o_prt = bb.add(P->prv_rt);
  size_t o_dae = bb.add(P->diag_ae), o_oae = bb.add(P->off_ae), o_pae = bb.add(P->prv_ae),
         o_cae = bb.add(P->csr_to_ell);
  size_t o_pk = bb.add(P->prv_k), o_p1 = bb.add(P->prv_s1), o_p2 = bb.add(P->prv_s2),
         o_po = bb.add(P->prv_pos_off), o_hs = bb.add(P->prv_hset), o_pkm = bb.add(P->prv_km),
         o_pkd = bb.add(P->prv_kind);
  size_t o_tk = bb.add(P->tl_k), o_tt = bb.add(P->tl_tank), o_to = bb.add(P->tl_other),
         o_tb = bb.add(P->tl_bits), o_tx = bb.add(P->tl_hmax), o_tn = bb.add(P->tl_hmin);
  size_t o_jgp = bb.add(P->jg_ptr), o_jg = bb.add(P->jg), o_jm = bb.add(P->jadj_mid);
  size_t o_wpp = bb.add(P->wp_ptr), o_wpn = bb.add(P->wp_node), o_wpd = bb.add(P->wp_diag),
         o_wtp = bb.add(P->wt_ptr), o_wt = bb.add(P->wt), o_wpr = bb.add(P->wpair),
         o_wq = bb.add(P->wq), o_wqn = bb.add(P->wq_n);
  size_t o_c2l = bb.add(P->csr_to_L), o_dp = bb.add(P->diag_pos), o_rs = bb.add(P->row_slot);
  size_t fi, fs, fl, wi, ws, bi, bs;
  prog_vectors(S.factor, fi, fs, fl);
  dot_vectors(S.fwd, wi, ws);
  dot_vectors(S.bwd, bi, bs);
  size_t o_lsc = bb.add(P->l_scale);
  void* d = nullptr;
  e = hipMalloc(&d, bb.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(plan)");
  e = hipMemcpy(d, bb.bytes.data(), bb.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "hipMemcpy(plan)");
  }
  DevPlan& D = P->dev;
  D.arena = P->arena;
  D.N = P->N;
  D.L = P->L;
  D.Nj = P->Nj;
  D.Nf = P->Nf;
  D.nnzL = (int32_t)S.nnzL;
  D.p = P->p;
  D.prv_disjoint = P->prv_disjoint;
  D.ntl = (int32_t)P->tl_k.size();
  D.headloss = P->headloss;
  D.hexp = P->hexp;
  D.qexp = P->qexp;
  D.rqtol = P->rqtol;
  D.viscos = P->viscos;
  D.htol = P->htol;
  D.qtol = P->qtol;
  D.checkfreq = P->checkfreq;
  D.maxcheck = P->maxcheck;
  D.extra_iter = P->extra_iter;
  D.has_cv = P->has_cv;
  D.has_pump = P->has_pump;
  D.has_fcv = P->has_fcv;
  D.node_slot = at<int32_t>(d, o_ns);
  D.slot_node = at<int32_t>(d, o_sn);
  D.el_slot = at<double>(d, o_el);
  D.l_s1 = at<int32_t>(d, o_s1);
  D.l_s2 = at<int32_t>(d, o_s2);
  D.l_flags = at<int32_t>(d, o_fl);
  D.l_st0 = at<int8_t>(d, o_st0);
  D.l_diam = at<double>(d, o_di);
  D.l_km = at<double>(d, o_km);
  D.l_km_tcv = at<double>(d, o_kt);
  D.l_km_valve = at<double>(d, o_kv);
  D.l_kc = at<double>(d, o_kc);
  D.l_q0 = at<double>(d, o_q0);
  D.pu_h0 = at<double>(d, o_ph);
  D.pu_r = at<double>(d, o_pr);
  D.pu_n = at<double>(d, o_pn);
  D.pu_speed = at<double>(d, o_ps);
  D.pu_hmax = at<double>(d, o_pm);
  D.pu_cptr = at<int32_t>(d, o_cp);
  D.pu_cx = at<double>(d, o_cx);
  D.pu_cy = at<double>(d, o_cy);
  D.ucf_flow = P->ucf_flow;
  D.ucf_head = P->ucf_head;
  D.jadj_ptr = at<int32_t>(d, o_jp);
  D.jadj = at<int32_t>(d, o_ja);
  D.fadj_ptr = at<int32_t>(d, o_fap);
  D.fadj = at<int32_t>(d, o_fa);
  D.noff = (int32_t)P->off_pos.size();
  D.off_pos = at<int32_t>(d, o_op);
  D.off_ptr = at<int32_t>(d, o_optr);
  D.off_link = at<int32_t>(d, o_ol);
  D.nfill = (int32_t)P->fill_pos.size();
  D.fill_pos = at<int32_t>(d, o_fp);
  D.spmv_ptr = at<int32_t>(d, o_sp);
  D.spmv_ent = at<int2>(d, o_se);
  D.rs_share = at<int4>(d, o_rb);
  D.rs_ell = at<int2>(d, o_re);
  for (int i = 0; i < 3; ++i) D.rs_len[i] = P->rs_len[i];
  D.as_blk = at<int32_t>(d, o_ab);
  D.as_ell = at<int2>(d, o_ae);
  D.off4 = at<int4>(d, o_o4);
  D.ps_blk = at<int32_t>(d, o_pb);
  D.ps_ell = at<int32_t>(d, o_pe);
  D.ps_len = at<int32_t>(d, o_pln);
  D.na = (int32_t)std::max<size_t>(S.nnzL, P->rs_ell.size() / 2);
  D.diag_ae = at<int32_t>(d, o_dae);
  D.off_ae = at<int2>(d, o_oae);
  D.prv_ae = at<int4>(d, o_pae);
  D.prv_rt = at<int4>(d, o_prt);
  D.csr_to_ell = at<int32_t>(d, o_cae);
  D.prv_k = at<int32_t>(d, o_pk);
  D.prv_kind = at<int32_t>(d, o_pkd);
  D.prv_s1 = at<int32_t>(d, o_p1);
  D.prv_s2 = at<int32_t>(d, o_p2);
  D.prv_pos_off = at<int32_t>(d, o_po);
  D.prv_hset = at<double>(d, o_hs);
  D.prv_km = at<double>(d, o_pkm);
  D.jg_ptr = at<int32_t>(d, o_jgp);
  D.jadj_mid = at<int32_t>(d, o_jm);
  D.jg = at<int2>(d, o_jg);
  D.np = 2 * P->p;
  D.wp_tot = (int32_t)P->wp_node.size();
  D.wp_ptr = at<int32_t>(d, o_wpp);
  D.wp_node = at<int32_t>(d, o_wpn);
  D.wp_diag = at<int32_t>(d, o_wpd);
  D.wt_ptr = at<int32_t>(d, o_wtp);
  D.wt = at<int2>(d, o_wt);
  D.wpair = at<int4>(d, o_wpr);
  D.wq_ok = P->wq_ok;
  D.wq = at<int2>(d, o_wq);
  D.wq_n = at<int32_t>(d, o_wqn);
  D.tl_k = at<int32_t>(d, o_tk);
  D.tl_tank = at<int32_t>(d, o_tt);
  D.tl_other = at<int32_t>(d, o_to);
  D.tl_bits = at<int32_t>(d, o_tb);
  D.tl_hmax = at<double>(d, o_tx);
  D.tl_hmin = at<double>(d, o_tn);
  D.nnz_csr = (int32_t)P->csr_to_L.size();
  D.csr_to_L = at<int32_t>(d, o_c2l);
  D.diag_pos = at<int32_t>(d, o_dp);
  D.row_slot = at<int32_t>(d, o_rs);
  auto set_prog = [&](ProgDev& pg, const Program& pr, size_t oi, size_t os, bool dot) {
    pg.instr = at<int4>(d, oi);
    (void)os;
    pg.nsteps = exec_steps(pr);
    pg.dummy = dot ? sink_b : sink_i;
    pg.nlevels = (int32_t)pr.level_end.size();
  };
  (void)fs;
  (void)fl;
  set_prog(D.factor, S.factor, fi, 0, false);
  set_prog(D.fwd, S.fwd, wi, ws, true);
  set_prog(D.bwd, S.bwd, bi, bs, true);
  D.l_scale = at<int32_t>(d, o_lsc);
  P->d_blob = d;
  P->device = dev;
  return WDS_OK;
}

extern "C" void wds_plan_destroy(wds_plan* P) {
  if (!P) return;
  if (P->d_blob) (void)hipFree(P->d_blob);
  delete P;
}

extern "C" int64_t wds_workspace_bytes(const wds_plan* P, int32_t B) {
  if (!P || B <= 0) return WDS_EINVAL;
  return (int64_t)B *
         (ws_doubles(P->L, P->Nj, (int32_t)std::max<size_t>(P->sym.nnzL, P->rs_ell.size() / 2),
                     (int32_t)P->wp_node.size(), 2 * P->p) +
          P->arena) *
         8;
}

extern "C" int wds_counters_get(wds_plan* P, wds_counters* out, int32_t reset) {
  if (!P || !out) return WDS_EINVAL;
  *out = P->cnt;
  if (reset) P->cnt = wds_counters{};
  return WDS_OK;
}

static int check_ready(wds_plan* P);

extern "C" int wds_plan_set_pumps(wds_plan* P, void* stream, const double* speed,
                                  const double* h0, const double* r, const double* hmax,
                                  const double* q0) {
  int rc = check_ready(P);
  if (rc) return rc;
  const int L = P->L;
  for (int k = 0; k < L; ++k) {
    int32_t& f = P->l_flags[k];
    if (!(f & LF_PUMP)) continue;
    if (speed) {
      if (!(speed[k] >= 0.0) || !std::isfinite(speed[k])) {
        set_error("wds_plan_set_pumps: pump speeds must be finite and >= 0");
        return WDS_EINVAL;
      }
      P->pu_speed[k] = speed[k];
      f &= ~(LF_ZERO_SPEED | LF_PUMP_LIVE);
      if (speed[k] == 0.0) f |= LF_ZERO_SPEED;
      if (speed[k] > 0.0) f |= LF_PUMP_LIVE;
    }
    if (h0) P->pu_h0[k] = h0[k];
    if (r) P->pu_r[k] = r[k];
    if (hmax) P->pu_hmax[k] = hmax[k];
    if (q0) P->l_q0[k] = q0[k];
  }
  // queued behind the work already on the stream; pageable sources are
  // staged by the runtime before the call returns
  hipStream_t st = (hipStream_t)stream;
  const wds::DevPlan& D = P->dev;
  struct {
    const void* dst;
    const void* src;
    size_t bytes;
  } cp[] = {{D.l_flags, P->l_flags.data(), sizeof(int32_t) * (size_t)L},
            {D.pu_speed, P->pu_speed.data(), sizeof(double) * (size_t)L},
            {D.pu_h0, P->pu_h0.data(), sizeof(double) * (size_t)L},
            {D.pu_r, P->pu_r.data(), sizeof(double) * (size_t)L},
            {D.pu_hmax, P->pu_hmax.data(), sizeof(double) * (size_t)L},
            {D.l_q0, P->l_q0.data(), sizeof(double) * (size_t)L}};
  for (const auto& c : cp) {
    hipError_t e = hipMemcpyAsync(const_cast<void*>(c.dst), c.src, c.bytes,
                                  hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "wds_plan_set_pumps");
  }
  return WDS_OK;
}

static int check_ready(wds_plan* P) {
  if (!P) {
    set_error("null plan");
    return WDS_EINVAL;
  }
  if (!P->d_blob) {
    int r = wds_plan_upload(P);
    if (r) return r;
  }
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != P->device) {
    set_error("plan was uploaded to another device");
    return WDS_EINVAL;
  }
  if (P->arena == 0 && P->stats.lds_bytes > kMaxLds) {
    set_error("network needs " + std::to_string(P->stats.lds_bytes) +
              " bytes of LDS per scenario (> " + std::to_string(kMaxLds) + ")");
    return WDS_EUNSUPPORTED;
  }
  // the solve programs address LDS by absolute byte offsets from the start of
  // the dynamic allocation: no kernel may carry static LDS
  static bool lds_checked = false;
  if (!lds_checked) {
    const void* ks[] = {(const void*)k_solve, (const void*)k_newton_iter,
                        (const void*)k_factor_solve, (const void*)k_adjoint,
                        (const void*)k_solve_multi, (const void*)k_adjoint_multi};
    for (const void* k : ks) {
      hipFuncAttributes at{};
      hipError_t e = hipFuncGetAttributes(&at, k);
      if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes");
      if (at.sharedSizeBytes != 0) {
        set_error("a solver kernel has static LDS; the solve programs assume none");
        return WDS_EDEVICE;
      }
    }
    lds_checked = true;
  }
  return WDS_OK;
}

static int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return WDS_OK;
}

// dynamic LDS beyond the default 64 KiB needs the kernel's attribute raised
template <class K>
static hipError_t set_lds(K kernel, int64_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
#define WDS_SET_LDS(kernel, bytes)                                          \
  do {                                                                      \
    const hipError_t le_ = set_lds(kernel, bytes);                          \
    if (le_ != hipSuccess) return hip_fail(le_, "hipFuncSetAttribute(" #kernel ")"); \
  } while (0)

extern "C" int wds_factor_solve(wds_plan* P, void* stream, int32_t B, const double* csr,
                                const double* F, double* H, int32_t n_refine, double* factor_out,
                                int32_t* info, void* workspace) {
  int r = check_ready(P);
  if (r) return r;
  if (B <= 0 || !csr || !F || !H || !info || !workspace || n_refine < 0) {
    set_error("wds_factor_solve: bad argument");
    return WDS_EINVAL;
  }
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_factor_solve(P->dev, csr, F, H, n_refine, factor_out, info, (double*)workspace, B,
                       (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_factor_solve, lds);
    hipLaunchKernelGGL(k_factor_solve, dim3(B), dim3(64), lds, (hipStream_t)stream, P->dev, csr,
                       F, H, n_refine, factor_out, info, (double*)workspace);
  }
  P->cnt.factorize += 1;
  P->cnt.solve += 1 + n_refine;
  P->cnt.launches += 1;
  return launch_check("k_factor_solve");
}

extern "C" int wds_factor(wds_plan* P, void* stream, int32_t B, const double* csr,
                          double* factor_out, int32_t* info, void* workspace) {
  int r = check_ready(P);
  if (r) return r;
  if (B <= 0 || !csr || !factor_out || !info || !workspace) {
    set_error("wds_factor: bad argument");
    return WDS_EINVAL;
  }
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_factor(P->dev, csr, factor_out, info, (double*)workspace, B, (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_factor, lds);
    hipLaunchKernelGGL(k_factor, dim3(B), dim3(64), lds, (hipStream_t)stream, P->dev, csr,
                       factor_out, info, (double*)workspace);
  }
  P->cnt.factorize += 1;
  P->cnt.launches += 1;
  return launch_check("k_factor");
}

extern "C" int wds_trisolve(wds_plan* P, void* stream, int32_t B, const double* factor,
                            const double* csr, const double* F, double* H, int32_t n_refine,
                            void* workspace) {
  int r = check_ready(P);
  if (r) return r;
  if (B <= 0 || !factor || !csr || !F || !H || !workspace || n_refine < 0) {
    set_error("wds_trisolve: bad argument");
    return WDS_EINVAL;
  }
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_trisolve(P->dev, factor, csr, F, H, n_refine, (double*)workspace, B,
                   (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_trisolve, lds);
    hipLaunchKernelGGL(k_trisolve, dim3(B), dim3(64), lds, (hipStream_t)stream, P->dev, factor,
                       csr, F, H, n_refine, (double*)workspace);
  }
  P->cnt.solve += 1 + n_refine;
  P->cnt.launches += 1;
  return launch_check("k_trisolve");
}

// record_schedule buffers: all or none, status machine on, event words wide
// enough for the iteration and link numbers
static int check_sched(const wds_plan* P, const wds_solve_args* a) {
  // the polish and the terminal factor's adjoint solve the row-replaced
  // systems of ACTIVE PRVs only
  if (P->has_psv_fcv && a->polish_steps >= 0) {
    set_error("the polish (and the implicit adjoint) of networks with PSVs or FCVs with a "
              "setting is not built; solve without polish (polish_steps < 0)");
    return WDS_EUNSUPPORTED;
  }
  const bool any = a->sched_flags || a->sched_events || a->sched_count;
  if (!any) return WDS_OK;
  const int64_t maxtrials =
      (int64_t)a->max_iter + (P->extra_iter > 0 ? P->extra_iter : 0);
  if (!a->sched_flags || !a->sched_events || !a->sched_count || a->sched_cap < 0) {
    set_error("schedule recording needs sched_flags, sched_events, sched_count and cap >= 0");
    return WDS_EINVAL;
  }
  if (!a->status_machine) {
    set_error("schedule recording covers the batched status machine only "
              "(dgga/solver.py:2802-2806)");
    return WDS_EUNSUPPORTED;
  }
  if (maxtrials >= 2048 || P->L >= (1 << 19)) {
    set_error("schedule recording needs maxtrials < 2048 and L < 524288");
    return WDS_EINVAL;
  }
  return WDS_OK;
}

extern "C" int wds_solve(wds_plan* P, void* stream, const wds_solve_args* a) {
  int r = check_ready(P);
  if (r) return r;
  if (!a || a->B <= 0 || !a->demand || !a->res_head || !a->ke || !a->r_hw || !a->head ||
      !a->flow || !a->emitter || !a->iters || !a->relerr || !a->status || !a->info ||
      !a->workspace || a->max_iter < 0 || a->n_refine < 0) {
    set_error("wds_solve: bad argument");
    return WDS_EINVAL;
  }
  if (P->p > 0 && !a->status_machine) {
    set_error("networks with PRVs need status_machine=1 (dgga/solver.py:2786-2797)");
    return WDS_EUNSUPPORTED;
  }
  if (P->has_cv && !a->status_machine) {
    set_error("networks with CV pipes need status_machine=1");
    return WDS_EUNSUPPORTED;
  }
  r = check_sched(P, a);
  if (r) return r;
  SolveParams sp = make_params(a, P->N);
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_solve(P->dev, sp, a->B, (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_solve, lds);
    hipLaunchKernelGGL(k_solve, dim3(a->B), dim3(64 * kSolveWaves), lds, (hipStream_t)stream,
                       P->dev, sp);
  }
  P->cnt.launches += 1;
  return launch_check("k_solve");
}

// ---------------------------------------------------------------------------
// wds_solve_noalloc: the fused forward with its workspace in a file-scope
// device array instead of caller memory, for callers that cannot allocate.
// Every GPU the library runs on holds its own copy of the array until the
// process ends; host code addresses it only through the pointer
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[WDS_NOALLOC_SCRATCH_BYTES / 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// device address of this device's scratch array, looked up once per device
// with that device current
int scratch_of(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("wds_solve_noalloc: device index out of range");
    return WDS_EDEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return WDS_OK;
}
}  // namespace

extern "C" int wds_solve_noalloc(wds_plan* P, void* stream, const wds_solve_args* a) {
  if (!P || !a) {
    set_error("wds_solve_noalloc: null argument");
    return WDS_EINVAL;
  }
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  // the plan upload allocates: it must have happened already
  if (!P->d_blob || P->device != dev) {
    set_error("wds_solve_noalloc: the plan is not uploaded to the current device "
              "(call wds_plan_upload first)");
    return WDS_EINVAL;
  }
  if (a->workspace) {
    set_error("wds_solve_noalloc: workspace must be NULL (the library's scratch is used)");
    return WDS_EINVAL;
  }
  if (a->B <= 0) {
    set_error("wds_solve_noalloc: bad argument");
    return WDS_EINVAL;
  }
  const int64_t need = wds_workspace_bytes(P, a->B);
  if (need > (int64_t)WDS_NOALLOC_SCRATCH_BYTES) {
    set_error("wds_solve_noalloc: " + std::to_string(a->B) + " scenarios need " +
              std::to_string(need) + " bytes of scratch, more than the " +
              std::to_string((int64_t)WDS_NOALLOC_SCRATCH_BYTES) + " available (at most " +
              std::to_string((int64_t)WDS_NOALLOC_SCRATCH_BYTES / (need / a->B)) +
              " scenarios of this network per call)");
    return WDS_EINVAL;
  }
  void* ptr = nullptr;
  int r = scratch_of(dev, &ptr);
  if (r) return r;
  wds_solve_args b = *a;
  b.workspace = ptr;
  return wds_solve(P, stream, &b);
}

// ---------------------------------------------------------------------------
// Ragged batches of networks (wds_multi_*): one persistent launch per kernel
// mode over a work list of (network, scenario) items.
struct wds_multi {
  std::vector<wds_plan*> plans;
  std::vector<int32_t> B;
  std::vector<int32_t> arena;     // per network: 1 arena mode, 0 LDS mode
  std::vector<int64_t> ws_bytes;  // per network workspace (mode-aware)
  int64_t lds_bytes = 0;          // dynamic LDS of the LDS-mode launch
  int32_t n_lds = 0, n_arena = 0, grid_lds = 0, grid_arena = 0;
  int device = -1;
  void* d_mem = nullptr;
  DevPlan* d_plans = nullptr;
  SolveParams* d_sp = nullptr;
  AdjParams* d_ap = nullptr;
  int2* d_items = nullptr;
  int32_t* d_counter = nullptr;
  std::vector<SolveParams> h_sp;
  std::vector<AdjParams> h_ap;
};

namespace {
__global__ void k_reset_counters(int32_t* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

// latency proxy of one scenario of a network: program steps of a Newton
// iteration (factor + three solve pairs) plus its link and row sweeps
int64_t scenario_cost(const wds_plan* P) {
  const wds_plan_stats& st = P->stats;
  return (int64_t)st.factor_steps + 3 * ((int64_t)st.fwd_steps + st.bwd_steps) +
         ((int64_t)P->L + P->Nj) / 64;
}
}  // namespace

extern "C" int wds_multi_create(wds_plan* const* plans, int32_t nnet, const int32_t* B,
                                int64_t lds_budget, wds_multi** out) {
  if (!plans || !B || !out || nnet <= 0) {
    set_error("wds_multi_create: bad argument");
    return WDS_EINVAL;
  }
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  wds_multi* M = new wds_multi();
  M->device = dev;
  std::vector<DevPlan> dp(nnet);
  std::vector<int64_t> cost(nnet);
  for (int i = 0; i < nnet; ++i) {
    wds_plan* P = plans[i];
    if (!P || B[i] <= 0) {
      delete M;
      set_error("wds_multi_create: null plan or non-positive batch");
      return WDS_EINVAL;
    }
    int r = wds_plan_upload(P);
    if (r) {
      delete M;
      return r;
    }
    const bool ar = P->arena > 0 || lds_budget <= 0 || P->stats.lds_bytes > lds_budget ||
                    P->stats.lds_bytes + kItemSlotBytes > kMaxLds;
    dp[i] = P->dev;
    dp[i].arena = ar ? (P->arena > 0 ? P->arena : (P->stats.lds_bytes / 8 + 31) / 32 * 32) : 0;
    M->plans.push_back(P);
    M->B.push_back(B[i]);
    M->arena.push_back(ar ? 1 : 0);
    M->ws_bytes.push_back((int64_t)B[i] * ws_stride(dp[i]) * 8);
    if (!ar) M->lds_bytes = std::max<int64_t>(M->lds_bytes, P->stats.lds_bytes);
    cost[i] = scenario_cost(P);
  }
  // items: the LDS-mode group first, then the arena-mode group, each with its
  // costliest networks' scenarios first
  std::vector<int> order(nnet);
  for (int i = 0; i < nnet; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<int2> items;
  for (int pass = 0; pass < 2; ++pass)
    for (int i : order)
      if (M->arena[i] == pass)
        for (int b = 0; b < B[i]; ++b) items.push_back(make_int2(i, b));
  for (int i = 0; i < nnet; ++i) (M->arena[i] ? M->n_arena : M->n_lds) += B[i];
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) {
    delete M;
    return hip_fail(e, "hipGetDeviceProperties");
  }
  const int cus = prop.multiProcessorCount;
  // one wavefront per SIMD (four per CU), fewer where the LDS layout is
  // larger than a quarter of the CU's LDS
  const int per_cu_lds =
      M->lds_bytes > 0
          ? (int)std::max<int64_t>(1, std::min<int64_t>(4, kMaxLds / (M->lds_bytes + kItemSlotBytes)))
          : 4;
  M->grid_lds = std::min(M->n_lds, cus * per_cu_lds);
  M->grid_arena = std::min(M->n_arena, cus * 4);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_plans = 0, o_sp = al(o_plans + nnet * sizeof(DevPlan)),
               o_ap = al(o_sp + nnet * sizeof(SolveParams)),
               o_items = al(o_ap + nnet * sizeof(AdjParams)),
               o_cnt = al(o_items + items.size() * sizeof(int2)), total = o_cnt + 256;
  e = hipMalloc(&M->d_mem, total);
  if (e != hipSuccess) {
    delete M;
    return hip_fail(e, "hipMalloc(multi)");
  }
  char* base = (char*)M->d_mem;
  M->d_plans = (DevPlan*)(base + o_plans);
  M->d_sp = (SolveParams*)(base + o_sp);
  M->d_ap = (AdjParams*)(base + o_ap);
  M->d_items = (int2*)(base + o_items);
  M->d_counter = (int32_t*)(base + o_cnt);
  e = hipMemcpy(M->d_plans, dp.data(), nnet * sizeof(DevPlan), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(M->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(M->d_mem);
    delete M;
    return hip_fail(e, "hipMemcpy(multi)");
  }
  M->h_sp.resize(nnet);
  M->h_ap.resize(nnet);
  *out = M;
  return WDS_OK;
}

extern "C" void wds_multi_destroy(wds_multi* M) {
  if (!M) return;
  if (M->d_mem) (void)hipFree(M->d_mem);
  delete M;
}

extern "C" int64_t wds_multi_workspace_bytes(const wds_multi* M, int32_t net) {
  if (!M || net < 0 || net >= (int32_t)M->plans.size()) return WDS_EINVAL;
  return M->ws_bytes[net];
}

extern "C" int wds_multi_info(const wds_multi* M, int32_t* n_lds, int32_t* n_arena,
                              int64_t* lds_bytes) {
  if (!M) return WDS_EINVAL;
  if (n_lds) *n_lds = M->n_lds;
  if (n_arena) *n_arena = M->n_arena;
  if (lds_bytes) *lds_bytes = M->lds_bytes;
  return WDS_OK;
}

static int multi_check(wds_multi* M) {
  if (!M) {
    set_error("null multi batch");
    return WDS_EINVAL;
  }
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != M->device) {
    set_error("multi batch built on another device");
    return WDS_EINVAL;
  }
  for (wds_plan* P : M->plans) {
    int r = check_ready(P);
    if (r && !(P->arena == 0 && P->stats.lds_bytes > kMaxLds)) return r;
  }
  return WDS_OK;
}

extern "C" int wds_multi_solve(wds_multi* M, void* stream, const wds_solve_args* args) {
  int r = multi_check(M);
  if (r) return r;
  const int nnet = (int)M->plans.size();
  for (int i = 0; i < nnet; ++i) {
    const wds_solve_args* a = &args[i];
    wds_plan* P = M->plans[i];
    if (a->B != M->B[i] || !a->demand || !a->res_head || !a->ke || !a->r_hw || !a->head ||
        !a->flow || !a->emitter || !a->iters || !a->relerr || !a->status || !a->info ||
        !a->workspace || a->max_iter < 0 || a->n_refine < 0) {
      set_error("wds_multi_solve: bad argument for network " + std::to_string(i));
      return WDS_EINVAL;
    }
    if ((P->p > 0 || P->has_cv) && !a->status_machine) {
      set_error("networks with PRVs or CV pipes need status_machine=1");
      return WDS_EUNSUPPORTED;
    }
    r = check_sched(P, a);
    if (r) return r;
    M->h_sp[i] = make_params(a, P->N);
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(M->d_sp, M->h_sp.data(), nnet * sizeof(SolveParams),
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(multi params)");
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, M->d_counter);
  if (M->n_lds) {
    // the item index passes between the wavefronts in a word after the layout
    WDS_SET_LDS(k_solve_multi, M->lds_bytes + kItemSlotBytes);
    hipLaunchKernelGGL(k_solve_multi, dim3(M->grid_lds), dim3(64 * kSolveWaves),
                       M->lds_bytes + kItemSlotBytes, st, M->d_plans, M->d_sp, M->d_items,
                       M->n_lds, M->d_counter, (int32_t)M->lds_bytes);
  }
  if (M->n_arena)
    arena_solve_multi(M->d_plans, M->d_sp, M->d_items + M->n_lds, M->n_arena,
                      M->d_counter + 1, M->grid_arena, st);
  for (wds_plan* P : M->plans) P->cnt.launches += 1;
  return launch_check("k_solve_multi");
}

extern "C" int wds_multi_adjoint(wds_multi* M, void* stream, const wds_adjoint_args* args) {
  int r = multi_check(M);
  if (r) return r;
  const int nnet = (int)M->plans.size();
  for (int i = 0; i < nnet; ++i) {
    const wds_adjoint_args* a = &args[i];
    if (a->B != M->B[i] || !a->factor || !a->g_demand || !a->g_res_head || !a->g_ke ||
        !a->g_r || !a->info || !a->workspace || !a->head || !a->flow || !a->emitter ||
        !a->status || !a->ke || !a->r_hw || a->n_refine < 0) {
      set_error("wds_multi_adjoint: bad argument for network " + std::to_string(i));
      return WDS_EINVAL;
    }
    if (M->plans[i]->has_psv_fcv) {
      set_error("the implicit adjoint of networks with PSVs or FCVs with a setting is not built");
      return WDS_EUNSUPPORTED;
    }
    M->h_ap[i] = make_adj_params(a);
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(M->d_ap, M->h_ap.data(), nnet * sizeof(AdjParams),
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(multi params)");
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, M->d_counter);
  if (M->n_lds) {
    WDS_SET_LDS(k_adjoint_multi, M->lds_bytes + kItemSlotBytes);
    hipLaunchKernelGGL(k_adjoint_multi, dim3(M->grid_lds), dim3(64 * kSolveWaves),
                       M->lds_bytes + kItemSlotBytes, st, M->d_plans, M->d_ap, M->d_items,
                       M->n_lds, M->d_counter, (int32_t)M->lds_bytes);
  }
  if (M->n_arena)
    arena_adjoint_multi(M->d_plans, M->d_ap, M->d_items + M->n_lds, M->n_arena,
                        M->d_counter + 1, M->grid_arena, st);
  for (wds_plan* P : M->plans) {
    P->cnt.bwd_solve += 1;
    P->cnt.bwd_reuse += 1;
    P->cnt.launches += 1;
  }
  return launch_check("k_adjoint_multi");
}

extern "C" int wds_newton_iter(wds_plan* P, void* stream, const wds_solve_args* a, int32_t it,
                               int32_t* ctrl) {
  int r = check_ready(P);
  if (r) return r;
  if (!a || a->B <= 0 || !ctrl || it < 1 || !a->workspace) {
    set_error("wds_newton_iter: bad argument");
    return WDS_EINVAL;
  }
  if ((P->p > 0 || P->has_cv) && !a->status_machine) {
    set_error("networks with PRVs or CV pipes need status_machine=1");
    return WDS_EUNSUPPORTED;
  }
  r = check_sched(P, a);
  if (r) return r;
  SolveParams sp = make_params(a, P->N);
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_newton_iter(P->dev, sp, it, ctrl, a->B, (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_newton_iter, lds);
    hipLaunchKernelGGL(k_newton_iter, dim3(a->B), dim3(64), lds, (hipStream_t)stream, P->dev,
                       sp, it, ctrl);
  }
  P->cnt.launches += 1;
  return launch_check("k_newton_iter");
}

extern "C" int wds_adjoint(wds_plan* P, void* stream, const wds_adjoint_args* a) {
  int r = check_ready(P);
  if (r) return r;
  if (P->has_psv_fcv) {
    set_error("the implicit adjoint of networks with PSVs or FCVs with a setting is not built");
    return WDS_EUNSUPPORTED;
  }
  if (!a || a->B <= 0 || !a->factor || !a->g_demand ||
      !a->g_res_head || !a->g_ke || !a->g_r || !a->info || !a->workspace || !a->head ||
      !a->flow || !a->emitter || !a->status || !a->ke || !a->r_hw || a->n_refine < 0) {
    set_error("wds_adjoint: bad argument");
    return WDS_EINVAL;
  }
  AdjParams ap = make_adj_params(a);
  const int64_t lds = P->stats.lds_bytes;
  if (P->arena) {
    arena_adjoint(P->dev, ap, a->B, (hipStream_t)stream);
  } else {
    WDS_SET_LDS(k_adjoint, lds);
    hipLaunchKernelGGL(k_adjoint, dim3(a->B), dim3(64 * kSolveWaves), lds, (hipStream_t)stream,
                       P->dev, ap);
  }
  P->cnt.bwd_solve += 1;
  P->cnt.bwd_reuse += 1;
  P->cnt.launches += 1;
  return launch_check("k_adjoint");
}

// ---------------------------------------------------------------------------
// Link coefficients at a given state (dgga/autodiff.py:239-362
// _link_coeffs_np): one lane per (scenario, link), the device functions the
// Newton loop and the adjoint use.  P = 1 / hgrad as assembled into the GGA
// matrix (1/CBIG closed, 0 for an ACTIVE PRV whose row is replaced),
// hloss = Y / P where P > 0 (0 otherwise), dr = d hloss / d r of an open pipe
// (0 on the RQtol clamp and at every other link).
namespace {
__global__ void k_link_coeffs(DevPlan pl, int32_t B, const double* q, const double* r,
                              int32_t r_stride, const int8_t* status, double* P_out,
                              double* hl_out, double* dr_out) {
  const int64_t n = (int64_t)B * pl.L;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / pl.L);
    const int k = (int)(i - (int64_t)b * pl.L);
    const int f = pl.l_flags[k];
    const double qk = q[i];
    const double rk = r[(int64_t)b * r_stride + k];
    const int8_t st = status[i];
    LinkPY py = link_PY(pl, k, qk, rk, st);
    if (f & LF_PCV) {
      // a PRV with a setting, as the PRV stage assembles it: 0 when ACTIVE
      // (its row is replaced), closed, or the open valve's coefficients
      if (st == ST_ACTIVE) {
        py = LinkPY{0.0, 0.0};
      } else if (st <= ST_CLOSED) {
        py = LinkPY{1.0 / CBIG, qk};
      } else {
        valve_open(pl, qk, pl.l_km_valve[k], py.P, py.Y);
      }
    }
    P_out[i] = py.P;
    if (hl_out) hl_out[i] = py.P > 0.0 ? py.Y / py.P : 0.0;
    if (!dr_out) continue;
    double dr = 0.0;
    const bool closed_now = (st <= ST_CLOSED || (f & LF_ZERO_SPEED)) && !(f & LF_PCV);
    if ((f & LF_PIPE) && !closed_now && pl.headloss != 0) {
      dr = dw_dr(pl, qk, pl.l_diam[k], pl.l_kc[k]);
    } else if ((f & LF_PIPE) && !closed_now) {
      const double sgn = qk < 0.0 ? -1.0 : 1.0;
      const double qa = qk * sgn;
      const double hgr = pl.hexp * rk * det_pow(fmax(qa, 1e-30), pl.hexp - 1.0);
      if (hgr >= pl.rqtol) dr = det_pow(qa, pl.hexp) * sgn;
    }
    dr_out[i] = dr;
  }
}
}  // namespace

extern "C" int wds_link_coeffs(wds_plan* P, void* stream, int32_t B, const double* q,
                               const double* r_hw, int32_t r_stride, const int8_t* status,
                               double* P_out, double* hloss, double* dr) {
  int rc = check_ready(P);
  if (rc) return rc;
  if (B <= 0 || !q || !r_hw || !status || !P_out || (r_stride != 0 && r_stride != P->L)) {
    set_error("wds_link_coeffs: bad argument");
    return WDS_EINVAL;
  }
  const int64_t n = (int64_t)B * P->L;
  const int64_t g = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_link_coeffs, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(256), 0,
                     (hipStream_t)stream, P->dev, B, q, r_hw, r_stride, status, P_out, hloss,
                     dr);
  P->cnt.launches += 1;
  return launch_check("k_link_coeffs");
}

// ---------------------------------------------------------------------------
// Parameter maps (dgga/calib.py:64-108, dgga/parse.py:557-568): elementwise,
// one lane per entry, grid-stride over n.
namespace {

__global__ void k_hw_resistance(int64_t n, int32_t L, const double* C, const double* len,
                                const double* diam, const int8_t* mask, double* r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % L);
    const bool on = !mask || mask[k];
    // 4.727 L / C^1.852 / D^4.871 in the reference's operation order
    r[i] = on ? 4.727 * len[k] / pow(C[i], 1.852) / pow(diam[k], 4.871) : 0.0;
  }
}

__global__ void k_hw_resistance_grad(int64_t n, const double* C, const double* r,
                                     const double* g_r, double* g_C) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    g_C[i] = g_r[i] * (-1.852 * r[i] / C[i]);  // dr/dC = -1.852 r / C (calib.py:89-108)
}

__global__ void k_emitter_ke(int64_t n, const double* C, double qexp, double coef,
                             double* ke) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double c = C[i];
    ke[i] = c > 0.0 ? coef / pow(c, qexp) : 0.0;
  }
}

__global__ void k_emitter_ke_grad(int64_t n, const double* C, const double* ke, double qexp,
                                  const double* g_ke, double* g_C) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double c = C[i];
    g_C[i] = c > 0.0 ? g_ke[i] * (-qexp * ke[i] / c) : 0.0;
  }
}

// deterministic power laws on arrays (op 0: x^y, op 1: log x)
__global__ void k_detmath(int32_t op, int64_t n, const double* x, const double* y,
                          double* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = op == 0 ? det_pow(x[i], y[i]) : det_log(x[i]);
}

dim3 ew_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return dim3((unsigned)std::min<int64_t>(std::max<int64_t>(g, 1), 4096));
}

}  // namespace

extern "C" int wds_hw_resistance(void* stream, int64_t n, int32_t L, const double* C,
                                 const double* len, const double* diam, const int8_t* mask,
                                 double* r) {
  if (n < 0 || L <= 0 || (n > 0 && (!C || !len || !diam || !r)) || n % L) {
    set_error("wds_hw_resistance: bad argument");
    return WDS_EINVAL;
  }
  if (n == 0) return WDS_OK;
  hipLaunchKernelGGL(k_hw_resistance, ew_grid(n), dim3(256), 0, (hipStream_t)stream, n, L, C,
                     len, diam, mask, r);
  return launch_check("k_hw_resistance");
}

extern "C" int wds_hw_resistance_grad(void* stream, int64_t n, const double* C, const double* r,
                                      const double* g_r, double* g_C) {
  if (n < 0 || (n > 0 && (!C || !r || !g_r || !g_C))) {
    set_error("wds_hw_resistance_grad: bad argument");
    return WDS_EINVAL;
  }
  if (n == 0) return WDS_OK;
  hipLaunchKernelGGL(k_hw_resistance_grad, ew_grid(n), dim3(256), 0, (hipStream_t)stream, n, C,
                     r, g_r, g_C);
  return launch_check("k_hw_resistance_grad");
}

struct BatchSums {
  const double* src[4];
  double* dst[4];
  int64_t n[4];
  int32_t narr, B;
};

// a block sums 16 columns of the concatenated arrays: thread (column c =
// t % 16, slice k = t / 16) adds rows k * chunk .. (k + 1) * chunk - 1 in
// order, then the 16 slice sums of a column add in slice order (a fixed
// tree: the same bits on every run)
__global__ void k_batch_sums(BatchSums a, int64_t tot) {
  __shared__ double part[16][16];
  const int c = threadIdx.x & 15, k = threadIdx.x >> 4;
  const int chunk = (a.B + 15) / 16;
  for (int64_t g0 = (int64_t)blockIdx.x * 16; g0 < tot; g0 += (int64_t)gridDim.x * 16) {
    int64_t col = g0 + c;
    int arr = -1;
    for (int i = 0; i < a.narr && col < tot; ++i) {
      if (col < a.n[i]) {
        arr = i;
        break;
      }
      col -= a.n[i];
    }
    double acc = 0.0;
    if (arr >= 0) {
      const double* sp = a.src[arr] + col;
      const int b1 = min(a.B, (k + 1) * chunk);
      const int64_t st = a.n[arr];
      int b = k * chunk;
      // sixteen rows' loads in flight at a time, added in row order
      for (; b + 16 <= b1; b += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = sp[(int64_t)(b + u) * st];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += v[u];
      }
      for (; b < b1; ++b) acc += sp[(int64_t)b * st];
    }
    part[k][c] = acc;
    __syncthreads();
    if (k == 0 && arr >= 0) {
      double s = 0.0;
      for (int q = 0; q < 16; ++q) s += part[q][c];
      a.dst[arr][col] = s;
    }
    __syncthreads();
  }
}

extern "C" int wds_batch_sums(void* stream, int32_t B, int32_t narr, const double* const* src,
                              const int64_t* n, double* const* dst) {
  if (B < 0 || narr < 0 || narr > 4 || (narr > 0 && (!src || !n || !dst))) {
    set_error("wds_batch_sums: bad argument");
    return WDS_EINVAL;
  }
  BatchSums a{};
  int64_t tot = 0;
  for (int i = 0; i < narr; ++i) {
    if (n[i] < 0 || (n[i] > 0 && (!src[i] || !dst[i]))) {
      set_error("wds_batch_sums: bad array " + std::to_string(i));
      return WDS_EINVAL;
    }
    a.src[i] = src[i];
    a.dst[i] = dst[i];
    a.n[i] = n[i];
    tot += n[i];
  }
  a.narr = narr;
  a.B = B;
  if (tot == 0) return WDS_OK;
  const int64_t blocks = std::min<int64_t>((tot + 15) / 16, 4096);
  hipLaunchKernelGGL(k_batch_sums, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a,
                     tot);
  return launch_check("k_batch_sums");
}

extern "C" int wds_emitter_ke(void* stream, int64_t n, const double* C, double qexp,
                              double coef, double* ke) {
  if (n < 0 || (n > 0 && (!C || !ke))) {
    set_error("wds_emitter_ke: bad argument");
    return WDS_EINVAL;
  }
  if (n == 0) return WDS_OK;
  hipLaunchKernelGGL(k_emitter_ke, ew_grid(n), dim3(256), 0, (hipStream_t)stream, n, C, qexp,
                     coef, ke);
  return launch_check("k_emitter_ke");
}

extern "C" int wds_emitter_ke_grad(void* stream, int64_t n, const double* C, const double* ke,
                                   double qexp, const double* g_ke, double* g_C) {
  if (n < 0 || (n > 0 && (!C || !ke || !g_ke || !g_C))) {
    set_error("wds_emitter_ke_grad: bad argument");
    return WDS_EINVAL;
  }
  if (n == 0) return WDS_OK;
  hipLaunchKernelGGL(k_emitter_ke_grad, ew_grid(n), dim3(256), 0, (hipStream_t)stream, n, C,
                     ke, qexp, g_ke, g_C);
  return launch_check("k_emitter_ke_grad");
}

extern "C" int wds_detmath_eval(void* stream, int32_t op, int64_t n, const double* x,
                                const double* y, double* out) {
  if (n < 0 || (op != 0 && op != 1) || (n > 0 && (!x || !out || (op == 0 && !y)))) {
    set_error("wds_detmath_eval: bad argument");
    return WDS_EINVAL;
  }
  if (n == 0) return WDS_OK;
  hipLaunchKernelGGL(k_detmath, ew_grid(n), dim3(256), 0, (hipStream_t)stream, op, n, x, y,
                     out);
  return launch_check("k_detmath");
}
