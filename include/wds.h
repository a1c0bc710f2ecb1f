/*
 * wds.h -- C ABI of the MI355X batched GGA library (libwds.so).
 *
 * Drop-in boundary for dgga's batched forward GGA and its implicit adjoint.
 * The reference exposes this path through Python keyword APIs, not an FFI;
 * each entry point below names the reference function it replaces.  The
 * Python host layer (wdsgpu_amd) binds these with ctypes.
 *
 * Conventions: every function returns 0 on success or a negative wds_status;
 * wds_last_error() gives the message of the last failure on this thread.
 * Device pointers are borrowed (the caller owns them, e.g. torch tensors);
 * the plan is owned by the library and released with wds_plan_destroy.
 * Arrays are row-major [B][n] with B scenarios.  Units are EPANET internal
 * units (ft, cfs).  All arithmetic is IEEE float64.
 */
#ifndef WDS_H
#define WDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wds_plan wds_plan;

enum wds_status {
  WDS_OK = 0,
  WDS_EINVAL = -1,      /* bad argument            -> ValueError */
  WDS_EUNSUPPORTED = -2,/* unsupported feature      -> NotImplementedError */
  WDS_ENOTSPD = -3,     /* non-positive pivot       -> RuntimeError naming samples */
  WDS_EUNCONVERGED = -4,/* adjoint on unconverged   -> RuntimeError */
  WDS_EDEVICE = -5,     /* HIP runtime failure      -> RuntimeError */
  WDS_ESTALE = -6       /* factor buffer not produced by this plan */
};

/* Static network description (file order, EPANET internal units).  Mirrors
 * the host precompute of GGASolver.__init__ (dgga/solver.py:160-495,
 * 1272-1390): internal status codes (CLOSED=2, OPEN=3, ACTIVE=4), internal
 * valve settings (MISSING = -1e10 for fixed valves), initial flows. */
typedef struct wds_network_desc {
  int32_t N, L;
  const int32_t* node_type;    /* [N] 0 junction, 1 reservoir, 2 tank */
  const double* elev;          /* [N] ft */
  const int32_t* link_n1;      /* [L] */
  const int32_t* link_n2;      /* [L] */
  const int32_t* link_type;    /* [L] EN_LinkType: 0 CVPIPE 1 PIPE 2 PUMP 3 PRV 4 PSV 6 FCV
                                  7 TCV (control valves need junction ends) */
  const int32_t* init_status;  /* [L] internal status code */
  const double* diam;          /* [L] ft */
  const double* km;            /* [L] pipe minor loss 0.02517 K / D^4 */
  const double* km_tcv;        /* [L] TCV loss coefficient */
  const double* km_valve;      /* [L] valve minor loss */
  const double* kc;            /* [L] Darcy-Weisbach roughness (ft) */
  const double* q0;            /* [L] initial flow (cfs) */
  const double* setting;       /* [L] internal setting; pumps: speed */
  const int32_t* pump_kind;    /* [L] -1 not a pump, 0 CONST_HP, 1 POWER_FUNC, 2 CUSTOM,
                                  3 NOCURVE */
  const double* pump_h0;       /* [L] */
  const double* pump_r;        /* [L] */
  const double* pump_n;        /* [L] (N snapped to 1 when |N-1| < 1e-6) */
  const double* pump_hmax;     /* [L] pumpstatus cut-off head (BIG for CONST_HP) */
  int32_t ntank;
  const int32_t* tank_node;    /* [ntank] */
  const double* tank_hmin;     /* [ntank] ft */
  const double* tank_hmax;     /* [ntank] ft */
  const int32_t* tank_overflow;/* [ntank] */
  int32_t headloss;            /* 0 Hazen-Williams, 1 Darcy-Weisbach */
  double hexp, qexp, rqtol, viscos, htol, qtol;
  int32_t checkfreq, maxcheck, extra_iter;
  /* CUSTOM pump curves (dgga/solver.py:1346-1370, 985-1011): the points of
   * link k are [pump_curve_ptr[k], pump_curve_ptr[k+1]), flows strictly
   * increasing, in the input's user units, converted on every evaluation
   * with ucf_flow (Ucf[FLOW]) and ucf_head (Ucf[HEAD]) as curvecoeff does
   * (hydcoeffs.c:794-831).  pump_curve_ptr may be null when no pump is
   * CUSTOM. */
  const int32_t* pump_curve_ptr;  /* [L+1] */
  const double* pump_curve_x;
  const double* pump_curve_y;
  double ucf_flow, ucf_head;
} wds_network_desc;

typedef struct wds_plan_stats {
  int32_t N, L, Nj, nnzL, nlevels, npcv;
  int64_t nnz_csr;             /* reference CSR pattern size (full symmetric) */
  int64_t nnz_lower;           /* lower pattern of A incl. diagonal */
  int64_t factor_updates;      /* multiply-adds of one numeric factorisation */
  int32_t factor_steps, fwd_steps, bwd_steps;
  int64_t factor_instr, fwd_instr, bwd_instr;
  int64_t lds_bytes;           /* LDS per scenario of the fused kernels */
  int64_t arena_bytes;         /* arena mode: global memory per scenario holding that
                                  layout (networks whose lds_bytes exceed a CU's
                                  160 KB); 0 in LDS mode */
  int32_t ordering;            /* the ordering the plan uses (0, 1 or 3 + d) */
  int32_t iteration_cost;      /* scheduler cost of one factorisation and three
                                  solve pairs (model units) */
} wds_plan_stats;

/* Counters (GGASolver.cudss_counters, dgga/solver.py:2580-2593): batched
 * numeric factorisations, triangular solve pairs, backward solves, backward
 * reuse of the forward's terminal factor, backward refactorisations. */
typedef struct wds_counters {
  int64_t factorize, solve, bwd_solve, bwd_reuse, bwd_refactorize, launches;
} wds_counters;

/* Symbolic analysis, once per topology (replaces cuDSS plan(),
 * dgga/solver.py:2474-2480, and models genmmd/etree/col_counts,
 * dgga/smatrix.py:352-547, dgga/symbolic.py:101-230).  Host only: no GPU
 * call.  ordering: 0 minimum degree, 1 natural, 2 automatic (the cheapest of
 * minimum degree and independent-set rounds with degree caps 3..6 by the
 * programs' scheduled cost, at the best LDS occupancy any of them reaches),
 * 3 + d independent-set rounds with degree cap d (0..8); | 0x100 runs the
 * kernels in arena mode (the per-scenario layout in global memory) even when
 * it would fit in LDS, which a network larger than a CU's LDS gets anyway. */
int wds_plan_create(const wds_network_desc* net, int32_t ordering, wds_plan** out);
void wds_plan_destroy(wds_plan* plan);
int wds_plan_stats_get(const wds_plan* plan, wds_plan_stats* out);

/* Copy a host-side plan array for inspection (tests): what = 0 perm,
 * 1 colptr, 2 rowidx, 3 parent, 4 node_slot, 5 csr indptr, 6 csr indices,
 * 7 csr_to_L, 8.. programs (8+3*p: step-major instr as int32 x4, +1 reserved,
 * +2 level_end; p = 0 factor, 1 fwd, 2 bwd), 17 the lane passes of the
 * symmetric rows (3 x 64 x int4: first slot, end slot, first row, rows) and
 * 18 their (pos, col) pairs in lane layout (slot t of lane l at t * 64 + l,
 * a row's last column word with bit 30 set), 19 row_slot (slot of each
 * junction row).  Returns the element count
 * (int32 units) or a negative status; copies min(count, cap). */
int64_t wds_plan_export(const wds_plan* plan, int32_t what, int32_t* dst, int64_t cap);

/* Upload the plan's static data to the current HIP device (idempotent). */
int wds_plan_upload(wds_plan* plan);

/* Device workspace the solver entry points need for B scenarios. */
int64_t wds_workspace_bytes(const wds_plan* plan, int32_t B);

/* Batched factor + solve with refinement on the reference's CSR values
 * (replaces _cudss_forward / dense cholesky + cholesky_solve + refinement,
 * dgga/solver.py:2630-2651, 3037-3064).  csr [B][nnz_csr] in the CSR pattern
 * of GGASolver (dgga/solver.py:336-369), F [B][Nj] and H [B][Nj] in original
 * junction order.  factor_out [B][nnzL] (optional) receives the factor.
 * info [B] (int32) gets 0 or the first non-positive pivot (1-based). */
int wds_factor_solve(wds_plan* plan, void* stream, int32_t B, const double* csr,
                     const double* F, double* H, int32_t n_refine,
                     double* factor_out, int32_t* info, void* workspace);

/* The two phases of wds_factor_solve on their own (the reference's cuDSS
 * factorize and solve phases, dgga/solver.py:2630-2651).  wds_factor writes
 * factor_out [B][nnzL] (the plan's internal LDL^T layout) and info [B] (0,
 * or 1 on a non-positive pivot).  wds_trisolve solves A H = F with that
 * factor plus n_refine refinement steps whose residuals use the CSR values
 * csr [B][nnz_csr]; F and H [B][Nj] in original junction order. */
int wds_factor(wds_plan* plan, void* stream, int32_t B, const double* csr,
               double* factor_out, int32_t* info, void* workspace);
int wds_trisolve(wds_plan* plan, void* stream, int32_t B, const double* factor,
                 const double* csr, const double* F, double* H, int32_t n_refine,
                 void* workspace);

/* Per-scenario inputs and outputs of the forward solve. */
typedef struct wds_solve_args {
  int32_t B;
  const double* demand;        /* [B][N] cfs */
  const double* res_head;      /* [B][N], or [N] with res_head_shared: ft at fixed-head
                                  nodes (junction entries ignored) */
  const double* ke;            /* [B][N] or [N] emitter coefficients */
  int32_t ke_stride;           /* N (per scenario) or 0 (shared) */
  const double* r_hw;          /* [B][L] or [L] resistance */
  int32_t r_stride;            /* L or 0 */
  const double* q0;            /* [B][L] warm start or NULL (initial flow) */
  const double* e0;            /* [B][N] warm start or NULL (1 at emitters) */
  int32_t max_iter;
  double accuracy;
  int32_t status_machine;      /* 1: batched status machine (dgga/solver.py:3126-3179) */
  int32_t n_refine;            /* refinement steps per linear solve (reference: 2) */
  int32_t polish_steps;        /* >=0: that many reduced Newton polish steps after
                                  convergence, then the terminal factorisation
                                  (kept in factor when non-NULL) and the residual
                                  (dgga/autodiff.py:1053-1104); <0: none */
  /* outputs */
  double* head;                /* [B][N] */
  double* flow;                /* [B][L] */
  double* emitter;             /* [B][N] */
  int32_t* iters;              /* [B] */
  double* relerr;              /* [B] */
  int8_t* status;              /* [B][L] internal status codes */
  double* diag_ratio;          /* [B] or NULL */
  int32_t* info;               /* [B] 0 ok, >0 non-SPD at that iteration, <0 polish failure */
  double* resid;               /* [B] polish residual inf-norm or NULL */
  double* factor;              /* [B][nnzL] terminal factor for the adjoint, or NULL */
  void* workspace;             /* wds_workspace_bytes(plan, B) */
  const int8_t* status0;       /* [B][L] initial internal link statuses (a warm start
                                  between extended-period frames) or NULL (the
                                  network's initial statuses) */
  /* Status-machine schedule (record_schedule, dgga/solver.py:2740-2745,
   * 3136-3160), all NULL when not recorded.  sched_flags [B][maxtrials]:
   * bit0 the scenario ran iteration it (index it-1), bit1 linkstatus ran on
   * the convergence branch, bit2 on the periodic branch.  sched_events
   * [B][sched_cap]: one int32 per status change, (it << 20) | (kind << 19)
   * | link with kind 0 prvstatus, 1 linkstatus; sched_count [B] counts the
   * changes (entries past sched_cap are dropped, the count keeps growing).
   * Needs status_machine and maxtrials < 2048. */
  uint8_t* sched_flags;
  int32_t* sched_events;
  int32_t* sched_count;
  int32_t sched_cap;
  int32_t res_head_shared;     /* 1: res_head is one [N] vector for every scenario */
} wds_solve_args;

/* Fused forward: the whole Newton loop of every scenario in one persistent
 * launch, one wavefront per scenario with its factor resident in LDS
 * (replaces GGASolver.solve mode='dense', dgga/solver.py:2703-3212, and
 * ImplicitGGASolveGPU.forward's polish + terminal factor,
 * dgga/autodiff.py:1171-1243). */
int wds_solve(wds_plan* plan, void* stream, const wds_solve_args* args);

/* wds_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  args->workspace must be NULL: the workspace is a file-scope
 * device array of WDS_NOALLOC_SCRATCH_BYTES (32 MiB) that every GPU the
 * library runs on holds until the process ends; calls needing more
 * (wds_workspace_bytes(plan, B) > WDS_NOALLOC_SCRATCH_BYTES: about 230
 * scenarios of a 782-junction network) fail with WDS_EINVAL.  The plan must
 * already be uploaded to the current device (wds_plan_upload allocates).
 * All calls on one GPU share that GPU's array, and its contents are those of
 * the last call: calls that overlap on one GPU (different streams, threads
 * or plans) race and must be serialised by the caller.  Otherwise identical
 * to wds_solve, bit for bit. */
#define WDS_NOALLOC_SCRATCH_BYTES (32ll << 20)
int wds_solve_noalloc(wds_plan* plan, void* stream, const wds_solve_args* args);

/* One Newton iteration of every active scenario with the state kept in
 * global memory (the unfused path: one launch per iteration, host loop).
 * state: q [B][L], e [B][N] (by node), head [B][N], status [B][L],
 * ctrl [B][4] int32 (active, iters, nextcheck, info), relerr [B].
 * it is the 1-based iteration number. */
int wds_newton_iter(wds_plan* plan, void* stream, const wds_solve_args* args,
                    int32_t it, int32_t* ctrl);

/* Backward of the implicit solve (replaces ImplicitGGASolveGPU.backward,
 * dgga/autodiff.py:1245-1312, and the theta contraction of the CPU adjoint
 * ImplicitGGASolve.backward, autodiff.py:658-732): reuses the forward's
 * terminal factor.  gH/gE [B][N], gQ [B][L] (each may be NULL: a zero
 * cotangent); outputs g_demand/g_res_head/g_ke [B][N], g_r [B][L] and, when
 * g_pump is not NULL, g_pump [B][3][L]: d loss / d (speed, pump_h0, pump_r)
 * of every pump link (pumpcoeff derivatives of _pump_coeffs_np,
 * autodiff.py:122-218), 0 at other links. */
typedef struct wds_adjoint_args {
  int32_t B;
  const double* demand;        /* [B][N] */
  const double* ke;            /* [B][N] or [N] */
  int32_t ke_stride;
  const double* r_hw;          /* [B][L] or [L] */
  int32_t r_stride;
  const double* head;          /* [B][N] converged heads */
  const double* flow;          /* [B][L] */
  const double* emitter;       /* [B][N] */
  const int8_t* status;        /* [B][L] */
  const double* factor;        /* [B][nnzL] from wds_solve */
  const double* gH;            /* [B][N] */
  const double* gQ;            /* [B][L] */
  const double* gE;            /* [B][N] */
  int32_t n_refine;
  double* g_demand;            /* [B][N] */
  double* g_res_head;          /* [B][N] */
  double* g_ke;                /* [B][N] */
  double* g_r;                 /* [B][L] */
  int32_t* info;               /* [B] */
  void* workspace;
  double* g_pump;              /* [B][3][L] or NULL */
} wds_adjoint_args;

int wds_adjoint(wds_plan* plan, void* stream, const wds_adjoint_args* args);

/* Link coefficients at a given state, as the kernels compute them
 * (replaces _link_coeffs_np, dgga/autodiff.py:239-362, for the consumers
 * that read the converged link gradients: the Schur-diagonal preconditioner
 * of dgga/optim2.py:137-215 and the RQtol clamp mask of dgga/calib.py:264-300).
 * q, status [B][L]; r_hw [B][L] (r_stride L) or [L] (r_stride 0).  P_out
 * [B][L] = 1 / hgrad as assembled into the GGA matrix (1/CBIG for a closed
 * link, 0 for an ACTIVE PRV, whose row is replaced); hloss [B][L] (optional)
 * the head loss;
 * dr [B][L] (optional) d hloss / d r of an open pipe, 0 on the RQtol clamp
 * and at other links.  The pump parameters are the plan's current ones. */
int wds_link_coeffs(wds_plan* plan, void* stream, int32_t B, const double* q,
                    const double* r_hw, int32_t r_stride, const int8_t* status,
                    double* P_out, double* hloss, double* dr);

/* Replace the pump parameters of the plan's pump links, batch-shared, as the
 * reference's solve_polished(speed=, pump_h0=, pump_r=) does for a call
 * (dgga/autodiff.py:539-603 with _override_pump): each array is [L] indexed
 * by link and read at the pump links only, NULL keeps the parameter.
 * speed: relative speed (LinkSetting; 0 closes the pump); h0, r: curve
 * coefficients (internal units); hmax: the pumpstatus cut-off head
 * speed^2 Hmax; q0: the initial flow speed Q0.  Stream-ordered: work queued
 * on `stream` after this call sees the new values, work queued before it
 * the old ones. */
int wds_plan_set_pumps(wds_plan* plan, void* stream, const double* speed, const double* h0,
                       const double* r, const double* hmax, const double* q0);

/* Ragged batches: several networks, each with its own plan, scenario count
 * and arrays, solved in one persistent launch per kernel mode.  No
 * counterpart in the reference (whose solver serves one network per call,
 * dgga/solver.py:2703-3212); it serves workloads that solve many networks
 * at once.  wds_multi_create uploads the plans and fixes, per network, the
 * kernel mode: LDS mode when its layout (wds_plan_stats.lds_bytes) is at
 * most lds_budget bytes, arena mode otherwise (lds_budget <= 0: every
 * network in arena mode), and the work list of (network, scenario) items,
 * costliest networks first.  Each network's args (B[net] scenarios, its own
 * workspace of wds_multi_workspace_bytes(multi, net) bytes) are those of
 * wds_solve / wds_adjoint; the results are the same bits as per-network
 * calls.  wds_multi_info reports the scenarios in each mode and the LDS
 * launch's dynamic LDS. */
typedef struct wds_multi wds_multi;
int wds_multi_create(wds_plan* const* plans, int32_t nnet, const int32_t* B,
                     int64_t lds_budget, wds_multi** out);
void wds_multi_destroy(wds_multi* multi);
int64_t wds_multi_workspace_bytes(const wds_multi* multi, int32_t net);
int wds_multi_info(const wds_multi* multi, int32_t* n_lds, int32_t* n_arena, int64_t* lds_bytes);
int wds_multi_solve(wds_multi* multi, void* stream, const wds_solve_args* args /* [nnet] */);
int wds_multi_adjoint(wds_multi* multi, void* stream, const wds_adjoint_args* args /* [nnet] */);

int wds_counters_get(wds_plan* plan, wds_counters* out, int32_t reset);
const char* wds_last_error(void);

/* Parameter maps of the calibration path, elementwise on the stream.
 * Hazen-Williams resistance r = 4.727 L / C^1.852 / D^4.871 (replaces
 * hw_resistance, dgga/calib.py:64-86) over n = B*L entries of C with len/diam
 * [L] and an optional [L] mask (0 -> r = 0); its gradient
 * g_C = g_r * (-1.852 r / C) (dr_dC, dgga/calib.py:89-108).
 * Emitter coefficient Ke = coef / C^qexp for C > 0, else 0, with
 * coef = Ucf_Q^qexp / Ucf_P (ke_int_of_C, dgga/parse.py:557-568) and
 * g_C = g_ke * (-qexp Ke / C). */
int wds_hw_resistance(void* stream, int64_t n, int32_t L, const double* C, const double* len,
                      const double* diam, const int8_t* mask, double* r);
int wds_hw_resistance_grad(void* stream, int64_t n, const double* C, const double* r,
                           const double* g_r, double* g_C);
int wds_emitter_ke(void* stream, int64_t n, const double* C, double qexp, double coef,
                   double* ke);
int wds_emitter_ke_grad(void* stream, int64_t n, const double* C, const double* ke,
                        double qexp, const double* g_ke, double* g_C);

/* Column sums over the batch, src[i] [B][n[i]] -> dst[i] [n[i]] for up to
 * four arrays in one launch (each column in 16 consecutive batch slices,
 * each slice in batch order, the slice sums in slice order): the
 * gradients of the batch-shared inputs of an implicit solve (reservoir
 * heads, emitter coefficients, resistances, pump parameters), which the
 * reference leaves to autograd's expand backward (a sum over dim 0 per input,
 * dgga/autodiff.py:1305-1312 returning per-scenario gradients). */
int wds_batch_sums(void* stream, int32_t B, int32_t narr, const double* const* src,
                   const int64_t* n, double* const* dst);

/* The deterministic power laws the kernels use in place of libm pow/log
 * (the reference's calls at dgga/solver.py:58-69, 2176-2377), elementwise on
 * device arrays: op 0 out = x^y, op 1 out = log x (y unused).  For checking
 * the kernels' arithmetic against the CPU checker bit for bit. */
int wds_detmath_eval(void* stream, int32_t op, int64_t n, const double* x, const double* y,
                     double* out);

#ifdef __cplusplus
}
#endif
#endif /* WDS_H */
