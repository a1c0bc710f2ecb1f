// Device building blocks shared by the GGA kernels (one wavefront = one
// scenario; every function below is called by all 64 lanes of the wave).
#pragma once
#include <hip/hip_runtime.h>

#include "detmath.h"
#include "devplan.h"
#include "symbolic.h"

namespace wds {

constexpr double CSMALL = 1.0e-6;
constexpr double CBIG = 1.0e8;
constexpr int8_t ST_XHEAD = 0, ST_TEMPCLOSED = 1, ST_CLOSED = 2, ST_OPEN = 3, ST_ACTIVE = 4,
                 ST_XFCV = 6, ST_XPRESSURE = 7;
// Darcy-Weisbach friction-factor constants (EPANET hydcoeffs.c; quoted in
// dgga/solver.py:40-45)
constexpr double DW_A1 = 3.14159265358979323850e+03;
constexpr double DW_A2 = 1.57079632679489661930e+03;
constexpr double DW_A8 = 4.61841319859066668690e+00;
constexpr double DW_A9 = -8.68588963806503655300e-01;
constexpr double DW_AB = 3.28895476345399058690e-03;
constexpr double DW_AC = -5.14214965799093883760e-03;
constexpr double PI_EPANET = 3.141592654;

// One wavefront per scenario whose LDS share admits a single wavefront per
// SIMD: let the compiler schedule for latency with the whole register file
// instead of squeezing registers for an occupancy the LDS rules out.
// (Arena mode keeps the same bound: at two wavefronts per SIMD the fused
// forward would spill to scratch.)
#define WDS_SCENARIO_KERNEL __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
// The fused forward runs two wavefronts per scenario (a 128-lane workgroup):
// the second shares the link, row and node loops of the Newton iteration, and
// a SIMD holds two wavefronts, so one hides the other's memory waits.  The
// register budget is then half the file.
constexpr int kSolveWaves = 2;
#define WDS_SOLVE_KERNEL __launch_bounds__(64 * kSolveWaves) \
    __attribute__((amdgpu_waves_per_eu(kSolveWaves, kSolveWaves)))

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
// the wavefront's index in its scenario's workgroup
__device__ __forceinline__ int wave_idx() { return (int)(threadIdx.x >> 6); }
// barrier of the scenario's workgroup: the LDS and global writes of every
// wavefront before it are visible to every wavefront after it
__device__ __forceinline__ void wg_sync() { __syncthreads(); }

// Make this wave's global stores visible to its own later loads.
__device__ __forceinline__ void mem_sync() {
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

// Two builds of the same kernels (the WDS_K names differ):
//  * LDS mode: a scenario's factor, head and solve vectors, sinks and
//    Woodbury block live in the workgroup's dynamic LDS;
//  * arena mode (WDS_GLOBAL_ARENA), for networks whose scenario exceeds the
//    CU's LDS: the same layout, at the same byte offsets, in a per-scenario
//    arena at the end of its global workspace slice.  The program words
//    address it by byte offset from the arena base.
#ifndef WDS_GLOBAL_ARENA
#define WDS_K(name) name
// Order LDS traffic between program levels of one wavefront: wait for the
// wave's outstanding LDS operations only (lgkmcnt(0); vector-memory loads of
// prefetched instructions stay in flight) and stop the compiler from moving
// memory operations across.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
}

// Order LDS traffic of one wavefront without waiting: the hardware performs a
// wave's LDS operations in issue order, so a wavefront-scope fence only has
// to stop the compiler from reordering (it emits no s_waitcnt).  Used
// between the levels of the sparse programs, where the next level's loads
// only need to see this wave's own stores.
__device__ __forceinline__ void lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
#else
#define WDS_K(name) name##_arena
// arena mode: the scenario's vectors are global memory; a wave's stores are
// complete and visible to its own later loads after vmcnt(0)
__device__ __forceinline__ void lds_sync() { mem_sync(); }
__device__ __forceinline__ void lds_order() { mem_sync(); }
#endif

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ bool wave_any(bool p) { return __any(p); }

// v of the lane the DPP control CTRL selects, for a double (both halves)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v of lane ^ 16 and of lane ^ 32: v_permlane16_swap / v_permlane32_swap
// exchange rows between two registers; with both holding v, one output holds
// the partner row for the even rows (halves), the other for the odd ones.
__device__ __forceinline__ double xor_lanes16(double v) {
  const long long b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
                                                   false, false);
  const bool odd = lane_id() & 16;
  const unsigned l = odd ? lo[0] : lo[1], h = odd ? hi[0] : hi[1];
  return __longlong_as_double(((long long)h << 32) | l);
}
__device__ __forceinline__ double xor_lanes32(double v) {
  const long long b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32),
                                                   false, false);
  const bool upper = lane_id() & 32;
  const unsigned l = upper ? lo[0] : lo[1], h = upper ? hi[0] : hi[1];
  return __longlong_as_double(((long long)h << 32) | l);
}

// Sum over aligned lane groups of per-lane width 2^lw (lw = 0: not in a
// group, the value is left alone; lwm: the widest group's lw, the number of
// stages), bitwise equal in every grouped lane to an xor butterfly over its
// group.  The stages within 16 lanes use DPP: quad permutes for xor 1 and
// xor 2, then row half-mirror and row mirror, which pair a lane with one of
// the partner half; after each stage every lane of a block holds the same
// bits (the additions are commutative), so the partner's value equals the
// butterfly's.  A lane adds only in the stages below its own group width,
// and a partner in those stages lies inside the same aligned group.  The
// stages across rows use the permlane swaps.
__device__ __forceinline__ double group_sum(double v, int lw, int lwm) {
  // nested so that each further stage is a fall-through
  double p = dpp_f64<0xB1>(v);                 // quad_perm [1, 0, 3, 2]
  v = lw >= 1 ? v + p : v;
  if (lwm >= 2) {
    p = dpp_f64<0x4E>(v);                      // quad_perm [2, 3, 0, 1]
    v = lw >= 2 ? v + p : v;
    if (lwm >= 3) {
      p = dpp_f64<0x141>(v);                   // row_half_mirror
      v = lw >= 3 ? v + p : v;
      if (lwm >= 4) {
        p = dpp_f64<0x140>(v);                 // row_mirror
        v = lw >= 4 ? v + p : v;
        if (lwm >= 5) {
          p = xor_lanes16(v);
          v = lw >= 5 ? v + p : v;
          if (lwm >= 6) {
            p = xor_lanes32(v);
            v = lw >= 6 ? v + p : v;
          }
        }
      }
    }
  }
  return v;
}

// ---------------------------------------------------------------------------
// Level-scheduled program executors, one step per pass.
// Instruction words are prefetched kRing steps ahead into a ring of kRing + 1
// register slots: step g lives in slot g % (kRing + 1), and the load of step
// g + kRing goes into the slot step g - 1 just released, so no slot is
// copied.  Instruction addresses depend on nothing loaded; the programs run
// in whole turns of the ring (NOP padded) and the instruction array carries
// kRing more NOP steps, so the ring can over-fetch without bounds checks.
// (A guard per step instead of the padding costs more: the wait-count pass
// cannot tell that a skipped step ends the loop and drains the whole
// prefetch at every turn.)
//
// The operands of step g + 1 are read while step g finishes: inside a level
// the tasks are independent, so when step g does not end its level, step
// g + 1's LDS reads are issued before step g's store (the wavefront performs
// its LDS operations in issue order, and no read of a level whose value is
// used touches a store of the same level: only a lane's own sink slot is
// both stored and read, and a sink read is discarded); after a level end
// they are issued after the level's last stores.
constexpr int kRing = 4;
constexpr int kSlots = kRing + 1;

// 1/x for a positive pivot: v_rcp_f64 refined by two Newton steps
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Explicit address spaces for the executor: LDS operands compile to ds_*
// (counted by lgkmcnt only) and instruction words to global loads (vmcnt
// only), so waiting for an operand never drains the instruction prefetch, as
// it would with flat accesses through generic pointers.
#ifndef WDS_GLOBAL_ARENA
typedef __attribute__((address_space(3))) double lds_f64;
#else
typedef __attribute__((address_space(1))) double lds_f64;
#endif
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) i32x4 glb_i32x4;

// The ring's first kRing steps, issued oldest first.  The scheduling
// barriers keep that order: the wait counter is in-order, so if the
// scheduler issued the first step's load last, the loop header would have
// to wait for all but the newest few loads and the prefetch would collapse.
__device__ __forceinline__ void fill_ring(i32x4 (&ring)[kSlots], const glb_i32x4* ip) {
#pragma unroll
  for (int d = 0; d < kRing; ++d) {
    ring[d] = ip[(int64_t)d * 64];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Numeric LDL^T factorisation in place in V: unscaled columns w (w_ij =
// L_ij D_j) and reciprocal pivots on the diagonal; update
// w_ij -= w_ik * w_jk * invD_k.
__device__ inline void run_factor(const ProgDev& pg, double* Vg, int& bad,
                                  double* prof = nullptr) {
  const int lane = lane_id();
  const int ns = pg.nsteps;
  if (ns == 0) return;
  lds_f64* const V = (lds_f64*)Vg;
  // read once: stores through V could alias a generic ProgDev reference;
  // each lane has its own sink slot
  lds_f64* const dummy = V + pg.dummy + lane;
  const glb_i32x4* ip = (const glb_i32x4*)pg.instr + lane;
  i32x4 ring[kSlots];
  fill_ring(ring, ip);
  double acc = 0.0;
  // grouped tasks: the target this lane's slice belongs to
  int fin_t = 0;
  bool fin_diag = false;
#ifdef WDS_PROFILE_STAGES
  long long prof_t0 = clock64();
  int prof_lvl = 0;
#endif
  double va = V[ring[0].x], vb = V[ring[0].y], vc = V[ring[0].z], vt = V[ring[0].w & kTMask];
  for (int g0 = 0; g0 < ns; g0 += kSlots) {
#pragma unroll
    for (int d = 0; d < kSlots; ++d) {
      const int g = g0 + d;  // ns is a multiple of kSlots
      {
        const i32x4 cur = ring[d];
        ring[(d + kRing) % kSlots] = ip[(int64_t)(g + kRing) * 64];
        const i32x4 nxt = ring[(d + 1) % kSlots];
        const int w = cur.w;
        const int t = w & kTMask;
        const bool first = w & kFirst, last = w & kLast, diag = w & kDiag;
        if (first) {
          fin_t = t;
          fin_diag = diag;
        }
        // a slice other than its group's first starts at zero
        const double init = (w & kZeroInit) ? 0.0 : vt;
        acc = fma(-va, vb * vc, first ? init : acc);
        const double r = recip(acc);
        if (diag && last && !(acc > 0.0)) bad = 1;
        // branch-free store: to the target when the instruction is a task's
        // last, else to the lane's sink
        lds_f64* const dst = last ? V + t : dummy;
        const double val = diag ? r : acc;
        const int f = __builtin_amdgcn_readfirstlane(w);
        if (!(f & kLevelEnd)) {
          va = V[nxt.x];
          vb = V[nxt.y];
          vc = V[nxt.z];
          vt = V[nxt.w & kTMask];
          *dst = val;
        } else {
          *dst = val;
          // lane 0 is in the level's widest group
          const int lwm = (f >> kWidthShift) & 7;
          if (lwm) {
            const int lw = (w >> kWidthShift) & 7;
            const double red = group_sum(acc, lw, lwm);
            if (lw && (lane & ((1 << lw) - 1)) == 0) {
              if (fin_diag && !(red > 0.0)) bad = 1;
              V[fin_t] = fin_diag ? recip(red) : red;
            }
          }
          lds_order();
#ifdef WDS_PROFILE_STAGES
          if (prof) {
            const long long t1 = clock64();
            if (lane == 0 && prof_lvl < 32) prof[prof_lvl] += (double)(t1 - prof_t0);
            prof_t0 = t1;
            ++prof_lvl;
          }
#endif
          va = V[nxt.x];
          vb = V[nxt.y];
          vc = V[nxt.z];
          vt = V[nxt.w & kTMask];
        }
      }
    }
  }
}

// The scenario's double at a byte address: LDS mode, from the start of the
// dynamic LDS (address 0; base unused); arena mode, from the arena base.
#ifndef WDS_GLOBAL_ARENA
__device__ __forceinline__ lds_f64& lds_at(const char* base, int byte_addr) {
  (void)base;
  return *(lds_f64*)(size_t)(unsigned)byte_addr;
}
#else
__device__ __forceinline__ lds_f64& lds_at(const char* base, int byte_addr) {
  return *(lds_f64*)(base + (unsigned)byte_addr);
}
#endif

// Triangular solve with the unit factor on the solve vector S (LDS):
// acc = S[t] - sum V[a] S[b] per task, stored back into S[t].  Instruction
// words are LDS byte addresses (a, b, init, store): init is S[t] on a task's
// first step (the zero slot for a group slice other than the group's first,
// the lane's own sink slot on every other step) and store is S[t] on its
// last step (the lane's sink otherwise).  The store word of a level's last
// step carries the level end and the lane's log2 group width in its high
// bits.
// The ring of a program's first kRing steps, loaded ahead of its run (by the
// caller, before other work whose time covers the loads' latency); nothing
// for an empty program.
__device__ __forceinline__ void prefetch_prog(const ProgDev& pg, i32x4 (&ring)[kSlots]) {
  if (pg.nsteps > 0) fill_ring(ring, (const glb_i32x4*)pg.instr + lane_id());
}

// run_dot with the ring prefetched (prefetch_prog of the same program)
__device__ inline void run_dot(const ProgDev& pg, const double* arena, i32x4 (&ring)[kSlots],
                               double* prof = nullptr) {
  const char* ab = (const char*)arena;
  const int lane = lane_id();
  const int ns = pg.nsteps;
  if (ns == 0) return;
  const int sent = pg.dummy + 8 * lane;  // byte address of this lane's sink slot
  const glb_i32x4* ip = (const glb_i32x4*)pg.instr + lane;
  double acc = 0.0;
  int fin = sent;  // grouped tasks: byte address of the slice's target
#ifdef WDS_PROFILE_STAGES
  long long prof_t0 = clock64();
  int prof_lvl = 0;
#endif
  double va = lds_at(ab, ring[0].x), vb = lds_at(ab, ring[0].y), vi = lds_at(ab, ring[0].z);
  for (int g0 = 0; g0 < ns; g0 += kSlots) {
#pragma unroll
    for (int d = 0; d < kSlots; ++d) {
      const int g = g0 + d;  // ns is a multiple of kSlots
      {
        const i32x4 cur = ring[d];
        ring[(d + kRing) % kSlots] = ip[(int64_t)(g + kRing) * 64];
        const i32x4 nxt = ring[(d + 1) % kSlots];
        const bool init = cur.z != sent;
        if (init) fin = cur.z;
        acc = fma(-va, vb, init ? vi : acc);
        const int f = __builtin_amdgcn_readfirstlane(cur.w);
        if (!(f & kDotLevelEnd)) {
          va = lds_at(ab, nxt.x);
          vb = lds_at(ab, nxt.y);
          vi = lds_at(ab, nxt.z);
          lds_at(ab, cur.w) = acc;
        } else {
          lds_at(ab, cur.w & kDotAddrMask) = acc;
          const int lwm = (f >> kDotWidthShift) & 7;
          if (lwm) {
            const int lw = (cur.w >> kDotWidthShift) & 7;
            const double red = group_sum(acc, lw, lwm);
            if (lw && (lane & ((1 << lw) - 1)) == 0) lds_at(ab, fin) = red;
          }
          lds_order();
#ifdef WDS_PROFILE_STAGES
          if (prof) {
            const long long t1 = clock64();
            if (lane == 0 && prof_lvl < 32) prof[prof_lvl] += (double)(t1 - prof_t0);
            prof_t0 = t1;
            ++prof_lvl;
          }
#endif
          va = lds_at(ab, nxt.x);
          vb = lds_at(ab, nxt.y);
          vi = lds_at(ab, nxt.z);
        }
      }
    }
  }
}

// b - (A x)_i evaluated in double-double: every product is split exactly
// with an FMA (TwoProduct) and the running sum is carried as an unevaluated
// pair with Knuth's TwoSum, so the residual is correct to about one rounding
// of its own magnitude.  Iterative refinement with this residual converges
// to the correctly rounded solution while kappa * eps < 1.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// R_j *= V[diag_pos[j]] (the reciprocal pivots) for every row, V and R in
// LDS: the pivot positions of up to kRowG rows per lane load before use.
constexpr int kRowG = 8;
constexpr int kRowG8 = 4;   // rows per lane per round where each row loads several values
// The pivot positions of the first kRowG rows per lane, loaded ahead (before
// the forward program runs) so that the scaling after it waits on no memory.
__device__ __forceinline__ void pivots_prefetch(const DevPlan& pl, int (&dp)[kRowG]) {
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < kRowG; ++u) dp[u] = pl.diag_pos[min(lane + 64 * u, pl.Nj - 1)];
}
__device__ inline void scale_by_pivots(const DevPlan& pl, const double* V, double* R,
                                       const int (&dp0)[kRowG]) {
  const int lane = lane_id();
  const lds_f64* Vl = (const lds_f64*)V;
  lds_f64* Rl = (lds_f64*)R;
#pragma unroll
  for (int u = 0; u < kRowG; ++u) {
    const int j = lane + 64 * u;
    if (j < pl.Nj) Rl[j] *= Vl[dp0[u]];
  }
  for (int j0 = lane + 64 * kRowG; j0 < pl.Nj; j0 += 64 * kRowG) {
    int dp[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) dp[u] = pl.diag_pos[min(j0 + 64 * u, pl.Nj - 1)];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      const int j = j0 + 64 * u;
      if (j < pl.Nj) Rl[j] *= Vl[dp[u]];
    }
  }
}

// Every junction row at once over the lane layout rs_ell, in whose slots A
// itself is stored.  Lane l walks its rows l, l + 64, ... as one packed
// sequence of slots (each row's entries in column order, the last one with
// kRowEndFlag in its column word), so a lane carries no padding between its
// rows; a pass (all of a lane's rows, or one of the two halves a pair of
// wavefronts takes, rs_share) runs to the longest lane's end, rs_len.  A
// sweep loads kSweep slots at a time (column words and matrix values,
// unconditionally from clamped slots, so that no load waits for another),
// then walks them in order, closing a row at its flagged entry with the
// next row's first value read one row ahead.  Slots past a lane's range
// read as column -1 and contribute exact zeros (p = 0: two_sum(hi, -0)
// returns hi and error 0).
constexpr int kSweep = 16;

// column words and matrix values of slots t0 + s0 .. t0 + s0 + kSweep - 1 of
// this lane (column -1 from t1 on; the loads come from slots clamped into
// the lane's range)
__device__ __forceinline__ void sweep_load(const DevPlan& pl, const double* A, int t0, int s0,
                                           int t1, int (&cl)[kSweep], double (&av)[kSweep]) {
  const int lane = lane_id();
  const int* cw = (const int*)pl.rs_ell;
#pragma unroll
  for (int u = 0; u < kSweep; ++u) {
    const int t = t0 + s0 + u;
    const int slot = max(min(t, t1 - 1), 0) * 64 + lane;
    const int w = cw[2 * slot + 1];
    cl[u] = t < t1 ? w : -1;
    av[u] = A[slot];
  }
}

// vector entries of a loaded chunk; entries past a lane's range read as zeros
__device__ __forceinline__ void sweep_values(const lds_f64* X, const int (&cl)[kSweep],
                                             double (&av)[kSweep], double (&xv)[kSweep]) {
#pragma unroll
  for (int u = 0; u < kSweep; ++u) xv[u] = X[cl[u] >= 0 ? (cl[u] & kColMask) : 0];
#pragma unroll
  for (int u = 0; u < kSweep; ++u) {
    const bool ok = cl[u] >= 0;
    av[u] = ok ? av[u] : 0.0;
    xv[u] = ok ? xv[u] : 0.0;
  }
}

// the last entry of a row (never a slot past the lane's range)
__device__ __forceinline__ bool row_end(int cl) { return cl >= 0 && (cl & kRowEndFlag); }

// The first row block of share i of nw over the row blocks of a row-block
// table whose block b starts at slot blk[b] (blk[nb] = E): the smallest b
// with blk[b] nw >= i E, so the shares hold about E / nw slots each.
__device__ __forceinline__ int block_split(const int32_t* blk, int nb, int E, int i, int nw) {
  if (i <= 0) return 0;
  if (i >= nw) return nb;
  int b = 0;
  while (b < nb && (int64_t)blk[b] * nw < (int64_t)i * E) ++b;
  return b;
}

// dst_j = b_j - (A x)_j in double-double, x and dst in LDS, b global; with
// nw = 2 wavefronts each takes one half of every lane's rows
__device__ inline void residual_all(const DevPlan& pl, const double* A, const double* x,
                                    const double* b, double* dst, int w = 0, int nw = 1) {
  const int lane = lane_id();
  const lds_f64* X = (const lds_f64*)x;
  lds_f64* D = (lds_f64*)dst;
  const int c = nw == 1 ? 0 : 1 + w;
  const int4 sh = pl.rs_share[c * 64 + lane];  // (t0, t1, first row, rows)
  const int T = pl.rs_len[c];
  const int jl = pl.Nj - 1;
  int cl[kSweep];
  double av[kSweep];
  sweep_load(pl, A, sh.x, 0, sh.y, cl, av);
  __builtin_amdgcn_sched_barrier(0);
  // D = b on this lane's rows of the pass (each lane copies and later reads
  // only its own rows)
  for (int r0 = 0; r0 < sh.w; r0 += kRowG) {
    double bv[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) bv[u] = b[min(sh.z + 64 * (r0 + u), jl)];
#pragma unroll
    for (int u = 0; u < kRowG; ++u)
      if (r0 + u < sh.w) D[sh.z + 64 * (r0 + u)] = bv[u];
  }
  int j = sh.z;
  double hi = D[min(j, jl)], lo = 0.0;
  // the next row's first value; past the lane's rows it is never used
  double hn = D[min(j + 64, jl)];
  for (int s0 = 0; s0 < T; s0 += kSweep) {
    double xv[kSweep];
    sweep_values(X, cl, av, xv);
#pragma unroll
    for (int u = 0; u < kSweep; ++u) {
      const double p = av[u] * xv[u];
      const double pe = fma(av[u], xv[u], -p);
      double sm, err;
      two_sum(hi, -p, sm, err);
      hi = sm;
      lo += err - pe;
      if (row_end(cl[u])) {
        D[j] = hi + lo;
        j += 64;
        hi = hn;
        lo = 0.0;
        hn = D[min(j + 64, jl)];
      }
    }
    if (s0 + kSweep < T) sweep_load(pl, A, sh.x, s0 + kSweep, sh.y, cl, av);
  }
}

// dst = A x, x in LDS, dst in global memory
__device__ inline void spmv_all(const DevPlan& pl, const double* A, const double* x,
                                double* dst) {
  const int lane = lane_id();
  const lds_f64* X = (const lds_f64*)x;
  __attribute__((address_space(1))) double* D = (__attribute__((address_space(1))) double*)dst;
  const int4 sh = pl.rs_share[lane];
  const int T = pl.rs_len[0];
  int cl[kSweep];
  double av[kSweep];
  sweep_load(pl, A, sh.x, 0, sh.y, cl, av);
  int j = sh.z;
  double sum = 0.0;
  for (int s0 = 0; s0 < T; s0 += kSweep) {
    double xv[kSweep];
    sweep_values(X, cl, av, xv);
#pragma unroll
    for (int u = 0; u < kSweep; ++u) {
      sum = fma(av[u], xv[u], sum);
      if (row_end(cl[u])) {
        D[j] = sum;
        j += 64;
        sum = 0.0;
      }
    }
    if (s0 + kSweep < T) sweep_load(pl, A, sh.x, s0 + kSweep, sh.y, cl, av);
  }
}

// ---------------------------------------------------------------------------
// Link coefficients (headlosscoeffs, dgga/solver.py:2924-2956 and the
// pipe/TCV/pump/valve helpers 2176-2326, 834-848; D-W 1404-1450).
struct LinkPY {
  double P, Y;
};

__device__ __forceinline__ void hw_pipe(const DevPlan& pl, double q, double r, double km,
                                        double& hloss, double& hgrad) {
  const double sgn = q < 0.0 ? -1.0 : 1.0;
  const double qa = q * sgn;
  double hg = pl.hexp * r * det_pow(fmax(qa, 1e-30), pl.hexp - 1.0);
  const bool lin = hg < pl.rqtol;
  if (lin) hg = pl.rqtol;
  double hl = lin ? hg * qa : hg * qa / pl.hexp;
  if (km > 0.0) {
    hl = hl + km * qa * qa;
    hg = hg + 2.0 * km * qa;
  }
  hloss = hl * sgn;
  hgrad = hg;
}

// Friction factor f and df/d|q| of a turbulent or transitional D-W pipe
// (w = |q| / s >= A2): Swamee-Jain above A1, Dunlop's cubic below
// (DWpipecoeff, hydcoeffs.c:640-668).
__device__ __forceinline__ void dw_friction(double qa, double s, double e, double& f,
                                            double& dfdq) {
  const double w = qa / s;
  if (w >= DW_A1) {
    const double y1 = DW_A8 / det_pow(w, 0.9);
    const double y2 = e / 3.7 + y1;
    const double y3 = DW_A9 * det_log(y2);
    f = 1.0 / (y3 * y3);
    dfdq = 1.8 * f * y1 * DW_A9 / y2 / y3 / qa;
  } else {
    const double y2 = e / 3.7 + DW_AB;
    const double y3 = DW_A9 * det_log(y2);
    const double fa = 1.0 / (y3 * y3);
    const double fb = (2.0 + DW_AC / (y2 * y3)) * fa;
    const double r2 = w / DW_A2;
    const double x1 = 7.0 * fa - fb;
    const double x2 = 0.128 - 17.0 * fa + 2.5 * fb;
    const double x3 = -0.128 + 13.0 * fa - (fb + fb);
    const double x4 = 0.032 - 3.0 * fa + 0.5 * fb;
    f = x1 + r2 * (x2 + r2 * (x3 + r2 * x4));
    dfdq = (x2 + r2 * (2.0 * x3 + r2 * 3.0 * x4)) / s / DW_A2;
  }
}

__device__ __forceinline__ void dw_pipe(const DevPlan& pl, double q, double r, double km,
                                        double diam, double kc, double& hloss, double& hgrad) {
  const double qa = fabs(q);
  const double s = pl.viscos * diam;
  const double e = kc / diam;
  if (qa <= DW_A2 * s) {
    const double r16 = 16.0 * PI_EPANET * s;
    hloss = q * (r16 * r + km * qa);
    hgrad = r16 * r + 2.0 * km * qa;
    return;
  }
  double f, dfdq;
  dw_friction(qa, s, e, f, dfdq);
  const double r1 = f * r + km;
  hloss = r1 * qa * q;
  hgrad = (2.0 * r1 * qa) + (dfdq * r * qa * qa);
}

// d hloss / d r of an open D-W pipe (dgga/autodiff.py:249-296): laminar
// 16 pi s Q; otherwise f |Q| Q (f depends on Q, not on r).
__device__ __forceinline__ double dw_dr(const DevPlan& pl, double q, double diam, double kc) {
  const double qa = fabs(q);
  const double s = pl.viscos * diam;
  if (qa <= DW_A2 * s) return 16.0 * PI_EPANET * s * q;
  double f, dfdq;
  dw_friction(qa, s, kc / diam, f, dfdq);
  return f * qa * q;
}

__device__ __forceinline__ void valve_open(const DevPlan& pl, double q, double km, double& P,
                                           double& Y) {
  if (km > 0.0) {
    const double qa = fabs(q);
    double hg = 2.0 * km * qa;
    const bool lin = hg < pl.rqtol;
    if (lin) hg = pl.rqtol;
    const double hl = lin ? q * hg : q * hg / 2.0;
    P = 1.0 / hg;
    Y = hl / hg;
  } else {
    P = 1.0 / CSMALL;
    Y = q;
  }
}

// CUSTOM pump (dgga/solver.py:2288-2318, 1452-1474; hydcoeffs.c:716-733,
// 794-831): the local line of the curve segment holding |q|/w in user
// units; the segment's right end is the first point with x >= |q|/w, the
// first and last segments extend beyond the curve.  N = 1: hgrad = R w,
// hloss = H0 w^2 + hgrad q with H0 = -h0, R = -r of the segment.
__device__ __forceinline__ void pump_custom(const DevPlan& pl, int k, double q, double& P,
                                            double& Y) {
  const double spc = fmax(pl.pu_speed[k], 1e-12);
  const double qq = (fabs(q) / spc) * pl.ucf_flow;
  const int a = pl.pu_cptr[k], npts = pl.pu_cptr[k + 1] - a;
  const double* x = pl.pu_cx + a;
  const double* y = pl.pu_cy + a;
  int k2 = 0;
  while (k2 < npts && x[k2] < qq) ++k2;
  if (k2 == 0) k2 = 1;
  else if (k2 == npts) k2 = npts - 1;
  const int k1 = k2 - 1;
  double r = (y[k2] - y[k1]) / (x[k2] - x[k1]);
  double h0 = y[k1] - r * x[k1];
  h0 = h0 / pl.ucf_head;
  r = r * pl.ucf_flow / pl.ucf_head;
  const double hg = -r * spc;
  const double hl = -h0 * (spc * spc) + hg * q;
  P = 1.0 / hg;
  Y = hl / hg;
}

__device__ __forceinline__ void pump_PY(const DevPlan& pl, int k, int f, double q, double& P,
                                        double& Y) {
  if (f & LF_PUMP_NOCURVE) {
    P = 1.0 / CSMALL;
    Y = q;
    return;
  }
  if (f & LF_PUMP_CUSTOM) {
    pump_custom(pl, k, q, P, Y);
    return;
  }
  const double sgn = q < 0.0 ? -1.0 : 1.0;
  const double qa = q * sgn;
  const double spc = fmax(pl.pu_speed[k], 1e-12);
  const double n = pl.pu_n[k];
  const double h0e = spc * spc * pl.pu_h0[k];
  const double re = pl.pu_r[k] * det_pow(spc, 2.0 - n);
  double hg, hl;
  if (f & LF_PUMP_CHP) {
    const double qa_c = fmax(qa, 1e-30);
    const double raw = -re / qa_c / qa_c;
    const bool big = raw > CBIG, sml = raw < pl.rqtol;
    hg = big ? CBIG : (sml ? pl.rqtol : raw);
    hl = (big || sml) ? -hg * q : re / (sgn * qa_c);
  } else if (f & LF_PUMP_LIN) {
    hg = re;
    hl = h0e + re * q;
  } else {
    const double nm1 = n - 1.0;
    double pw;
    if (qa == 0.0)
      pw = nm1 > 0.0 ? 0.0 : (nm1 < 0.0 ? __builtin_huge_val() : 1.0);
    else
      pw = det_pow(qa, nm1);
    const double raw = n * re * pw;
    const bool sml = raw < pl.rqtol;
    hg = sml ? pl.rqtol : raw;
    hl = sml ? h0e + hg * q : h0e + hg * q / fmax(n, 1e-30);
  }
  P = 1.0 / hg;
  Y = hl / hg;
}

// d hloss / d (speed, h0, r) of open pump link k at flow q (the pumpcoeff
// derivatives of dgga/autodiff.py:122-218), on the branches pump_PY takes:
// NOCURVE and the CBIG / RQtol clamps of CONST_HP carry none, the RQtol
// clamp of a power curve none in r, a CUSTOM pump only the speed (its
// segment line is fixed by the curve points).
__device__ __forceinline__ void pump_dtheta(const DevPlan& pl, int k, int f, double q,
                                            double& dw, double& dh0, double& drp) {
  dw = 0.0;
  dh0 = 0.0;
  drp = 0.0;
  if (f & LF_PUMP_NOCURVE) return;
  const double spc = fmax(pl.pu_speed[k], 1e-12);
  if (f & LF_PUMP_CUSTOM) {
    // the segment of pump_custom: hloss = -h0 w^2 - r w q
    const double qq = (fabs(q) / spc) * pl.ucf_flow;
    const int a = pl.pu_cptr[k], npts = pl.pu_cptr[k + 1] - a;
    const double* x = pl.pu_cx + a;
    const double* y = pl.pu_cy + a;
    int k2 = 0;
    while (k2 < npts && x[k2] < qq) ++k2;
    if (k2 == 0) k2 = 1;
    else if (k2 == npts) k2 = npts - 1;
    const int k1 = k2 - 1;
    double r = (y[k2] - y[k1]) / (x[k2] - x[k1]);
    double h0 = y[k1] - r * x[k1];
    h0 = h0 / pl.ucf_head;
    r = r * pl.ucf_flow / pl.ucf_head;
    dw = -2.0 * spc * h0 - r * q;
    return;
  }
  const double sgn = q < 0.0 ? -1.0 : 1.0;
  const double qa = q * sgn;
  const double n = pl.pu_n[k];
  const double H0 = pl.pu_h0[k], R = pl.pu_r[k];
  const double w2n = det_pow(spc, 2.0 - n);
  const double re = R * w2n;
  if (f & LF_PUMP_CHP) {
    // hloss = R w^(2-n) / q
    const double qa_c = fmax(qa, 1e-30);
    const double raw = -re / qa_c / qa_c;
    if (raw > CBIG || raw < pl.rqtol) return;
    const double qs = sgn * qa_c;
    dw = (2.0 - n) * R * det_pow(spc, 1.0 - n) / qs;
    drp = w2n / qs;
  } else if (f & LF_PUMP_LIN) {
    // hloss = w^2 H0 + R w q
    dw = 2.0 * spc * H0 + R * q;
    dh0 = spc * spc;
    drp = spc * q;
  } else {
    // hloss = w^2 H0 + R w^(2-n) |q|^(n-1) q, or w^2 H0 + RQtol q clamped
    const double nm1 = n - 1.0;
    double pw;
    if (qa == 0.0)
      pw = nm1 > 0.0 ? 0.0 : (nm1 < 0.0 ? __builtin_huge_val() : 1.0);
    else
      pw = det_pow(qa, nm1);
    dh0 = spc * spc;
    if (n * re * pw < pl.rqtol) {
      dw = 2.0 * spc * H0;
      return;
    }
    dw = 2.0 * spc * H0 + (2.0 - n) * R * det_pow(spc, 1.0 - n) * pw * q;
    drp = w2n * pw * q;
  }
}

// P, Y of link k at flow q and status st (pcv PRVs -> 0, handled later)
// f = l_flags[k], km = l_km[k] preloaded by the caller
__device__ __forceinline__ LinkPY link_PY_f(const DevPlan& pl, int k, int f, double km,
                                            double q, double r, int8_t st) {
  double P, Y;
  if (f & LF_PIPE) {
    double hl, hg;
    if (pl.headloss == 0)
      hw_pipe(pl, q, r, km, hl, hg);
    else
      dw_pipe(pl, q, r, km, pl.l_diam[k], pl.l_kc[k], hl, hg);
    P = 1.0 / hg;
    Y = hl / hg;
  } else if (f & LF_TCV) {
    valve_open(pl, q, pl.l_km_tcv[k], P, Y);
  } else if (f & LF_PUMP) {
    pump_PY(pl, k, f, q, P, Y);
  } else if (f & LF_PRV_FIXED) {
    valve_open(pl, q, pl.l_km_valve[k], P, Y);
  } else {
    P = 0.0;
    Y = 0.0;
  }
  if (f & LF_PCV) {
    P = 0.0;
    Y = 0.0;
  } else if (st <= ST_CLOSED || (f & LF_ZERO_SPEED)) {
    P = 1.0 / CBIG;
    Y = q;
  }
  return LinkPY{P, Y};
}

__device__ __forceinline__ LinkPY link_PY(const DevPlan& pl, int k, double q, double r,
                                          int8_t st) {
  return link_PY_f(pl, k, pl.l_flags[k], pl.l_km[k], q, r, st);
}

// Batched gathers: a loop over items issues all loads of kGather items before
// it consumes any, so one memory latency is paid per batch instead of per
// item; results are accumulated in the original item order.
constexpr int kGather = 4;

// Signed gather over one side of a junction row: for entries e in [e0, e1)
// of jadj, acc_a += sgn * a[k] and acc_b += sgn * (b[k] * w[k]) in entry
// order (b and w may be null: acc_b untouched); sgn = +1 adds the values,
// -1 adds their negation, as the reference's scatter of -x does.
__device__ inline void row_gather(const DevPlan& pl, int e0, int e1, double sgn, const double* a,
                                  const double* b, const double* w, double& acc_a,
                                  double& acc_b) {
  for (int eb = e0; eb < e1; eb += kGather) {
    double av[kGather], bv[kGather], wv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = pl.jadj[min(eb + u, e1 - 1)] & 0x7fffffff;
      av[u] = a[k];
      if (b) {
        bv[u] = b[k];
        wv[u] = w[k];
      }
    }
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      if (eb + u >= e1) break;
      acc_a += sgn > 0.0 ? av[u] : -av[u];
      if (b) acc_b += sgn > 0.0 ? bv[u] * wv[u] : -(bv[u] * wv[u]);
    }
  }
}

// dst[j] = a[j], or a[j] - b[j] when b is given, for j < n; loads batched
__device__ inline void vec_copy(double* dst, const double* a, const double* b, int n,
                                int w = 0, int nw = 1) {
  const int lane = lane_id();
  // kRowG rows per lane per round: one memory round trip per round; the
  // rounds alternate between the nw wavefronts
  for (int j0 = lane + 64 * kRowG * w; j0 < n; j0 += 64 * kRowG * nw) {
    double av[kRowG], bv[kRowG];
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      const int j = min(j0 + 64 * u, n - 1);
      av[u] = a[j];
      bv[u] = b ? b[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kRowG; ++u) {
      const int j = j0 + 64 * u;
      if (j < n) dst[j] = b ? av[u] - bv[u] : av[u];
    }
  }
}

// link coefficients P, Y of every link (K1 and the polish), batched
__device__ inline void link_coeffs_all(const DevPlan& pl, const double* q, const double* r,
                                       const int8_t* S, double* P, double* Y, int w = 0,
                                       int nw = 1) {
  const int lane = lane_id();
  for (int k0 = lane + 64 * kGather * w; k0 < pl.L; k0 += 64 * kGather * nw) {
    double qv[kGather], rv[kGather], kmv[kGather];
    int fv[kGather];
    int8_t sv[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = min(k0 + 64 * u, pl.L - 1);
      qv[u] = q[k];
      rv[u] = r[k];
      sv[u] = S[k];
      fv[u] = pl.l_flags[k];
      kmv[u] = pl.l_km[k];
    }
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int k = k0 + 64 * u;
      if (k < pl.L) {
        const LinkPY py = link_PY_f(pl, k, fv[u], kmv[u], qv[u], rv[u], sv[u]);
        P[k] = py.P;
        Y[k] = py.Y;
      }
    }
  }
}

// emitterheadloss (dgga/solver.py:2368-2377)
__device__ __forceinline__ void emitter_hl(const DevPlan& pl, double e, double ke, double& hloss,
                                           double& hgrad) {
  const double ke_adj = fmax(ke, CSMALL);
  double hg = pl.qexp * ke_adj * det_pow(fmax(fabs(e), 1e-30), pl.qexp - 1.0);
  const bool lin = hg < pl.rqtol;
  if (lin) hg = pl.rqtol;
  hloss = lin ? hg * e : hg * e / pl.qexp;
  hgrad = hg;
}

}  // namespace wds
