// Arena-mode build of the scenario kernels: the kernel bodies of
// kernels_impl.h / kernels_polish.h compiled with WDS_GLOBAL_ARENA, so that a
// scenario's factor, head and solve vectors, sinks and Woodbury block live in
// a global-memory arena (at the end of its workspace slice) instead of LDS,
// at the same byte offsets the plan's programs address.  Used for networks
// whose layout exceeds a CU's 160 KB of LDS (several thousand junctions).
#define WDS_GLOBAL_ARENA 1
#include <hip/hip_runtime.h>

#include "../../include/wds.h"
#include "arena.h"
#include "device_common.h"
#include "kernels_impl.h"

namespace wds {

void arena_solve(const DevPlan& pl, const SolveParams& sp, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_arena, dim3(B), dim3(64 * kSolveWaves), 0, st, pl, sp);
}

void arena_newton_iter(const DevPlan& pl, const SolveParams& sp, int it, int32_t* ctrl, int B,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_newton_iter_arena, dim3(B), dim3(64), 0, st, pl, sp, it, ctrl);
}

void arena_factor_solve(const DevPlan& pl, const double* csr, const double* F, double* H,
                        int n_refine, double* factor_out, int32_t* info, double* ws, int B,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_factor_solve_arena, dim3(B), dim3(64), 0, st, pl, csr, F, H, n_refine,
                     factor_out, info, ws);
}

void arena_factor(const DevPlan& pl, const double* csr, double* factor_out, int32_t* info,
                  double* ws, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_factor_arena, dim3(B), dim3(64), 0, st, pl, csr, factor_out, info, ws);
}

void arena_trisolve(const DevPlan& pl, const double* factor, const double* csr, const double* F,
                    double* H, int n_refine, double* ws, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_trisolve_arena, dim3(B), dim3(64), 0, st, pl, factor, csr, F, H, n_refine,
                     ws);
}

void arena_adjoint(const DevPlan& pl, const AdjParams& ap, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_adjoint_arena, dim3(B), dim3(64 * kSolveWaves), 0, st, pl, ap);
}

void arena_solve_multi(const DevPlan* plans, const SolveParams* params, const int2* items,
                       int nitems, int32_t* counter, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_multi_arena, dim3(grid), dim3(64 * kSolveWaves), 0, st, plans,
                     params, items, nitems, counter, 0);
}

void arena_adjoint_multi(const DevPlan* plans, const AdjParams* params, const int2* items,
                         int nitems, int32_t* counter, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_adjoint_multi_arena, dim3(grid), dim3(64 * kSolveWaves), 0, st, plans,
                     params, items, nitems, counter, 0);
}

}  // namespace wds
