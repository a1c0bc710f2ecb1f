// Launchers of the arena-mode build of the scenario kernels (kernels_arena.hip):
// the same kernels with each scenario's LDS layout in a global-memory arena,
// for networks too large for a CU's LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "devplan.h"

namespace wds {

struct SolveParams;
struct AdjParams;

void arena_solve(const DevPlan& pl, const SolveParams& sp, int B, hipStream_t st);
void arena_newton_iter(const DevPlan& pl, const SolveParams& sp, int it, int32_t* ctrl, int B,
                       hipStream_t st);
void arena_factor_solve(const DevPlan& pl, const double* csr, const double* F, double* H,
                        int n_refine, double* factor_out, int32_t* info, double* ws, int B,
                        hipStream_t st);
void arena_factor(const DevPlan& pl, const double* csr, double* factor_out, int32_t* info,
                  double* ws, int B, hipStream_t st);
void arena_trisolve(const DevPlan& pl, const double* factor, const double* csr, const double* F,
                    double* H, int n_refine, double* ws, int B, hipStream_t st);
void arena_adjoint(const DevPlan& pl, const AdjParams& ap, int B, hipStream_t st);
void arena_solve_multi(const DevPlan* plans, const SolveParams* params, const int2* items,
                       int nitems, int32_t* counter, int grid, hipStream_t st);
void arena_adjoint_multi(const DevPlan* plans, const AdjParams* params, const int2* items,
                         int nitems, int32_t* counter, int grid, hipStream_t st);

}  // namespace wds
