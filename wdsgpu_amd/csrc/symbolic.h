// Host-side symbolic analysis for the batched sparse LDL^T factorisation of
// the GGA node matrix A (junction x junction, symmetric positive definite).
//
// Once per topology: minimum-degree ordering, elimination tree, the pattern
// of L, tree levels, and the per-lane instruction streams that a 64-lane
// wavefront executes to factor and solve with L resident in LDS.
#pragma once
#include <cstdint>
#include <vector>

namespace wds {

constexpr int kLanes = 64;

// One instruction of a level-scheduled sparse dot-product program.  Each lane
// of the wavefront owns its own stream; streams are stored step-major with a
// fixed stride of 64, so lane l reads instr[s * 64 + l] at step s: the loads
// of a step are one contiguous 1 KiB line set, their addresses depend on
// nothing loaded, and a deep prefetch can run ahead.  Idle lanes hold NOPs.
//   factor: acc -= V[a] * V[b] * V[c]     (w_ik * w_jk * invD_k)
//   fwd:    acc -= V[a] * X[b]            (w_jk * z_k)
//   bwd:    acc += V[a] * X[b]            (w_ij * x_i)
// flags live in the top bits of t.
struct Instr {
  int32_t a, b, c, t;
};
constexpr int32_t kFirst = 1 << 30;     // acc starts from the target (or 0)
constexpr int32_t kLast = 1 << 29;      // store acc into the target after this op
constexpr int32_t kDiag = 1 << 28;      // factor: target is a diagonal -> store 1/acc
constexpr int32_t kNop = 1 << 27;       // padding, no load/store
constexpr int32_t kLevelEnd = 1 << 26;  // last step of a level (set in every lane)
constexpr int kWidthShift = 23;         // bits 23-25 at a level end: log2 of the lane's
                                        // group width (0: not in a group)
constexpr int32_t kZeroInit = 1 << 22;  // a group slice's partial sum starts at 0
constexpr int32_t kTMask = (1 << 22) - 1;

// one dot-product task of a level: its updates (flags are added by the
// scheduler) and whether its target is a diagonal pivot
struct Task {
  std::vector<Instr> ins;
  bool diag;
};

// Padding instruction: multiplies by V[zero] == 0.0 (X[0] is a valid
// operand of the solves) and stores nowhere, so the executor needs no branch.
inline Instr nop_instr(int32_t zero) { return Instr{zero, 0, zero, kNop}; }

struct Program {
  std::vector<Instr> instr;            // step-major, nsteps * 64
  std::vector<int32_t> level_end;      // cumulative steps at each level end
  int64_t useful = 0;                  // non-padding instructions
  int64_t cost = 0;                    // scheduler cost model summed over the levels
  int32_t nsteps() const { return level_end.empty() ? 0 : level_end.back(); }
};

struct Symbolic {
  int n = 0;                           // junction count
  std::vector<int32_t> perm, iperm;    // perm[new] = old, iperm[old] = new
  std::vector<int32_t> parent;         // elimination tree (new order), -1 root
  std::vector<int32_t> colptr, rowidx; // L pattern, CSC, diagonal first
  std::vector<int32_t> height;         // 0 at leaves
  int nlevels = 0;
  int64_t nnzL = 0;
  int64_t factor_updates = 0;          // sum over targets of update terms
  Program factor, fwd, bwd;
  // position of L(i,j) for a pattern entry (binary search in column j)
  int64_t pos(int i, int j) const;
};

// adjacency: for each vertex (original index) its neighbours (no self loops,
// duplicates allowed).  ordering: 0 = minimum degree, 1 = natural, 3 + d =
// independent-set rounds with degree cap d (0 <= d <= 8).
constexpr int kOrderRounds = 3;
void analyse(int n, const std::vector<std::vector<int32_t>>& adj, int ordering,
             Symbolic& s);

// weighted scheduler cost of one Newton iteration's programs: one
// factorisation and three forward/backward solve pairs (a solve and two
// refinement steps)
inline int64_t iteration_cost(const Symbolic& s) {
  return s.factor.cost + 3 * (s.fwd.cost + s.bwd.cost);
}

}  // namespace wds
