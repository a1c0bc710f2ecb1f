// Host symbolic analysis: minimum-degree ordering, elimination tree, pattern
// of L, tree levels and the level-scheduled per-lane programs.
//
// The ordering is an exact minimum-degree elimination on the explicit
// elimination graph (eliminating v turns its neighbours into a clique), ties
// broken by the lowest original index.  The reference models the same
// ordering family with EPANET's genmmd port (dgga/smatrix.py:352-430) and
// computes the elimination tree and column counts in dgga/symbolic.py:101-230;
// this file is an independent implementation of the textbook algorithms.
#include "symbolic.h"

#include <algorithm>
#include <queue>
#include <set>
#include <stdexcept>
#include <utility>

namespace wds {

int64_t Symbolic::pos(int i, int j) const {
  if (i == j) return colptr[j];
  const int32_t* b = rowidx.data() + colptr[j] + 1;
  const int32_t* e = rowidx.data() + colptr[j + 1];
  const int32_t* p = std::lower_bound(b, e, i);
  if (p == e || *p != i) return -1;
  return p - rowidx.data();
}

static void minimum_degree(int n, const std::vector<std::vector<int32_t>>& adj,
                           std::vector<int32_t>& order) {
  std::vector<std::set<int32_t>> g(n);
  for (int v = 0; v < n; ++v)
    for (int32_t u : adj[v])
      if (u != v) {
        g[v].insert(u);
        g[u].insert(v);
      }
  std::set<std::pair<int32_t, int32_t>> pq;
  std::vector<int32_t> deg(n);
  for (int v = 0; v < n; ++v) {
    deg[v] = (int32_t)g[v].size();
    pq.insert({deg[v], v});
  }
  order.clear();
  order.reserve(n);
  std::vector<int32_t> nb;
  while (!pq.empty()) {
    int32_t v = pq.begin()->second;
    pq.erase(pq.begin());
    order.push_back(v);
    nb.assign(g[v].begin(), g[v].end());
    for (int32_t u : nb) g[u].erase(v);
    for (size_t x = 0; x < nb.size(); ++x)
      for (size_t y = 0; y < nb.size(); ++y)
        if (x != y) g[nb[x]].insert(nb[y]);
    for (int32_t u : nb) {
      pq.erase({deg[u], u});
      deg[u] = (int32_t)g[u].size();
      pq.insert({deg[u], u});
    }
    g[v].clear();
  }
}

// Longest-processing-time assignment of the tasks of one level to 64
// lanes, lanes renumbered by decreasing load, then step-major compacted
// layout (only active lanes store an instruction).
static void emit_level(const std::vector<std::vector<Instr>>& tasks, Program& prog) {
  if (tasks.empty()) return;
  std::vector<int> idx(tasks.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
    return tasks[x].size() > tasks[y].size();
  });
  using Item = std::pair<int64_t, int>;  // (load, lane)
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (int l = 0; l < kLanes; ++l) heap.push({0, l});
  std::vector<std::vector<Instr>> lane(kLanes);
  for (int ti : idx) {
    Item it = heap.top();
    heap.pop();
    for (const Instr& in : tasks[ti]) lane[it.second].push_back(in);
    prog.useful += (int64_t)tasks[ti].size();
    heap.push({it.first + (int64_t)tasks[ti].size(), it.second});
  }
  std::stable_sort(lane.begin(), lane.end(),
                   [](const std::vector<Instr>& x, const std::vector<Instr>& y) {
                     return x.size() > y.size();
                   });
  size_t steps = lane[0].size();
  steps = (steps + kStepAlign - 1) / kStepAlign * kStepAlign;
  for (size_t s = 0; s < steps; ++s) {
    int cnt = 0;
    while (cnt < kLanes && lane[cnt].size() > s) ++cnt;
    prog.step_off.push_back((int32_t)prog.instr.size());
    prog.step_cnt.push_back(cnt);
    for (int l = 0; l < cnt; ++l) prog.instr.push_back(lane[l][s]);
  }
  const int32_t prev = prog.level_end.empty() ? 0 : prog.level_end.back();
  prog.level_end.push_back(prev + (int32_t)steps);
}

void analyse(int n, const std::vector<std::vector<int32_t>>& adj, int ordering,
             Symbolic& s) {
  s = Symbolic();
  s.n = n;
  if (n <= 0) throw std::runtime_error("empty matrix");
  if (ordering == 0) {
    minimum_degree(n, adj, s.perm);
  } else {
    s.perm.resize(n);
    for (int i = 0; i < n; ++i) s.perm[i] = i;
  }
  s.iperm.assign(n, -1);
  for (int i = 0; i < n; ++i) s.iperm[s.perm[i]] = i;
  // symmetric adjacency in the new order
  std::vector<std::vector<int32_t>> nadj(n);
  for (int v = 0; v < n; ++v)
    for (int32_t u : adj[v])
      if (u != v) {
        nadj[s.iperm[v]].push_back(s.iperm[u]);
        nadj[s.iperm[u]].push_back(s.iperm[v]);
      }
  // symbolic elimination: struct(j) = adj_>j(j) U (struct(children) \ {j});
  // parent(j) = min struct(j)
  std::vector<std::vector<int32_t>> cols(n), children(n);
  s.parent.assign(n, -1);
  for (int j = 0; j < n; ++j) {
    std::vector<int32_t> rows;
    for (int32_t u : nadj[j])
      if (u > j) rows.push_back(u);
    for (int32_t c : children[j])
      for (int32_t r : cols[c])
        if (r != j) rows.push_back(r);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    cols[j] = rows;
    if (!rows.empty()) {
      s.parent[j] = rows[0];
      children[rows[0]].push_back(j);
    }
  }
  s.colptr.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) s.colptr[j + 1] = s.colptr[j] + 1 + (int32_t)cols[j].size();
  s.nnzL = s.colptr[n];
  s.rowidx.resize(s.nnzL);
  for (int j = 0; j < n; ++j) {
    s.rowidx[s.colptr[j]] = j;
    std::copy(cols[j].begin(), cols[j].end(), s.rowidx.begin() + s.colptr[j] + 1);
  }
  // heights
  s.height.assign(n, 0);
  s.nlevels = 0;
  for (int j = 0; j < n; ++j) {
    int h = 0;
    for (int32_t c : children[j]) h = std::max(h, s.height[c] + 1);
    s.height[j] = h;
    s.nlevels = std::max(s.nlevels, h + 1);
  }
  // row structures (k < j with L(j,k) != 0), k ascending
  std::vector<std::vector<int32_t>> rowst(n);
  for (int k = 0; k < n; ++k)
    for (int32_t p = s.colptr[k] + 1; p < s.colptr[k + 1]; ++p)
      rowst[s.rowidx[p]].push_back(k);
  std::vector<std::vector<int32_t>> bylevel(s.nlevels);
  for (int j = 0; j < n; ++j) bylevel[s.height[j]].push_back(j);
  const int32_t ZERO = (int32_t)s.nnzL;  // V[nnzL] holds 0.0
  // ---- factor program ----
  s.factor_updates = 0;
  for (int lev = 0; lev < s.nlevels; ++lev) {
    std::vector<std::vector<Instr>> tasks;
    for (int32_t j : bylevel[lev]) {
      const int32_t cj = s.colptr[j];
      const int32_t nrow = s.colptr[j + 1] - cj;  // targets: diag + sub-diagonal
      std::vector<std::vector<Instr>> tgt(nrow);
      for (int32_t k : rowst[j]) {
        const int64_t b = s.pos(j, k);
        const int32_t ck = s.colptr[k];
        for (int64_t p = b; p < s.colptr[k + 1]; ++p) {
          const int32_t i = s.rowidx[p];
          const int64_t t = s.pos(i, j);
          if (t < 0) throw std::runtime_error("symbolic: missing fill entry");
          tgt[t - cj].push_back(Instr{(int32_t)p, (int32_t)b, ck, (int32_t)t});
        }
      }
      for (int32_t r = 0; r < nrow; ++r) {
        auto& v = tgt[r];
        const bool diag = (r == 0);
        if (v.empty()) {
          if (!diag) continue;  // off-diagonal without updates: w_ij = A_ij
          v.push_back(Instr{ZERO, ZERO, ZERO, cj});
        }
        s.factor_updates += (int64_t)v.size();
        v.front().t |= kFirst;
        v.back().t |= kLast | (diag ? kDiag : 0);
        tasks.push_back(std::move(v));
      }
    }
    emit_level(tasks, s.factor);
  }
  // ---- forward solve: z_j = invD_j * (b_j - sum_k w_jk z_k) ----
  for (int lev = 0; lev < s.nlevels; ++lev) {
    std::vector<std::vector<Instr>> tasks;
    for (int32_t j : bylevel[lev]) {
      std::vector<Instr> v;
      for (int32_t k : rowst[j])
        v.push_back(Instr{(int32_t)s.pos(j, k), k, s.colptr[j], j});
      if (v.empty()) v.push_back(Instr{ZERO, j, s.colptr[j], j});
      v.front().t |= kFirst;
      v.back().t |= kLast;
      tasks.push_back(std::move(v));
    }
    emit_level(tasks, s.fwd);
  }
  // ---- backward solve: x_j = z_j - invD_j * sum_i w_ij x_i (top level first) ----
  for (int lev = s.nlevels - 1; lev >= 0; --lev) {
    std::vector<std::vector<Instr>> tasks;
    for (int32_t j : bylevel[lev]) {
      std::vector<Instr> v;
      for (int32_t p = s.colptr[j] + 1; p < s.colptr[j + 1]; ++p)
        v.push_back(Instr{p, s.rowidx[p], s.colptr[j], j});
      if (v.empty()) continue;
      v.front().t |= kFirst;
      v.back().t |= kLast;
      tasks.push_back(std::move(v));
    }
    emit_level(tasks, s.bwd);
  }
}

}  // namespace wds
