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

// Depth-limited elimination: rounds of independent sets.  Every round
// eliminates a maximal independent set (in the current elimination graph,
// fill included) of the vertices whose degree is at most max(dmin, dcap),
// taken in ascending (degree, index) order.  No two vertices of one round
// are adjacent, so none is an elimination-tree descendant of another and the
// tree's height is at most the number of rounds; the price is fill from
// eliminating vertices of degree up to dcap before the minimum-degree
// vertices alone would.  On tree-like networks (long chains of degree-2
// vertices) a round halves every chain, as cyclic reduction does, where
// minimum degree eats chains from their ends one level at a time.
static void independent_rounds(int n, const std::vector<std::vector<int32_t>>& adj, int dcap,
                               std::vector<int32_t>& order) {
  std::vector<std::set<int32_t>> g(n);
  for (int v = 0; v < n; ++v)
    for (int32_t u : adj[v])
      if (u != v) {
        g[v].insert(u);
        g[u].insert(v);
      }
  std::vector<char> alive(n, 1), blocked(n, 0);
  order.clear();
  order.reserve(n);
  std::vector<std::pair<int32_t, int32_t>> cand;
  std::vector<int32_t> sel, nb;
  int left = n;
  while (left > 0) {
    int dmin = n;
    for (int v = 0; v < n; ++v)
      if (alive[v]) dmin = std::min(dmin, (int)g[v].size());
    const int thr = std::max(dmin, dcap);
    cand.clear();
    for (int v = 0; v < n; ++v)
      if (alive[v] && (int)g[v].size() <= thr) cand.push_back({(int32_t)g[v].size(), v});
    std::sort(cand.begin(), cand.end());
    sel.clear();
    for (const auto& c : cand) {
      const int v = c.second;
      if (blocked[v]) continue;
      sel.push_back(v);
      blocked[v] = 1;
      for (int32_t u : g[v]) blocked[u] = 1;
    }
    for (const auto& c : cand) blocked[c.second] = 0;
    for (int32_t v : sel)
      for (int32_t u : g[v]) blocked[u] = 0;
    for (int32_t v : sel) {
      order.push_back(v);
      alive[v] = 0;
      --left;
      nb.assign(g[v].begin(), g[v].end());
      for (int32_t u : nb) g[u].erase(v);
      for (size_t x = 0; x < nb.size(); ++x)
        for (size_t y = 0; y < nb.size(); ++y)
          if (x != y) g[nb[x]].insert(nb[y]);
      g[v].clear();
    }
  }
}

// Schedule the tasks of one level onto the 64 lanes and append its steps.
//
// A level of S steps gives every task the smallest power-of-two group of g
// lanes whose slices of its dot product fit in S steps (g = 1: the whole task
// on one lane).  Groups take aligned lane blocks, largest first from lane 0,
// so lane 0 always belongs to the level's widest group.  Single-lane tasks
// fill the lanes' remaining steps, longest first onto the lane with the most
// room; a grouped lane runs its single-lane tasks first and its slice last.
// A single-lane task's first instruction loads its target (kFirst), its last
// stores the result (kLast).  A slice starts from the target's value on the
// group's first lane and from zero on the others (kFirst | kZeroInit; a lane
// with no share contributes one explicit zero); at the level end the executor
// sums each group with a butterfly of cross-lane exchanges, lanes adding only
// within their own group, and the group's first lane finalises the target.
//
// Every S from the smallest that could fit to the single-lane schedule's
// length is tried; the cost model counts a level's first step (its operands
// wait for the previous level's stores) as kLevelCost, every further step
// as kStepCost, and each butterfly stage as kStageCost.
//
// Layout: step-major with a fixed stride of 64 (idle lanes hold NOPs); the
// level's last step carries kLevelEnd in every lane and each lane's log2
// group width (0 outside a group).
namespace {
constexpr int kLevelCost = 13, kStepCost = 5, kStageCost = 2;

struct LaneSched {
  std::vector<std::vector<Instr>> lane;
  int lw_max = 0;
  size_t steps = 0;
};

int ceil_log2(size_t x) {
  int l = 0;
  while (((size_t)1 << l) < x) ++l;
  return l;
}

// schedule of the level in S steps, or false if it does not fit
bool pack_level(const std::vector<Task>& tasks, size_t S, int32_t zero, LaneSched& out) {
  const size_t nt = tasks.size();
  std::vector<int> lw(nt, 0);
  for (size_t i = 0; i < nt; ++i) {
    const size_t m = tasks[i].ins.size();
    while ((m + ((size_t)1 << lw[i]) - 1) >> lw[i] > S) {
      if (++lw[i] > 6) return false;
    }
  }
  std::vector<int> grouped, single;
  for (size_t i = 0; i < nt; ++i) (lw[i] ? grouped : single).push_back((int)i);
  std::stable_sort(grouped.begin(), grouped.end(), [&](int x, int y) { return lw[x] > lw[y]; });
  std::stable_sort(single.begin(), single.end(), [&](int x, int y) {
    return tasks[x].ins.size() > tasks[y].ins.size();
  });
  out.lane.assign(kLanes, {});
  std::vector<std::vector<Instr>> slice(kLanes);
  std::vector<int> lane_lw(kLanes, 0);
  out.lw_max = 0;
  int next = 0;
  for (int ti : grouped) {
    const int W = 1 << lw[ti];
    if (next + W > kLanes) return false;
    const Task& tk = tasks[ti];
    const size_t m = tk.ins.size();
    const size_t sl = (m + W - 1) / W;
    const int32_t dflag = tk.diag ? kDiag : 0;
    for (int i = 0; i < W; ++i) {
      auto& ln = slice[next + i];
      lane_lw[next + i] = lw[ti];
      const size_t b = (size_t)i * sl, e = std::min(m, b + sl);
      if (b >= e) {
        // no share of the dot product: contribute an explicit zero
        Instr in = nop_instr(zero);
        in.t = (tk.ins.front().t & kTMask) | kFirst | kZeroInit | dflag;
        ln.push_back(in);
        continue;
      }
      for (size_t p = b; p < e; ++p) {
        Instr in = tk.ins[p];
        in.t |= dflag | (p == b ? (kFirst | (i == 0 ? 0 : kZeroInit)) : 0);
        ln.push_back(in);
      }
    }
    out.lw_max = std::max(out.lw_max, lw[ti]);
    next += W;
  }
  // single-lane tasks into the room each lane has left before its slice
  using Item = std::pair<int64_t, int>;  // (-room, lane): most room first
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (int l = 0; l < kLanes; ++l) heap.push({-(int64_t)(S - slice[l].size()), l});
  for (int ti : single) {
    Item it = heap.top();
    heap.pop();
    const int64_t room = -it.first;
    std::vector<Instr> v = tasks[ti].ins;
    if ((int64_t)v.size() > room) return false;
    v.front().t |= kFirst;
    v.back().t |= kLast | (tasks[ti].diag ? kDiag : 0);
    for (const Instr& in : v) out.lane[it.second].push_back(in);
    heap.push({-(room - (int64_t)v.size()), it.second});
  }
  out.steps = 0;
  for (int l = 0; l < kLanes; ++l) {
    for (const Instr& in : slice[l]) out.lane[l].push_back(in);
    out.steps = std::max(out.steps, out.lane[l].size());
  }
  // the level end carries each lane's group width
  for (int l = 0; l < kLanes; ++l) {
    out.lane[l].resize(out.steps, nop_instr(zero));
    out.lane[l].back().t |= kLevelEnd | (lane_lw[l] << kWidthShift);
  }
  return true;
}
}  // namespace

static void emit_level(const std::vector<Task>& tasks, Program& prog, int32_t zero) {
  if (tasks.empty()) return;
  size_t total = 0, mmax = 0;
  for (const Task& t : tasks) {
    total += t.ins.size();
    mmax = std::max(mmax, t.ins.size());
  }
  prog.useful += (int64_t)total;
  LaneSched best, cur;
  int64_t best_cost = -1;
  // the single-lane schedule needs at most sum(len) steps; larger S only
  // repeats it
  const size_t lo = std::max<size_t>(1, (total + kLanes - 1) / kLanes);
  for (size_t S = lo; S <= std::max(lo, total); ++S) {
    if (!pack_level(tasks, S, zero, cur)) continue;
    const int64_t cost = kLevelCost + kStepCost * (int64_t)(cur.steps - 1) +
                         kStageCost * cur.lw_max;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = cur;
    }
    if (cur.lw_max == 0) break;  // no groups: a longer level only costs more
  }
  if (best_cost < 0) throw std::runtime_error("symbolic: a level does not fit the wavefront");
  prog.cost += best_cost;
  for (size_t st = 0; st < best.steps; ++st)
    for (int l = 0; l < kLanes; ++l) prog.instr.push_back(best.lane[l][st]);
  const int32_t prev = prog.level_end.empty() ? 0 : prog.level_end.back();
  prog.level_end.push_back(prev + (int32_t)best.steps);
}

void analyse(int n, const std::vector<std::vector<int32_t>>& adj, int ordering,
             Symbolic& s) {
  s = Symbolic();
  s.n = n;
  if (n <= 0) throw std::runtime_error("empty matrix");
  if (ordering == 0) {
    minimum_degree(n, adj, s.perm);
  } else if (ordering >= 3 && ordering <= 3 + 8) {
    independent_rounds(n, adj, ordering - 3, s.perm);
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
  if (s.nnzL + 2 > kTMask) throw std::runtime_error("factor too large for the program encoding");
  for (int lev = 0; lev < s.nlevels; ++lev) {
    std::vector<Task> tasks;
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
        tasks.push_back(Task{std::move(v), diag});
      }
    }
    emit_level(tasks, s.factor, ZERO);
  }
  // The solves use the unit factor: after the numeric factorisation the
  // off-diagonal entries are scaled to l_ij = w_ij / D_j, so the forward
  // solve is y_j = b_j - sum_k l_jk y_k, then z = D^-1 y (a diagonal pass),
  // and the backward solve is x_j = z_j - sum_i l_ij x_i.  Both are the same
  // dot-product form; rows without updates need no task.
  // ---- forward solve ----
  for (int lev = 0; lev < s.nlevels; ++lev) {
    std::vector<Task> tasks;
    for (int32_t j : bylevel[lev]) {
      std::vector<Instr> v;
      for (int32_t k : rowst[j]) v.push_back(Instr{(int32_t)s.pos(j, k), k, 0, j});
      if (v.empty()) continue;
      tasks.push_back(Task{std::move(v), false});
    }
    if (tasks.empty()) continue;
    emit_level(tasks, s.fwd, ZERO);
  }
  // ---- backward solve (top level first) ----
  for (int lev = s.nlevels - 1; lev >= 0; --lev) {
    std::vector<Task> tasks;
    for (int32_t j : bylevel[lev]) {
      std::vector<Instr> v;
      for (int32_t p = s.colptr[j] + 1; p < s.colptr[j + 1]; ++p)
        v.push_back(Instr{p, s.rowidx[p], 0, j});
      if (v.empty()) continue;
      tasks.push_back(Task{std::move(v), false});
    }
    emit_level(tasks, s.bwd, ZERO);
  }
}

}  // namespace wds
