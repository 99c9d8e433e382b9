// Host-side program compiler: elimination-order planning, the step compiler that fixes
// every intermediate's axis order and arena offset, and the batched SAT front-end that
// turns CNF instances into conditioned contraction jobs (cubes).
//
// Reference anchors: the step records are the compiled form of the contract2 arguments
// precomputed by SimpleContraction._new (tensorsat/contractions/simple.py:134-194) and
// of cotengra's pairwise loop (tensorsat/contractions/cotengra.py:120-126); clause
// tensors and leaves follow tensorsat/lib/sat.py:208-262 and lib/bincirc.py:28-72.

#include <algorithm>
#include <climits>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "planner_internal.h"

namespace tsp {

namespace {

inline void merged(const std::vector<int>& a, const std::vector<int>& b,
                   const std::vector<int>& count, std::vector<int>& out, int& n_union) {
  out.clear();
  n_union = 0;
  size_t i = 0, j = 0;
  while (i < a.size() || j < b.size()) {
    int w, c;
    if (j >= b.size() || (i < a.size() && a[i] < b[j])) {
      w = a[i++];
      c = 1;
    } else if (i >= a.size() || b[j] < a[i]) {
      w = b[j++];
      c = 1;
    } else {
      w = a[i];
      ++i;
      ++j;
      c = 2;
    }
    ++n_union;
    if (count[w] > c) out.push_back(w);
  }
}

// Resize a list of wire lists to n empty lists, keeping the inner lists' capacity (the
// planner's per-thread scratch is reused for every cube, so steady state allocates nothing).
inline void reset_lists(std::vector<std::vector<int>>& v, size_t n) {
  if (v.size() != n) v.resize(n);
  for (auto& x : v) x.clear();
}

}  // namespace

bool path_legs(const BitNet& net, const std::vector<int32_t>& path,
               std::vector<std::vector<int>>& legs) {
  const int S = net.S;
  reset_lists(legs, (size_t)std::max(1, 2 * S - 1));
  thread_local std::vector<int> count;
  thread_local std::vector<char> alive;
  count.assign(net.W, 0);
  alive.assign(legs.size(), 0);
  for (int s = 0; s < S; ++s) {
    legs[s] = net.slot_wires[s];
    std::sort(legs[s].begin(), legs[s].end());
    for (int w : legs[s]) count[w]++;
    alive[s] = 1;
  }
  for (int w = 0; w < net.W; ++w)
    if (net.is_out[w]) count[w]++;
  if ((int)path.size() != 2 * std::max(0, S - 1)) return false;
  for (int i = 0; i + 1 < S; ++i) {
    int a = path[2 * i], b = path[2 * i + 1], k = S + i;
    if (a < 0 || b < 0 || a >= k || b >= k || a == b || !alive[a] || !alive[b]) return false;
    int nu;
    merged(legs[a], legs[b], count, legs[k], nu);
    for (int w : legs[a]) count[w]--;
    for (int w : legs[b]) count[w]--;
    for (int w : legs[k]) count[w]++;
    alive[a] = alive[b] = 0;
    alive[k] = 1;
  }
  return true;
}

TreeCost tree_cost(const BitNet& net, const std::vector<int32_t>& path) {
  TreeCost tc;
  const int S = net.S;
  thread_local std::vector<std::vector<int>> legs;
  thread_local std::vector<int> count;
  reset_lists(legs, (size_t)std::max(1, 2 * S - 1));
  count.assign(net.W, 0);
  double live = 0;
  for (int s = 0; s < S; ++s) {
    legs[s] = net.slot_wires[s];
    std::sort(legs[s].begin(), legs[s].end());
    for (int w : legs[s]) count[w]++;
    live += std::ldexp(1.0, (int)legs[s].size());
  }
  for (int w = 0; w < net.W; ++w)
    if (net.is_out[w]) count[w]++;
  double flops = 0, write = 0, peak = live;
  int mx = S == 1 ? (int)legs[0].size() : 0;
  for (int i = 0; i + 1 < S; ++i) {
    int a = path[2 * i], b = path[2 * i + 1], k = S + i;
    int nu;
    merged(legs[a], legs[b], count, legs[k], nu);
    for (int w : legs[a]) count[w]--;
    for (int w : legs[b]) count[w]--;
    for (int w : legs[k]) count[w]++;
    flops += std::ldexp(1.0, nu);
    write += std::ldexp(1.0, (int)legs[k].size());
    mx = std::max(mx, (int)legs[k].size());
    live += std::ldexp(1.0, (int)legs[k].size());
    peak = std::max(peak, live);
    live -= std::ldexp(1.0, (int)legs[a].size()) + std::ldexp(1.0, (int)legs[b].size());
  }
  tc.log2_flops = flops > 0 ? std::log2(flops) : 0;
  tc.log2_write = write > 0 ? std::log2(write) : 0;
  tc.log2_max = mx;
  tc.log2_peak = peak > 0 ? std::log2(peak) : 0;
  return tc;
}

namespace {

struct Bits {
  int nw = 0;
  std::vector<uint64_t> d;
  void init(int n, int nv) {
    nw = (nv + 63) / 64;
    d.assign((size_t)n * nw, 0);
  }
  uint64_t* row(int v) { return d.data() + (size_t)v * nw; }
  const uint64_t* row(int v) const { return d.data() + (size_t)v * nw; }
};

// Elimination order over `nv` vertices (cliques given), skipping `keep` vertices.
// Returns the order and its width (max degree at elimination).
// With `prior` (one value per vertex), the score is fill + lam * prior[v] and ties go to
// the smaller prior: a learned proposal steering the adaptive heuristic.  With
// `stop_above`, the pass stops as soon as the width exceeds it (the order is then partial
// and the returned width is some value > stop_above).
int elim_order(int nv, const std::vector<std::vector<int>>& cliques, const std::vector<char>& keep,
               bool min_fill, std::mt19937_64* rng, std::vector<int>& order, const float* prior = nullptr,
               float lam = 0.f, int stop_above = INT_MAX) {
  thread_local Bits adj;
  thread_local std::vector<char> gone;
  thread_local std::vector<int> deg, cand, nb;
  adj.init(nv, nv);
  for (const auto& c : cliques)
    for (int u : c)
      for (int v : c)
        if (u != v) adj.row(u)[v >> 6] |= 1ull << (v & 63);
  gone.assign(nv, 0);
  deg.assign(nv, 0);
  for (int v = 0; v < nv; ++v) {
    int d = 0;
    for (int i = 0; i < adj.nw; ++i) d += __builtin_popcountll(adj.row(v)[i]);
    deg[v] = d;
  }
  order.clear();
  int width = 0;
  int remaining = 0;
  for (int v = 0; v < nv; ++v)
    if (!keep[v]) ++remaining;
  if (!min_fill && !prior && !rng) {
    // Deterministic min-degree: vertices in one bit row per degree, so a pick is the
    // lowest set bit of the first non-empty row -- the first minimum in index order, as
    // the scan below finds it.
    thread_local Bits bucket;
    bucket.init(std::max(1, nv), nv);
    for (int v = 0; v < nv; ++v)
      if (!keep[v]) bucket.row(deg[v])[v >> 6] |= 1ull << (v & 63);
    while (remaining > 0) {
      int v = -1, d = 0;
      for (; d < nv && v < 0; ++d) {
        const uint64_t* b = bucket.row(d);
        for (int i = 0; i < adj.nw; ++i)
          if (b[i]) {
            v = i * 64 + __builtin_ctzll(b[i]);
            break;
          }
      }
      width = std::max(width, deg[v]);
      if (width > stop_above) return width;
      order.push_back(v);
      gone[v] = 1;
      --remaining;
      bucket.row(deg[v])[v >> 6] &= ~(1ull << (v & 63));
      uint64_t* rv = adj.row(v);
      nb.clear();
      for (int i = 0; i < adj.nw; ++i)
        for (uint64_t m = rv[i]; m; m &= m - 1) nb.push_back(i * 64 + __builtin_ctzll(m));
      for (int u : nb) {
        uint64_t* ru = adj.row(u);
        const bool listed = !keep[u] && !gone[u];
        if (listed) bucket.row(deg[u])[u >> 6] &= ~(1ull << (u & 63));
        ru[v >> 6] &= ~(1ull << (v & 63));
        for (int i = 0; i < adj.nw; ++i) ru[i] |= rv[i];
        ru[u >> 6] &= ~(1ull << (u & 63));
        int dd = 0;
        for (int i = 0; i < adj.nw; ++i) dd += __builtin_popcountll(ru[i]);
        deg[u] = dd;
        if (listed) bucket.row(dd)[u >> 6] |= 1ull << (u & 63);
      }
    }
    return width;
  }
  while (remaining > 0) {
    long best = -1;
    double bestp = 0.0;
    int bestv = -1;
    cand.clear();
    for (int v = 0; v < nv; ++v) {
      if (gone[v] || keep[v]) continue;
      long score;
      if (min_fill) {
        long fill = 0;
        const uint64_t* rv = adj.row(v);
        if (adj.nw == 1) {  // at most 64 vertices: one word per row
          const uint64_t r0 = rv[0];
          for (uint64_t m = r0; m; m &= m - 1) {
            const int u = __builtin_ctzll(m);
            fill += __builtin_popcountll(r0 & ~adj.d[u] & ~(1ull << u));
          }
        }
        for (int i = 0; i < adj.nw && adj.nw > 1; ++i) {
          uint64_t m = rv[i];
          while (m) {
            int u = i * 64 + __builtin_ctzll(m);
            m &= m - 1;
            const uint64_t* ru = adj.row(u);
            for (int t = 0; t < adj.nw; ++t) {
              uint64_t x = rv[t] & ~ru[t];
              if (t == (u >> 6)) x &= ~(1ull << (u & 63));
              fill += __builtin_popcountll(x);
            }
          }
        }
        score = fill * 4096 + deg[v];
      } else {
        score = deg[v];
      }
      if (prior) {
        // score = fill * 4096 + deg (min-fill) or deg; steer by lam * prior in fill units.
        const double p = (double)score / 4096.0 + (double)lam * prior[v];
        if (bestv < 0 || p < bestp || (p == bestp && prior[v] < prior[bestv])) {
          bestp = p;
          bestv = v;
        }
        continue;
      }
      if (!rng) {  // deterministic: the first minimum in index order, no tie list
        if (best < 0 || score < best) {
          best = score;
          bestv = v;
        }
        continue;
      }
      if (best < 0 || score < best) {
        best = score;
        cand.clear();
        cand.push_back(v);
      } else if (score == best) {
        cand.push_back(v);
      }
    }
    int v = bestv;
    if (!prior && rng) {
      v = cand[0];
      if (cand.size() > 1) v = cand[(*rng)() % cand.size()];
    }
    width = std::max(width, deg[v]);
    if (width > stop_above) return width;
    order.push_back(v);
    gone[v] = 1;
    --remaining;
    // Connect neighbours, remove v.
    uint64_t* rv = adj.row(v);
    nb.clear();
    for (int i = 0; i < adj.nw; ++i) {
      uint64_t m = rv[i];
      while (m) {
        nb.push_back(i * 64 + __builtin_ctzll(m));
        m &= m - 1;
      }
    }
    for (int u : nb) {
      uint64_t* ru = adj.row(u);
      ru[v >> 6] &= ~(1ull << (v & 63));
      for (int i = 0; i < adj.nw; ++i) ru[i] |= rv[i];
      ru[u >> 6] &= ~(1ull << (u & 63));
      int d = 0;
      for (int i = 0; i < adj.nw; ++i) d += __builtin_popcountll(ru[i]);
      deg[u] = d;
    }
  }
  return width;
}

// Tree from an elimination order: for each eliminated wire merge all live tensors
// holding it, smallest first; then merge what is left, smallest first.
void order_to_path(const BitNet& net, const std::vector<int>& order, std::vector<int32_t>& path) {
  const int S = net.S;
  path.clear();
  if (S <= 1) return;
  thread_local std::vector<std::vector<int>> legs, holders;
  thread_local std::vector<int> count, group;
  thread_local std::vector<char> alive;
  reset_lists(legs, (size_t)(2 * S - 1));
  reset_lists(holders, (size_t)net.W);
  count.assign(net.W, 0);
  alive.assign(2 * S - 1, 0);
  for (int s = 0; s < S; ++s) {
    legs[s] = net.slot_wires[s];
    std::sort(legs[s].begin(), legs[s].end());
    for (int w : legs[s]) {
      count[w]++;
      holders[w].push_back(s);
    }
    alive[s] = 1;
  }
  for (int w = 0; w < net.W; ++w)
    if (net.is_out[w]) count[w]++;
  int next = S;
  auto merge = [&](int a, int b) {
    int k = next++;
    int nu;
    merged(legs[a], legs[b], count, legs[k], nu);
    for (int w : legs[a]) count[w]--;
    for (int w : legs[b]) count[w]--;
    for (int w : legs[k]) {
      count[w]++;
      holders[w].push_back(k);
    }
    alive[a] = alive[b] = 0;
    alive[k] = 1;
    path.push_back(a);
    path.push_back(b);
    return k;
  };
  auto by_size = [&](int x, int y) {
    if (legs[x].size() != legs[y].size()) return legs[x].size() > legs[y].size();
    return x > y;
  };
  group.clear();
  for (int v : order) {
    group.clear();
    for (int t : holders[v])
      if (alive[t]) group.push_back(t);
    std::sort(group.begin(), group.end());
    group.erase(std::unique(group.begin(), group.end()), group.end());
    // Min-heap by size (pop the two smallest).
    std::make_heap(group.begin(), group.end(), by_size);
    while (group.size() > 1) {
      std::pop_heap(group.begin(), group.end(), by_size);
      int a = group.back();
      group.pop_back();
      std::pop_heap(group.begin(), group.end(), by_size);
      int b = group.back();
      group.pop_back();
      int k = merge(std::min(a, b), std::max(a, b));
      group.push_back(k);
      std::push_heap(group.begin(), group.end(), by_size);
    }
  }
  group.clear();
  for (int t = 0; t < next; ++t)
    if (alive[t]) group.push_back(t);
  std::make_heap(group.begin(), group.end(), by_size);
  while (group.size() > 1) {
    std::pop_heap(group.begin(), group.end(), by_size);
    int a = group.back();
    group.pop_back();
    std::pop_heap(group.begin(), group.end(), by_size);
    int b = group.back();
    group.pop_back();
    int k = merge(std::min(a, b), std::max(a, b));
    group.push_back(k);
    std::push_heap(group.begin(), group.end(), by_size);
  }
}

bool cost_less(const TreeCost& a, const TreeCost& b) {
  if (a.log2_max != b.log2_max) return a.log2_max < b.log2_max;
  return a.log2_flops < b.log2_flops - 1e-12;
}

}  // namespace

void order_plan(const BitNet& net, const std::vector<int>& order, std::vector<int32_t>& path,
                TreeCost* cost) {
  order_to_path(net, order, path);
  if (cost) *cost = tree_cost(net, path);
}

void elim_plan(const BitNet& net, int trials, uint64_t seed, std::vector<int32_t>& path,
               TreeCost* cost, std::vector<int>* best_order) {
  path.clear();
  if (best_order) best_order->clear();
  if (net.S <= 1) {
    if (cost) *cost = tree_cost(net, path);
    if (best_order) {
      std::vector<char> keep(net.W, 0);
      for (int w = 0; w < net.W; ++w) keep[w] = net.is_out[w];
      elim_order(net.W, net.slot_wires, keep, false, nullptr, *best_order);
    }
    return;
  }
  thread_local std::vector<char> keep;
  thread_local std::vector<int> order;
  thread_local std::vector<int32_t> cur;
  keep.assign(net.is_out.begin(), net.is_out.end());
  std::mt19937_64 rng(seed);
  TreeCost best{};
  bool have = false;
  const bool fill_ok = net.W <= 600;
  for (int t = 0; t < std::max(1, trials); ++t) {
    bool min_fill = fill_ok && (t % 2 == 0);
    elim_order(net.W, net.slot_wires, keep, min_fill, t == 0 ? nullptr : &rng, order);
    order_to_path(net, order, cur);
    TreeCost c = tree_cost(net, cur);
    if (!have || cost_less(c, best)) {
      best = c;
      path = cur;
      have = true;
      if (best_order) *best_order = order;
    }
  }
  if (cost) *cost = best;
}

void prior_plan(const BitNet& net, const float* prior, float lam, std::vector<int32_t>& path, TreeCost* cost,
                std::vector<int>* order_out) {
  std::vector<char> keep(net.W, 0);
  for (int w = 0; w < net.W; ++w) keep[w] = net.is_out[w];
  std::vector<int> order;
  elim_order(net.W, net.slot_wires, keep, net.W <= 600, nullptr, order, prior, lam);
  order_to_path(net, order, path);
  if (cost) *cost = tree_cost(net, path);
  if (order_out) *order_out = order;
}

int mindeg_width(int nv, const std::vector<std::vector<int>>& clauses_vars, int stop_above) {
  thread_local std::vector<char> keep;
  thread_local std::vector<int> order;
  keep.assign(nv, 0);
  return elim_order(nv, clauses_vars, keep, false, nullptr, order, nullptr, 0.f, stop_above);
}

namespace {

struct Allocator {
  // First-fit over a free list of [off, off+len) blocks sorted by offset (a few dozen
  // blocks at most: a sorted array beats a tree and is reused across jobs).
  std::vector<std::pair<int64_t, int64_t>> free_;  // (off, len), ascending off
  int64_t top = 0;
  void reset() {
    free_.clear();
    top = 0;
  }
  static int64_t align(int64_t n) { return (n + 15) & ~int64_t(15); }
  int64_t alloc(int64_t n) {
    n = align(n);
    for (size_t i = 0; i < free_.size(); ++i) {
      if (free_[i].second >= n) {
        const int64_t off = free_[i].first, len = free_[i].second;
        if (len > n) free_[i] = {off + n, len - n};  // stays below the next block
        else free_.erase(free_.begin() + (std::ptrdiff_t)i);
        return off;
      }
    }
    // Extend the top, merging with a free block that ends at the top.
    if (!free_.empty() && free_.back().first + free_.back().second == top) {
      const int64_t off = free_.back().first;
      free_.pop_back();
      top = off + n;
      return off;
    }
    const int64_t off = top;
    top += n;
    return off;
  }
  void release(int64_t off, int64_t n) {
    n = align(n);
    auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(off, int64_t(0)),
                               [](const std::pair<int64_t, int64_t>& a, const std::pair<int64_t, int64_t>& b) {
                                 return a.first < b.first;
                               });
    it = free_.insert(it, {off, n});
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

void permute_leaf(const uint8_t* src, const std::vector<int>& src_wires,
                  const std::vector<int>& dst_wires, uint8_t* dst) {
  const int r = (int)dst_wires.size();
  std::vector<int> spos(r);
  for (int i = 0; i < r; ++i) {
    int p = -1;
    for (size_t j = 0; j < src_wires.size(); ++j)
      if (src_wires[j] == dst_wires[i]) p = (int)(src_wires.size() - 1 - j);
    spos[i] = p;
  }
  for (uint32_t o = 0; o < (1u << r); ++o) {
    uint32_t s = 0;
    for (int i = 0; i < r; ++i)
      if ((o >> (r - 1 - i)) & 1) s |= 1u << spos[i];
    dst[o] = src[s];
  }
}

}  // namespace

int compile_job(const BitNet& net_in, const std::vector<int32_t>& path_in,
                const std::vector<const uint8_t*>& leaves_in, Job& job) {
  // Dangling out wires get an all-ones rank-1 leaf merged at the end
  // (SimpleContraction's spider(t, 1) products, simple.py:281-285).  The network, path
  // and leaves are copied only then.
  BitNet net_copy;
  std::vector<int32_t> path_copy;
  std::vector<const uint8_t*> leaves_copy;
  const BitNet* netp = &net_in;
  const std::vector<int32_t>* pathp = &path_in;
  const std::vector<const uint8_t*>* leavesp = &leaves_in;
  static const uint8_t kOnes[2] = {1, 1};
  static const uint8_t kOne[1] = {1};
  {
    thread_local std::vector<char> used;
    used.assign(net_in.W, 0);
    for (auto& ws : net_in.slot_wires)
      for (int w : ws) used[w] = 1;
    std::vector<int> dangling;
    for (int w : net_in.out_wires)
      if (!used[w]) dangling.push_back(w);
    if (!dangling.empty()) {
      net_copy = net_in;
      path_copy = path_in;
      leaves_copy = leaves_in;
      netp = &net_copy;
      pathp = &path_copy;
      leavesp = &leaves_copy;
      BitNet& net = net_copy;
      std::vector<int32_t>& path = path_copy;
      std::vector<const uint8_t*>& leaves = leaves_copy;
      const int S0 = net.S;
      int root = S0 == 0 ? -1 : (S0 == 1 ? 0 : 2 * S0 - 2);
      // Re-number merges: existing merge ids shift by the number of new leaves.
      const int add = (int)dangling.size();
      for (auto& p : path)
        if (p >= S0) p += add;
      if (root >= S0) root += add;
      for (int w : dangling) {
        net.slot_wires.push_back({w});
        leaves.push_back(kOnes);
      }
      net.S = S0 + add;
      int next = net.S + (int)path.size() / 2;
      for (int i = 0; i < add; ++i) {
        int leaf = S0 + i;
        if (root < 0) {
          root = leaf;
          continue;
        }
        path.push_back(std::min(root, leaf));
        path.push_back(std::max(root, leaf));
        root = next++;
      }
    }
  }
  const BitNet& net = *netp;
  const std::vector<int32_t>& path = *pathp;
  const std::vector<const uint8_t*>& leaves = *leavesp;
  const int S = net.S;
  job = Job();
  job.steps.reserve((size_t)std::max(1, S));
  for (auto& ws : net.slot_wires) {
    if ((int)ws.size() > TS_MAX_BITS) return TSP_EINVAL;
    int u[TS_MAX_BITS];
    std::copy(ws.begin(), ws.end(), u);
    std::sort(u, u + ws.size());
    if (std::adjacent_find(u, u + ws.size()) != u + ws.size()) return TSP_EINVAL;  // diagonals are extracted by the caller
  }
  if (S == 0) {
    job.leaf_bytes.assign(16, 0);
    job.leaf_bytes[0] = kOne[0];
    job.arena_elems = 16;
    job.result_off = 0;
    job.result_rank = 0;
    return TSP_OK;
  }
  thread_local std::vector<std::vector<int>> legs;
  if (!path_legs(net, path, legs)) return TSP_EINVAL;
  const int root = S == 1 ? 0 : 2 * S - 2;
  // The root must hold exactly the out wires (dangling ones were added above).
  {
    std::vector<int> o = net.out_wires;
    std::sort(o.begin(), o.end());
    if (std::adjacent_find(o.begin(), o.end()) != o.end()) return TSP_EINVAL;
    if (S > 1 && o != legs[root]) return TSP_EINVAL;
  }
  const int nodes = std::max(1, 2 * S - 1);
  thread_local std::vector<std::vector<int>> layout;
  reset_lists(layout, (size_t)nodes);
  // Unary root fix-up when a single leaf is not already the out tensor.
  bool unary = false;
  if (S == 1) {
    std::vector<int> l0 = net.slot_wires[0];
    if (l0 != net.out_wires) unary = true;
    layout[0] = l0;
  } else {
    layout[root] = net.out_wires;
  }
  thread_local std::vector<std::array<std::vector<int>, 3>> sums;  // Kc, La, Lb
  sums.resize(std::max(0, S - 1));
  thread_local std::vector<int> common, kc, la, lb, lay_a, lay_b, csum;
  for (int i = S - 2; i >= 0; --i) {
    const int a = path[2 * i], b = path[2 * i + 1], k = S + i;
    const auto &A = legs[a], &B = legs[b], &K = legs[k];
    common.clear();
    kc.clear();
    la.clear();
    lb.clear();
    lay_a.clear();
    lay_b.clear();
    std::set_intersection(A.begin(), A.end(), B.begin(), B.end(), std::back_inserter(common));
    for (int w : common)
      if (!std::binary_search(K.begin(), K.end(), w)) kc.push_back(w);
    for (int w : A)
      if (!std::binary_search(K.begin(), K.end(), w) && !std::binary_search(B.begin(), B.end(), w))
        la.push_back(w);
    for (int w : B)
      if (!std::binary_search(K.begin(), K.end(), w) && !std::binary_search(A.begin(), A.end(), w))
        lb.push_back(w);
    for (int w : layout[k]) {
      if (std::binary_search(A.begin(), A.end(), w)) lay_a.push_back(w);
      if (std::binary_search(B.begin(), B.end(), w)) lay_b.push_back(w);
    }
    lay_a.insert(lay_a.end(), la.begin(), la.end());
    lay_a.insert(lay_a.end(), kc.begin(), kc.end());
    lay_b.insert(lay_b.end(), lb.begin(), lb.end());
    lay_b.insert(lay_b.end(), kc.begin(), kc.end());
    layout[a].swap(lay_a);
    layout[b].swap(lay_b);
    sums[i][0].assign(kc.begin(), kc.end());
    sums[i][1].assign(la.begin(), la.end());
    sums[i][2].assign(lb.begin(), lb.end());
  }
  // Leaves in their layouts, packed at the start of the arena.
  thread_local Allocator al;
  al.reset();
  thread_local std::vector<int64_t> off;
  off.assign(nodes, -1);
  int64_t leaf_end = 0;
  for (int s = 0; s < S; ++s) {
    off[s] = al.alloc(int64_t(1) << layout[s].size());
    leaf_end = std::max(leaf_end, off[s] + (int64_t(1) << layout[s].size()));
  }
  job.leaf_bytes.assign((size_t)Allocator::align(leaf_end), 0);
  for (int s = 0; s < S; ++s) {
    if (!leaves[s]) return TSP_EINVAL;
    permute_leaf(leaves[s], net.slot_wires[s], layout[s], job.leaf_bytes.data() + off[s]);
  }
  int max_rank = 0;
  for (int s = 0; s < S; ++s) max_rank = std::max(max_rank, (int)layout[s].size());
  thread_local std::vector<uint8_t> pa, pb;  // wire -> bit position in lhs / rhs (TS_NOPOS: absent)
  pa.assign(net.W, TS_NOPOS);
  pb.assign(net.W, TS_NOPOS);
  auto emit = [&](int a, int b, int k, const std::vector<int>& csum) -> int {
    ts_step st;
    std::memset(&st, 0, sizeof(st));
    std::memset(st.o_lpos, TS_NOPOS, sizeof(st.o_lpos));
    std::memset(st.o_rpos, TS_NOPOS, sizeof(st.o_rpos));
    std::memset(st.c_lpos, TS_NOPOS, sizeof(st.c_lpos));
    std::memset(st.c_rpos, TS_NOPOS, sizeof(st.c_rpos));
    const auto& LO = layout[k];
    if ((int)LO.size() > TS_MAX_BITS || (int)csum.size() > TS_MAX_BITS) return TSP_EINVAL;
    st.lhs_off = (uint32_t)off[a];
    st.rhs_off = b >= 0 ? (uint32_t)off[b] : 0;
    st.out_off = (uint32_t)off[k];
    st.lhs_rank = (uint8_t)layout[a].size();
    st.rhs_rank = b >= 0 ? (uint8_t)layout[b].size() : 0;
    st.out_rank = (uint8_t)LO.size();
    st.sum_rank = (uint8_t)csum.size();
    const int ro = (int)LO.size();
    // Bit position of each wire in the operands' layouts (first occurrence, counted from
    // the innermost axis), through per-wire tables cleared again right after use.
    const auto& LA = layout[a];
    const std::vector<int>* LB = b >= 0 ? &layout[b] : nullptr;
    for (int i = (int)LA.size() - 1; i >= 0; --i) pa[LA[i]] = (uint8_t)(LA.size() - 1 - i);
    if (LB)
      for (int i = (int)LB->size() - 1; i >= 0; --i) pb[(*LB)[i]] = (uint8_t)(LB->size() - 1 - i);
    for (int j = 0; j < ro; ++j) {
      int w = LO[ro - 1 - j];
      st.o_lpos[j] = pa[w];
      st.o_rpos[j] = pb[w];
    }
    bool kinner = true;
    for (size_t j = 0; j < csum.size(); ++j) {
      int w = csum[j];
      st.c_lpos[j] = pa[w];
      st.c_rpos[j] = pb[w];
      if (st.c_lpos[j] != j || (b >= 0 && st.c_rpos[j] != j)) kinner = false;
    }
    for (int w : LA) pa[w] = TS_NOPOS;
    if (LB)
      for (int w : *LB) pb[w] = TS_NOPOS;
    // Every operand bit must be covered exactly once.
    for (int side = 0; side < (b >= 0 ? 2 : 1); ++side) {
      int rank = side ? st.rhs_rank : st.lhs_rank;
      uint32_t seen = 0;
      for (int j = 0; j < ro; ++j) {
        uint8_t p = side ? st.o_rpos[j] : st.o_lpos[j];
        if (p == TS_NOPOS) continue;
        if (seen & (1u << p)) return TSP_EINVAL;
        seen |= 1u << p;
      }
      for (size_t j = 0; j < csum.size(); ++j) {
        uint8_t p = side ? st.c_rpos[j] : st.c_lpos[j];
        if (p == TS_NOPOS) continue;
        if (seen & (1u << p)) return TSP_EINVAL;
        seen |= 1u << p;
      }
      if (seen != (rank == 32 ? 0xFFFFFFFFu : ((1u << rank) - 1))) return TSP_EINVAL;
    }
    st.flags = (kinner ? TS_STEP_KINNER : 0) | (b < 0 ? TS_STEP_NORHS : 0);
    job.steps.push_back(st);
    max_rank = std::max(max_rank, ro);
    return TSP_OK;
  };
  if (S == 1) {
    if (unary) {
      // out = sum over the leaf's non-out wires, in out order.
      csum.clear();
      for (int w : layout[0])
        if (std::find(net.out_wires.begin(), net.out_wires.end(), w) == net.out_wires.end())
          csum.push_back(w);
      // Put the summed wires of the unary step in the leaf's own order (innermost last).
      std::reverse(csum.begin(), csum.end());
      layout.resize(2);
      layout[1] = net.out_wires;
      off.resize(2);
      off[1] = al.alloc(int64_t(1) << layout[1].size());
      int rc = emit(0, -1, 1, csum);
      if (rc) return rc;
      job.result_off = (int32_t)off[1];
      job.result_rank = (int32_t)layout[1].size();
    } else {
      job.result_off = (int32_t)off[0];
      job.result_rank = (int32_t)layout[0].size();
    }
    job.arena_elems = Allocator::align(al.top);
    job.max_rank = max_rank;
    return TSP_OK;
  }
  for (int i = 0; i < S - 1; ++i) {
    const int a = path[2 * i], b = path[2 * i + 1], k = S + i;
    off[k] = al.alloc(int64_t(1) << layout[k].size());
    // Summed bits: contracted (innermost, LSB first), then lhs-only, then rhs-only.
    csum.clear();
    const auto& kc = sums[i][0];
    for (int j = (int)kc.size() - 1; j >= 0; --j) csum.push_back(kc[j]);
    for (int w : sums[i][1]) csum.push_back(w);
    for (int w : sums[i][2]) csum.push_back(w);
    int rc = emit(a, b, k, csum);
    if (rc) return rc;
    al.release(off[a], int64_t(1) << layout[a].size());
    al.release(off[b], int64_t(1) << layout[b].size());
  }
  job.result_off = (int32_t)off[root];
  job.result_rank = (int32_t)layout[root].size();
  job.arena_elems = Allocator::align(al.top);
  job.max_rank = max_rank;
  return TSP_OK;
}

}  // namespace tsp
