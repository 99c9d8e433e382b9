// Host-side contraction planner: greedy / hyper-optimised random-greedy trees over a
// hypergraph wiring, the cotengra cost model, and greedy index slicing.
//
// Replaces the reference's cotengra planning call (tensorsat/contractions/cotengra.py:78-85)
// and the slicing of notebooks/02-random-ksat.ipynb cell 32.  The merge rule is the one
// every executor in this package follows: a wire is summed at the first merge after which
// no other live tensor and no out port holds it; otherwise it is kept.

#include "../../include/tensorsat_plan.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <random>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

struct Net {
  int S = 0, W = 0;
  std::vector<std::vector<int>> legs;  // unique sorted wires per slot
  std::vector<double> lw;              // log2 extent per wire
  std::vector<char> is_out;
};

int load(const tsp_network* n, Net& net) {
  if (!n || n->num_slots < 0 || n->num_wires < 0 || n->num_out < 0) return TSP_EINVAL;
  if (n->num_slots > 0 && (!n->slot_offsets || !n->slot_wires)) return TSP_EINVAL;
  if (n->num_wires > 0 && !n->wire_dims) return TSP_EINVAL;
  if (n->num_out > 0 && !n->out_wires) return TSP_EINVAL;
  net.S = n->num_slots;
  net.W = n->num_wires;
  net.lw.resize(net.W);
  for (int w = 0; w < net.W; ++w) {
    if (n->wire_dims[w] < 1) return TSP_EINVAL;
    net.lw[w] = std::log2((double)n->wire_dims[w]);
  }
  net.is_out.assign(net.W, 0);
  for (int i = 0; i < n->num_out; ++i) {
    int w = n->out_wires[i];
    if (w < 0 || w >= net.W) return TSP_EINVAL;
    net.is_out[w] = 1;
  }
  net.legs.resize(net.S);
  if (net.S > 0 && n->slot_offsets[0] != 0) return TSP_EINVAL;
  for (int s = 0; s < net.S; ++s) {
    int a = n->slot_offsets[s], b = n->slot_offsets[s + 1];
    if (b < a) return TSP_EINVAL;
    auto& L = net.legs[s];
    for (int i = a; i < b; ++i) {
      int w = n->slot_wires[i];
      if (w < 0 || w >= net.W) return TSP_EINVAL;
      L.push_back(w);
    }
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
  }
  return TSP_OK;
}

inline double lsize(const std::vector<int>& legs, const std::vector<double>& lw) {
  double s = 0;
  for (int w : legs) s += lw[w];
  return s;
}

// Output legs of merging a and b given live counts (count[w] = live tensors holding w
// plus one if w is an out wire).  Also returns the log2 size of the union.
inline void merged_legs(const std::vector<int>& a, const std::vector<int>& b,
                        const std::vector<int>& count, const std::vector<double>& lw,
                        std::vector<int>& out, double& l_union, double& l_out) {
  out.clear();
  l_union = 0;
  l_out = 0;
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
    l_union += lw[w];
    if (count[w] > c) {
      out.push_back(w);
      l_out += lw[w];
    }
  }
}

struct Cand {
  double score;
  int a, b;
  bool operator<(const Cand& o) const {  // max-heap on "worse"; we want min score on top
    if (score != o.score) return score > o.score;
    if (a != o.a) return a > o.a;
    return b > o.b;
  }
};

int greedy_run(const Net& net, double costmod, double temperature, uint64_t seed,
               int32_t* path) {
  const int S = net.S;
  if (S <= 1) return TSP_OK;
  const int total = 2 * S - 1;
  std::vector<std::vector<int>> legs(total);
  std::vector<double> size(total, 0.0);
  std::vector<char> alive(total, 0);
  std::vector<int> count(net.W, 0);
  std::vector<std::vector<int>> wt(net.W);
  for (int s = 0; s < S; ++s) {
    legs[s] = net.legs[s];
    size[s] = std::exp2(lsize(legs[s], net.lw));
    alive[s] = 1;
    for (int w : legs[s]) {
      count[w]++;
      wt[w].push_back(s);
    }
  }
  for (int w = 0; w < net.W; ++w)
    if (net.is_out[w]) count[w]++;

  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  std::priority_queue<Cand> pq;
  std::unordered_set<uint64_t> seen;
  seen.reserve((size_t)S * 8);
  std::vector<int> tmp;
  auto score_of = [&](int a, int b) {
    double lu, lo;
    merged_legs(legs[a], legs[b], count, net.lw, tmp, lu, lo);
    return std::exp2(lo) - costmod * (size[a] + size[b]);
  };
  auto push_pair = [&](int a, int b) {
    if (a > b) std::swap(a, b);
    uint64_t key = ((uint64_t)a << 32) | (uint32_t)b;
    if (!seen.insert(key).second) return;
    pq.push(Cand{score_of(a, b), a, b});
  };
  for (int w = 0; w < net.W; ++w) {
    auto& v = wt[w];
    for (size_t i = 0; i < v.size(); ++i)
      for (size_t j = i + 1; j < v.size(); ++j) push_pair(v[i], v[j]);
  }

  int next = S;
  int step = 0;
  auto do_merge = [&](int a, int b) {
    const int k = next++;
    double lu, lo;
    merged_legs(legs[a], legs[b], count, net.lw, legs[k], lu, lo);
    size[k] = std::exp2(lo);
    for (int w : legs[a]) count[w]--;
    for (int w : legs[b]) count[w]--;
    for (int w : legs[k]) count[w]++;
    alive[a] = alive[b] = 0;
    alive[k] = 1;
    path[2 * step] = a;
    path[2 * step + 1] = b;
    ++step;
    for (int w : legs[k]) {
      auto& v = wt[w];
      size_t o = 0;
      for (size_t i = 0; i < v.size(); ++i)
        if (alive[v[i]]) v[o++] = v[i];
      v.resize(o);
      for (int t : v) push_pair(k, t);
      v.push_back(k);
    }
  };

  std::vector<Cand> top;
  while (!pq.empty()) {
    // Collect up to 8 valid candidates with up-to-date scores.
    top.clear();
    const size_t want = temperature > 0 ? 8 : 1;
    while (!pq.empty() && top.size() < want) {
      Cand c = pq.top();
      pq.pop();
      if (!alive[c.a] || !alive[c.b]) continue;
      double fresh = score_of(c.a, c.b);
      if (fresh > c.score + 1e-9 * std::fabs(c.score) + 1e-12) {  // got worse: re-rank
        c.score = fresh;
        pq.push(c);
        continue;
      }
      c.score = fresh;
      top.push_back(c);
    }
    if (top.empty()) break;
    size_t pick = 0;
    if (temperature > 0 && top.size() > 1) {
      std::sort(top.begin(), top.end(), [](const Cand& x, const Cand& y) {
        if (x.score != y.score) return x.score < y.score;
        if (x.a != y.a) return x.a < y.a;
        return x.b < y.b;
      });
      const double cmin = top[0].score;
      const double t = temperature * std::max(1.0, std::fabs(cmin));
      double tot = 0;
      std::vector<double> wts(top.size());
      for (size_t i = 0; i < top.size(); ++i) {
        wts[i] = std::exp(-(top[i].score - cmin) / t);
        tot += wts[i];
      }
      double r = unif(rng) * tot;
      for (pick = 0; pick + 1 < top.size(); ++pick) {
        r -= wts[pick];
        if (r <= 0) break;
      }
      for (size_t i = 0; i < top.size(); ++i)
        if (i != pick) pq.push(top[i]);
    }
    do_merge(top[pick].a, top[pick].b);
  }
  // Disconnected components: merge the two smallest live tensors until one is left.
  while (step < S - 1) {
    int a = -1, b = -1;
    for (int t = 0; t < next; ++t) {
      if (!alive[t]) continue;
      if (a < 0 || size[t] < size[a]) {
        b = a;
        a = t;
      } else if (b < 0 || size[t] < size[b]) {
        b = t;
      }
    }
    if (a < 0 || b < 0) return TSP_EINVAL;
    if (a > b) std::swap(a, b);
    do_merge(a, b);
  }
  return TSP_OK;
}

int cost_run(const Net& net, const int32_t* path, const std::vector<char>& sliced_mask,
             int num_sliced, tsp_cost* out, std::vector<double>* step_lout = nullptr) {
  const int S = net.S;
  std::vector<std::vector<int>> legs(std::max(1, 2 * S - 1));
  std::vector<double> lsz(legs.size(), 0.0);
  std::vector<int> count(net.W, 0);
  std::vector<char> alive(legs.size(), 0);
  double live = 0, peak = 0;
  for (int s = 0; s < S; ++s) {
    for (int w : net.legs[s])
      if (!sliced_mask[w]) legs[s].push_back(w);
    for (int w : legs[s]) count[w]++;
    lsz[s] = lsize(legs[s], net.lw);
    live += std::exp2(lsz[s]);
    alive[s] = 1;
  }
  peak = live;
  for (int w = 0; w < net.W; ++w)
    if (net.is_out[w]) count[w]++;
  double flops = 0, write = 0, mx = (S == 1 ? lsz[0] : -INFINITY);
  if (S == 0) mx = 0;
  for (int i = 0; i < S - 1; ++i) {
    int a = path[2 * i], b = path[2 * i + 1];
    int k = S + i;
    if (a < 0 || b < 0 || a >= k || b >= k || a == b || !alive[a] || !alive[b])
      return TSP_EINVAL;
    double lu, lo;
    merged_legs(legs[a], legs[b], count, net.lw, legs[k], lu, lo);
    for (int w : legs[a]) count[w]--;
    for (int w : legs[b]) count[w]--;
    for (int w : legs[k]) count[w]++;
    alive[a] = alive[b] = 0;
    alive[k] = 1;
    lsz[k] = lo;
    flops += std::exp2(lu);
    write += std::exp2(lo);
    mx = std::max(mx, lo);
    live += std::exp2(lo);
    peak = std::max(peak, live);
    live -= std::exp2(lsz[a]) + std::exp2(lsz[b]);
    if (step_lout) step_lout->push_back(lo);
  }
  const double ns = std::ldexp(1.0, num_sliced);
  out->log2_flops = flops > 0 ? std::log2(flops * ns) : -INFINITY;
  out->log2_write = write > 0 ? std::log2(write * ns) : -INFINITY;
  out->log2_max_size = mx;
  out->log2_peak = std::log2(peak);
  out->num_slices = ns;
  return TSP_OK;
}

double objective(const tsp_cost& c, double write_weight) {
  double f = std::exp2(c.log2_flops) + write_weight * std::exp2(c.log2_write);
  return f > 0 ? std::log2(f) : -INFINITY;
}

}  // namespace

extern "C" {

int tsp_greedy(const tsp_network* n, double costmod, double temperature, uint64_t seed,
               int32_t* path_out) {
  Net net;
  int rc = load(n, net);
  if (rc) return rc;
  if (net.S > 1 && !path_out) return TSP_EINVAL;
  try {
    return greedy_run(net, costmod, temperature, seed, path_out);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
}

int tsp_tree_cost(const tsp_network* n, const int32_t* path, int32_t num_sliced,
                  const int32_t* sliced, tsp_cost* cost_out) {
  Net net;
  int rc = load(n, net);
  if (rc) return rc;
  if (!cost_out || (net.S > 1 && !path) || num_sliced < 0) return TSP_EINVAL;
  std::vector<char> mask(net.W, 0);
  for (int i = 0; i < num_sliced; ++i) {
    if (sliced[i] < 0 || sliced[i] >= net.W) return TSP_EINVAL;
    mask[sliced[i]] = 1;
  }
  try {
    return cost_run(net, path, mask, num_sliced, cost_out);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
}

int tsp_hyper(const tsp_network* n, const tsp_hyper_opts* opts, int32_t* path_out,
              tsp_cost* cost_out) {
  Net net;
  int rc = load(n, net);
  if (rc) return rc;
  if (!opts || opts->trials < 1) return TSP_EINVAL;
  if (net.S > 1 && !path_out) return TSP_EINVAL;
  const int S = net.S;
  const int T = opts->trials;
  int threads = opts->threads > 0 ? opts->threads : (int)std::thread::hardware_concurrency();
  threads = std::max(1, std::min(threads, T));
  const size_t plen = (size_t)std::max(0, 2 * (S - 1));
  struct Best {
    double key0 = INFINITY, key1 = INFINITY;
    int trial = -1;
    std::vector<int32_t> path;
    tsp_cost cost{};
  };
  std::vector<Best> best(threads);
  std::atomic<int> next{0};
  std::atomic<int> failed{0};
  std::vector<char> nomask(net.W, 0);
  auto worker = [&](int tid) {
    std::vector<int32_t> path(plen);
    Best& b = best[tid];
    for (;;) {
      int t = next.fetch_add(1);
      if (t >= T) break;
      std::mt19937_64 prng(opts->seed * 0x9E3779B97F4A7C15ull + (uint64_t)t);
      std::uniform_real_distribution<double> u(0.0, 1.0);
      double cm = 1.0, temp = 0.0;
      if (t > 0) {
        cm = opts->costmod_min + (opts->costmod_max - opts->costmod_min) * u(prng);
        double lo = std::log(std::max(1e-6, opts->temperature_min));
        double hi = std::log(std::max(1e-6, opts->temperature_max));
        temp = std::exp(lo + (hi - lo) * u(prng));
      }
      try {
        if (greedy_run(net, cm, temp, prng(), path.data()) != TSP_OK) {
          failed++;
          continue;
        }
        tsp_cost c;
        if (cost_run(net, path.data(), nomask, 0, &c) != TSP_OK) {
          failed++;
          continue;
        }
        double over = c.log2_max_size > opts->log2_size_limit ? c.log2_max_size : 0.0;
        double obj = objective(c, opts->write_weight);
        bool better = over < b.key0 || (over == b.key0 && (obj < b.key1 || (obj == b.key1 && t < b.trial)));
        if (b.trial < 0 || better) {
          b.key0 = over;
          b.key1 = obj;
          b.trial = t;
          b.path = path;
          b.cost = c;
        }
      } catch (...) {
        failed++;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < threads; ++i) pool.emplace_back(worker, i);
  worker(0);
  for (auto& th : pool) th.join();
  int bi = -1;
  for (int i = 0; i < threads; ++i) {
    if (best[i].trial < 0) continue;
    if (bi < 0) {
      bi = i;
      continue;
    }
    const Best &x = best[i], &y = best[bi];
    if (x.key0 < y.key0 || (x.key0 == y.key0 && (x.key1 < y.key1 || (x.key1 == y.key1 && x.trial < y.trial))))
      bi = i;
  }
  if (bi < 0) return failed ? TSP_ENOMEM : TSP_EINVAL;
  if (plen) std::memcpy(path_out, best[bi].path.data(), plen * sizeof(int32_t));
  if (cost_out) *cost_out = best[bi].cost;
  return TSP_OK;
}

int tsp_slice(const tsp_network* n, const int32_t* path, double target_log2_size,
              int32_t max_sliced, int32_t* sliced_out, int32_t* num_sliced_out,
              tsp_cost* cost_out) {
  Net net;
  int rc = load(n, net);
  if (rc) return rc;
  if ((net.S > 1 && !path) || !num_sliced_out || max_sliced < 0) return TSP_EINVAL;
  if (max_sliced > 0 && !sliced_out) return TSP_EINVAL;
  std::vector<char> mask(net.W, 0);
  int ns = 0;
  tsp_cost cur;
  std::vector<double> louts;
  rc = cost_run(net, path, mask, 0, &cur, &louts);
  if (rc) return rc;
  const int S = net.S;
  while (cur.log2_max_size > target_log2_size + 1e-9 && ns < max_sliced) {
    // Candidates: wires of the current largest intermediates.
    std::vector<std::vector<int>> legs(std::max(1, 2 * S - 1));
    std::vector<int> count(net.W, 0);
    for (int s = 0; s < S; ++s) {
      for (int w : net.legs[s])
        if (!mask[w]) legs[s].push_back(w);
      for (int w : legs[s]) count[w]++;
    }
    for (int w = 0; w < net.W; ++w)
      if (net.is_out[w]) count[w]++;
    std::vector<char> cand(net.W, 0);
    for (int i = 0; i < S - 1; ++i) {
      int a = path[2 * i], b = path[2 * i + 1], k = S + i;
      double lu, lo;
      merged_legs(legs[a], legs[b], count, net.lw, legs[k], lu, lo);
      for (int w : legs[a]) count[w]--;
      for (int w : legs[b]) count[w]--;
      for (int w : legs[k]) count[w]++;
      if (lo >= cur.log2_max_size - 1e-9)
        for (int w : legs[k])
          if (!net.is_out[w]) cand[w] = 1;
    }
    int bw = -1;
    tsp_cost bc{};
    for (int w = 0; w < net.W; ++w) {
      if (!cand[w]) continue;
      mask[w] = 1;
      tsp_cost c;
      cost_run(net, path, mask, ns + 1, &c);
      mask[w] = 0;
      if (bw < 0 || c.log2_max_size < bc.log2_max_size - 1e-9 ||
          (std::fabs(c.log2_max_size - bc.log2_max_size) <= 1e-9 && c.log2_flops < bc.log2_flops)) {
        bw = w;
        bc = c;
      }
    }
    if (bw < 0) break;
    mask[bw] = 1;
    sliced_out[ns++] = bw;
    cur = bc;
  }
  *num_sliced_out = ns;
  if (cost_out) *cost_out = cur;
  return TSP_OK;
}

}  // extern "C"
