// Packing of CNF instances for the device cube search (tsc_search_pack): one pass over
// the literals, normalising each clause (duplicate literals dropped, tautologies removed)
// into the kernel's word layout, (variable << 1 | negated) per literal byte and the
// length in the top byte.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "tensorsat_prog.h"

namespace {
constexpr int kMaxVars = 128;
constexpr int64_t kMaxClauses = 64 * 12;
}  // namespace

namespace {

// Normalise the clauses of instance i; writes them at `out` when non-null.  Returns the
// number of clauses kept, or a negative TSP / TSC code.
int64_t pack_one(int n, int64_t ncl, const int32_t* len, const int32_t* L, uint32_t* out) {
  if (n < 0) return TSP_EINVAL;
  if (n > kMaxVars) return TSC_EUNSUPPORTED;
  int64_t w = 0, off = 0;
  const auto code = [](int l) { return ((uint32_t)(std::abs(l) - 1) << 1) | (l < 0 ? 1u : 0u); };
  for (int64_t c = 0; c < ncl; ++c) {
    const int k = len[c];
    if (k < 0) return TSP_EINVAL;
    if (k == 3) {
      // the common case: three literals of distinct variables, kept as they are
      const int a = L[off], b = L[off + 1], d = L[off + 2];
      const int va = std::abs(a), vb = std::abs(b), vd = std::abs(d);
      if (va != vb && va != vd && vb != vd && va && vb && vd && va <= n && vb <= n && vd <= n) {
        if (out) out[w] = (3u << 24) | code(a) | (code(b) << 8) | (code(d) << 16);
        ++w;
        off += 3;
        continue;
      }
    }
    int keep[3];
    int nk = 0;
    bool taut = false;
    for (int j = 0; j < k; ++j) {
      const int l = L[off + j];
      if (l == 0 || std::abs(l) > n) return TSP_EINVAL;
      bool dup = false;
      for (int q = 0; q < nk; ++q) {
        dup |= keep[q] == l;
        taut |= keep[q] == -l;
      }
      if (dup) continue;
      if (nk == 3) return TSC_EUNSUPPORTED;  // more than 3 distinct literals
      keep[nk++] = l;
    }
    off += k;
    if (taut) continue;  // always satisfied: dropped
    if (out) {
      uint32_t word = (uint32_t)nk << 24;
      for (int q = 0; q < nk; ++q) word |= code(keep[q]) << (8 * q);
      out[w] = word;
    }
    ++w;
  }
  return w > kMaxClauses ? (int64_t)TSC_EUNSUPPORTED : w;
}

template <typename F>
void parallel_for(int32_t num, int threads, F&& f) {
  threads = std::max(1, std::min<int>(threads, (num + 255) / 256));
  if (threads == 1) {
    for (int32_t i = 0; i < num; ++i) f(i);
    return;
  }
  std::atomic<int32_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (;;) {
        const int32_t b = next.fetch_add(256);
        if (b >= num) break;
        for (int32_t i = b; i < std::min(num, b + 256); ++i) f(i);
      }
    });
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int tsc_search_pack(int32_t num, const int32_t* nvars, const int64_t* nclauses, const int32_t* const* lens,
                               const int32_t* const* lits, uint32_t* words_out, int32_t* inst_out, int64_t* total_out,
                               int32_t* bad_out, int32_t threads) {
  if (num < 0 || (num > 0 && (!nvars || !nclauses || !lens || !lits || !inst_out)) || !total_out) return TSP_EINVAL;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  // One parallel pass packs instance i at its input clause offset (packing never adds
  // clauses, and words_out holds at least the input clause count when given); the
  // kept ranges are then moved down in order over the clauses dropped before them.
  std::vector<int64_t> in_off((size_t)num + 1, 0), kept((size_t)num);
  for (int32_t i = 0; i < num; ++i) {
    if (nclauses[i] < 0) {
      if (bad_out) *bad_out = i;
      return TSP_EINVAL;
    }
    in_off[i + 1] = in_off[i] + nclauses[i];
  }
  parallel_for(num, threads, [&](int32_t i) {
    kept[i] = pack_one(nvars[i], nclauses[i], lens[i], lits[i], words_out ? words_out + in_off[i] : nullptr);
  });
  int64_t w = 0;
  for (int32_t i = 0; i < num; ++i) {
    if (kept[i] < 0) {
      if (bad_out) *bad_out = i;
      return (int)kept[i];
    }
    if (words_out && w != in_off[i] && kept[i] > 0)
      std::memmove(words_out + w, words_out + in_off[i], (size_t)kept[i] * sizeof(uint32_t));
    inst_out[4 * i + 0] = (int32_t)w;
    inst_out[4 * i + 1] = (int32_t)kept[i];
    inst_out[4 * i + 2] = nvars[i];
    inst_out[4 * i + 3] = 0;
    w += kept[i];
  }
  if (bad_out) *bad_out = -1;
  *total_out = w;
  return TSP_OK;
}

// Start jobs of a batch (the search's job records, 64 bytes each).
namespace {

struct StartJob {
  int32_t instance, depth, split, own_from;
  uint32_t prefix, reserved[3];
  uint64_t assigned[2], value[2];
};
static_assert(sizeof(StartJob) == 64, "job record layout");

// Estimated log2 branch nodes of a random 3-SAT-like instance: linear in n with a slope
// that grows with the clause ratio up to the threshold (4.26) and falls past it.
double log2_search_size(int32_t num_vars, int32_t num_clauses) {
  const double n = num_vars, r = (double)num_clauses / std::max(n, 1.0);
  const double lo = 0.024, hi = 0.054;
  const double slope = r <= 4.26 ? lo + (hi - lo) * std::min(std::max((r - 3.5) / 0.76, 0.0), 1.0)
                                 : hi * 4.26 / std::max(r, 1e-9);
  return 2.5 + slope * n;
}

}  // namespace

extern "C" int tsc_search_plan(int32_t num, const int32_t* inst, double split_target, int32_t max_split,
                               double tail_fraction, int32_t tail_boost, void* jobs_out, int64_t capacity,
                               int64_t* njobs_out) {
  if (num < 0 || (num > 0 && !inst) || !njobs_out || max_split < 0 || capacity < 0) return TSP_EINVAL;
  if (!(tail_fraction >= 0.0 && tail_fraction <= 1.0) || tail_boost < 0 || tail_boost > 31) return TSP_EINVAL;
  const int32_t cap_split = std::min<int32_t>(max_split, 31);
  std::vector<int32_t> split((size_t)num, 0);
  std::vector<double> key((size_t)num), est((size_t)num);
  const double lt = split_target > 0 ? std::log2(split_target) : 0.0;
  for (int32_t i = 0; i < num; ++i) est[i] = log2_search_size(inst[4 * i + 2], inst[4 * i + 1]);
  // The tail: instances whose estimate is at least the k-th largest, k = ceil(tail_fraction
  // * num), get tail_boost more split levels -- the longest searches of a batch set its
  // time, and splitting only them shortens the critical path without multiplying the
  // work of the rest.
  double thr = std::numeric_limits<double>::infinity();
  if (split_target > 0 && tail_boost > 0 && tail_fraction > 0 && num > 0) {
    const int64_t k = std::min<int64_t>(num, std::max<int64_t>(1, (int64_t)std::ceil(tail_fraction * (double)num)));
    std::vector<double> sorted(est);
    std::nth_element(sorted.begin(), sorted.begin() + (num - k), sorted.end());
    thr = sorted[(size_t)(num - k)];
  }
  int64_t total = 0;
  for (int32_t i = 0; i < num; ++i) {
    if (split_target > 0) {
      const double s = std::ceil(est[i] - lt);
      int32_t v = s <= 0 ? 0 : (s >= cap_split ? cap_split : (int32_t)s);
      if (est[i] >= thr) v = std::min<int32_t>(v + tail_boost, cap_split);
      split[i] = v;
    }
    key[i] = -(est[i] - (double)split[i]);
    total += int64_t(1) << split[i];
  }
  *njobs_out = total;
  if (total > capacity) return TSP_ENOMEM;
  std::vector<int32_t> order((size_t)num);
  for (int32_t i = 0; i < num; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  StartJob* r = static_cast<StartJob*>(jobs_out);
  // One thread: starting threads costs more than writing these 1-2 MB.
  for (int32_t i : order) {
    const uint32_t nj = 1u << split[i];
    std::memset(static_cast<void*>(r), 0, (size_t)nj * sizeof(StartJob));
    for (uint32_t p = 0; p < nj; ++p) {
      r[p].instance = i;
      r[p].split = split[i];
      r[p].prefix = p;
      // A node on the prefix levels belongs to the job whose remaining prefix bits are
      // all 0: the job counts from one level below its highest set prefix bit.
      r[p].own_from = p ? 32 - __builtin_clz(p) : 0;
    }
    r += nj;
  }
  return TSP_OK;
}
