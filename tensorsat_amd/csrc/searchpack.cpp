// Packing of CNF instances for the device cube search (tsc_search_pack): one pass over
// the literals, normalising each clause (duplicate literals dropped, tautologies removed)
// into the kernel's word layout, (variable << 1 | negated) per literal byte and the
// length in the top byte.

#include <cstdlib>

#include "tensorsat_prog.h"

namespace {
constexpr int kMaxVars = 128;
constexpr int64_t kMaxClauses = 64 * 12;
}  // namespace

extern "C" int tsc_search_pack(int32_t num, const int32_t* nvars, const int64_t* nclauses, const int32_t* const* lens,
                               const int32_t* const* lits, uint32_t* words_out, int32_t* inst_out, int64_t* total_out,
                               int32_t* bad_out) {
  if (num < 0 || (num > 0 && (!nvars || !nclauses || !lens || !lits || !inst_out)) || !total_out) return TSP_EINVAL;
  int64_t w = 0;
  for (int32_t i = 0; i < num; ++i) {
    if (bad_out) *bad_out = i;
    const int n = nvars[i];
    if (n < 0) return TSP_EINVAL;
    if (n > kMaxVars) return TSC_EUNSUPPORTED;
    const int64_t begin = w;
    const int32_t* L = lits[i];
    const int32_t* len = lens[i];
    int64_t off = 0;
    for (int64_t c = 0; c < nclauses[i]; ++c) {
      const int k = len[c];
      if (k < 0) return TSP_EINVAL;
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
      uint32_t word = (uint32_t)nk << 24;
      for (int q = 0; q < nk; ++q) {
        const int l = keep[q];
        const uint32_t b = ((uint32_t)(std::abs(l) - 1) << 1) | (l < 0 ? 1u : 0u);
        word |= b << (8 * q);
      }
      if (words_out) words_out[w] = word;
      ++w;
    }
    if (w - begin > kMaxClauses) return TSC_EUNSUPPORTED;
    inst_out[4 * i + 0] = (int32_t)begin;
    inst_out[4 * i + 1] = (int32_t)(w - begin);
    inst_out[4 * i + 2] = n;
    inst_out[4 * i + 3] = 0;
  }
  if (bad_out) *bad_out = -1;
  *total_out = w;
  return TSP_OK;
}
