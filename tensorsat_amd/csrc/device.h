// Device-side step engine for gfx950, shared by the grid-wide step kernels and the
// batched persistent executor.
//
// A step (ts_step, include/tensorsat_prog.h) computes for every out index o
//     acc(o) = sum_c lhs[L(o, c)] * rhs[R(o, c)]
// where L/R scatter the bits of o and of the summed index c to operand bit positions.
// L(o, c) = Lo(o) + Lc(c) with both parts linear in the bits, evaluated with byte-chunk
// lookup tables held in LDS: Lo(o) = T0[o & 255] + T1[(o >> 8) & 255] + ...
// Out indices are walked in blocks of 256 consecutive values, one per thread, so the
// low chunk is a per-thread constant and the higher chunks are workgroup-uniform
// (LDS broadcast); stores are byte-coalesced.
//
// Paths:
//  * sum_rank == 0: product (or copy for unary steps);
//  * TS_STEP_KINNER: the summed bits are the lowest bits of both operands in the same
//    order, so each operand's 2^sum_rank terms are contiguous: 2/4/8/16-byte vector
//    loads and popcount(lhs & rhs) on 0/1 bytes;
//  * otherwise: loop over c with LDS tables for Lc/Rc.
// Epilogues: GT0 (acc > 0) and MOD256 ((acc mod 256) != 0) for 0/1 bytes; exact uint64
// with overflow detection for the counting semiring.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tensorsat_amd.h"

namespace tsd {

constexpr int WG = 256;

// NQ byte chunks per table: 4 cover every rank up to 27 bits (TS_MAX_BITS); 2 suffice
// when every tensor has at most 16 bits (LDS-resident arenas).
template <int NQ>
struct LutT {
  uint32_t ol[NQ][256];
  uint32_t orr[NQ][256];
  uint32_t cl[NQ][256];
  uint32_t cr[NQ][256];
};
using Lut = LutT<4>;

__device__ __forceinline__ int nchunks(int bits) { return (bits + 7) >> 3; }

// Build the chunk tables of one step (all threads of the workgroup participate; the
// caller synchronises afterwards).
template <int NQ>
__device__ inline void build_lut(const ts_step& st, LutT<NQ>& L, bool need_c, int tid) {
  const int ro = st.out_rank;
  // Chunk 0 is always built: a scalar step (out_rank 0) reads entry [0][0], which must
  // be 0 and not a value left in LDS by an earlier step.
  int nqo = nchunks(ro) > 0 ? nchunks(ro) : 1;
  nqo = nqo < NQ ? nqo : NQ;  // callers keep ranks within the table (16 bits for NQ = 2)
  for (int e = tid; e < nqo * 256; e += WG) {
    const int q = e >> 8, v = e & 255;
    uint32_t al = 0, ar = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int j = q * 8 + i;
      if (j < ro && ((v >> i) & 1)) {
        const uint8_t pl = st.o_lpos[j], pr = st.o_rpos[j];
        if (pl != TS_NOPOS) al |= 1u << pl;
        if (pr != TS_NOPOS) ar |= 1u << pr;
      }
    }
    L.ol[q][v] = al;
    L.orr[q][v] = ar;
  }
  if (need_c) {
    const int nc = st.sum_rank;
    int nqc = nchunks(nc) > 0 ? nchunks(nc) : 1;
    nqc = nqc < NQ ? nqc : NQ;
    for (int e = tid; e < nqc * 256; e += WG) {
      const int q = e >> 8, v = e & 255;
      uint32_t al = 0, ar = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int j = q * 8 + i;
        if (j < nc && ((v >> i) & 1)) {
          const uint8_t pl = st.c_lpos[j], pr = st.c_rpos[j];
          if (pl != TS_NOPOS) al |= 1u << pl;
          if (pr != TS_NOPOS) ar |= 1u << pr;
        }
      }
      L.cl[q][v] = al;
      L.cr[q][v] = ar;
    }
  }
}

// High-chunk part of an out offset (o >> 8), workgroup-uniform in the block loops.
template <int NQ>
__device__ __forceinline__ uint32_t hi_off(const uint32_t (*T)[256], uint32_t hi, int nq) {
  uint32_t r = 0;
  if (nq > 1) r += T[1][hi & 255];
  if (NQ > 2 && nq > 2) r += T[NQ > 2 ? 2 : 0][(hi >> 8) & 255];
  if (NQ > 3 && nq > 3) r += T[NQ > 3 ? 3 : 0][hi >> 16];
  return r;
}

template <int NQ>
__device__ __forceinline__ uint32_t c_off(const uint32_t (*T)[256], uint32_t c, int nq) {
  uint32_t r = T[0][c & 255];
  if (nq > 1) r += T[1][(c >> 8) & 255];
  if (NQ > 2 && nq > 2) r += T[NQ > 2 ? 2 : 0][(c >> 16) & 255];
  if (NQ > 3 && nq > 3) r += T[NQ > 3 ? 3 : 0][c >> 24];
  return r;
}

__device__ __forceinline__ uint32_t popc_and(const uint8_t* a, const uint8_t* b, int nc,
                                             bool norhs) {
  // 2^nc contiguous 0/1 bytes in each operand, aligned to min(16, 2^nc).
  switch (nc) {
    case 1: {
      uint32_t x = *reinterpret_cast<const uint16_t*>(a);
      if (!norhs) x &= *reinterpret_cast<const uint16_t*>(b);
      return __popc(x);
    }
    case 2: {
      uint32_t x = *reinterpret_cast<const uint32_t*>(a);
      if (!norhs) x &= *reinterpret_cast<const uint32_t*>(b);
      return __popc(x);
    }
    case 3: {
      uint64_t x = *reinterpret_cast<const uint64_t*>(a);
      if (!norhs) x &= *reinterpret_cast<const uint64_t*>(b);
      return __popcll(x);
    }
    default: {
      uint32_t acc = 0;
      const int n16 = 1 << (nc - 4);
      for (int i = 0; i < n16; ++i) {
        uint4 x = reinterpret_cast<const uint4*>(a)[i];
        if (!norhs) {
          const uint4 y = reinterpret_cast<const uint4*>(b)[i];
          x.x &= y.x;
          x.y &= y.y;
          x.z &= y.z;
          x.w &= y.w;
        }
        acc += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
      }
      return acc;
    }
  }
}

// The 2^nc contiguous bytes of one row, nc = 2 or 3, as an integer (aligned to 2^nc).
__device__ __forceinline__ uint64_t load_row(const uint8_t* p, int nc) {
  return nc == 2 ? *reinterpret_cast<const uint32_t*>(p) : *reinterpret_cast<const uint64_t*>(p);
}

// ---- vector path for short rows ------------------------------------------------------
//
// When, in each operand, the low t out bits it holds sit right above its summed bits in
// order (the others not held at all), the rows of 2^t consecutive outputs lie in one
// chunk of 2^(sum + held) bytes: a lane loads each operand's chunk once (at most 32
// bytes), picks the row of each of its 2^t outputs out of it, and stores the 2^t results
// at once.

// t (0 = no vector path, at most 4) and the held low out bits of each operand.
__host__ __device__ inline int vec_plan(const ts_step& st, uint32_t& ml, uint32_t& mr) {
  ml = mr = 0;
  const int nc = st.sum_rank;
  if (nc > 3 || !(nc == 0 || (st.flags & TS_STEP_KINNER))) return 0;
  const bool norhs = st.flags & TS_STEP_NORHS;
  const int tmax = st.out_rank < 4 ? st.out_rank : 4;
  int best = 0, hl = 0, hr = 0;
  uint32_t cl = 0, cr = 0;
  for (int t = 1; t <= tmax; ++t) {
    const int j = t - 1;
    if (st.o_lpos[j] != TS_NOPOS) {
      if (st.o_lpos[j] != nc + hl) break;
      cl |= 1u << j;
      ++hl;
    }
    if (!norhs && st.o_rpos[j] != TS_NOPOS) {
      if (st.o_rpos[j] != nc + hr) break;
      cr |= 1u << j;
      ++hr;
    }
    if (nc + hl > 5 || nc + hr > 5) break;  // chunks of at most 32 bytes
    best = t;
    ml = cl;
    mr = cr;
  }
  return best;
}

// 2^lg bytes (lg <= 5) into 64-bit words.
__device__ __forceinline__ void load_chunk(uint64_t (&w)[4], const uint8_t* p, int lg) {
  switch (lg) {
    case 0: w[0] = *p; break;
    case 1: w[0] = *reinterpret_cast<const uint16_t*>(p); break;
    case 2: w[0] = *reinterpret_cast<const uint32_t*>(p); break;
    case 3: w[0] = *reinterpret_cast<const uint64_t*>(p); break;
    case 4: {
      const uint4 v = *reinterpret_cast<const uint4*>(p);
      w[0] = ((uint64_t)v.y << 32) | v.x;
      w[1] = ((uint64_t)v.w << 32) | v.z;
      break;
    }
    default: {
      const uint4 v = reinterpret_cast<const uint4*>(p)[0], u = reinterpret_cast<const uint4*>(p)[1];
      w[0] = ((uint64_t)v.y << 32) | v.x;
      w[1] = ((uint64_t)v.w << 32) | v.z;
      w[2] = ((uint64_t)u.y << 32) | u.x;
      w[3] = ((uint64_t)u.w << 32) | u.z;
    }
  }
}

// Row idx (2^K bytes) of a chunk held as 64-bit words (word picked by selects: no
// indexed register access).
template <int K>
__device__ __forceinline__ uint64_t chunk_row(const uint64_t (&w)[4], uint32_t idx) {
  const uint32_t off = idx << K, q = off >> 3;
  const uint64_t word = q == 0 ? w[0] : q == 1 ? w[1] : q == 2 ? w[2] : w[3];
  if constexpr (K == 3) {
    return word;
  } else {
    constexpr uint64_t MASK = (1ull << (8 << K)) - 1;
    return (word >> (8 * (off & 7))) & MASK;
  }
}

// Bits of i (< 16) at the set positions of m, packed low.
__device__ __forceinline__ uint32_t pext4(uint32_t i, uint32_t m) {
  uint32_t r = 0, b = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if ((m >> j) & 1u) r |= ((i >> j) & 1u) << b++;
  return r;
}

template <int K, int V, int NQ>
__device__ void step_vec_kv(const ts_step& st, const LutT<NQ>& L, const uint8_t* __restrict__ lhs,
                            const uint8_t* __restrict__ rhs, uint8_t* __restrict__ out, uint64_t first,
                            uint64_t stride, uint32_t ml, uint32_t mr) {
  constexpr int T = V == 2 ? 1 : V == 4 ? 2 : V == 8 ? 3 : 4;
  const bool norhs = st.flags & TS_STEP_NORHS;
  const int lgl = K + __popc(ml), lgr = K + __popc(mr);
  const int nqo = nchunks(st.out_rank);
  const uint64_t nq = 1ull << (st.out_rank - T);
  if (first >= nq) return;
  const uint32_t lo = (uint32_t)((first << T) & 255);  // the same for every q of this lane
  const uint32_t lo_l = L.ol[0][lo], lo_r = L.orr[0][lo];
  for (uint64_t q = first; q < nq; q += stride) {
    const uint64_t o = q << T;
    const uint32_t hi = (uint32_t)(o >> 8);
    uint64_t a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    load_chunk(a, lhs + lo_l + hi_off<NQ>(L.ol, hi, nqo), lgl);
    if (!norhs) load_chunk(b, rhs + lo_r + hi_off<NQ>(L.orr, hi, nqo), lgr);
    uint32_t res[(V + 3) / 4];
#pragma unroll
    for (int i = 0; i < (V + 3) / 4; ++i) res[i] = 0;
#pragma unroll
    for (int i = 0; i < V; ++i) {
      uint64_t x = chunk_row<K>(a, pext4((uint32_t)i, ml));
      if (!norhs) x &= chunk_row<K>(b, pext4((uint32_t)i, mr));
      res[i / 4] |= (uint32_t)(x != 0) << (8 * (i % 4));
    }
    uint8_t* po = out + o;
    if constexpr (V == 16) *reinterpret_cast<uint4*>(po) = make_uint4(res[0], res[1], res[2], res[3]);
    else if constexpr (V == 8) *reinterpret_cast<uint2*>(po) = make_uint2(res[0], res[1]);
    else if constexpr (V == 4) *reinterpret_cast<uint32_t*>(po) = res[0];
    else *reinterpret_cast<uint16_t*>(po) = (uint16_t)res[0];
  }
}

// Grid-stride over groups of 2^t outputs, t = vec_plan(st, ml, mr) (first / stride as
// for step_u8).  Counts are at most 8, so both epilogues reduce to count != 0.
template <int NQ>
__device__ void step_u8_vec(const ts_step& st, const LutT<NQ>& L, const uint8_t* __restrict__ lhs,
                            const uint8_t* __restrict__ rhs, uint8_t* __restrict__ out, uint64_t first,
                            uint64_t stride, int t, uint32_t ml, uint32_t mr) {
#define TS_VEC(K, V) step_vec_kv<K, V, NQ>(st, L, lhs, rhs, out, first, stride, ml, mr)
  switch (st.sum_rank * 8 + t) {
    case 0 * 8 + 1: TS_VEC(0, 2); break;
    case 0 * 8 + 2: TS_VEC(0, 4); break;
    case 0 * 8 + 3: TS_VEC(0, 8); break;
    case 0 * 8 + 4: TS_VEC(0, 16); break;
    case 1 * 8 + 1: TS_VEC(1, 2); break;
    case 1 * 8 + 2: TS_VEC(1, 4); break;
    case 1 * 8 + 3: TS_VEC(1, 8); break;
    case 1 * 8 + 4: TS_VEC(1, 16); break;
    case 2 * 8 + 1: TS_VEC(2, 2); break;
    case 2 * 8 + 2: TS_VEC(2, 4); break;
    case 2 * 8 + 3: TS_VEC(2, 8); break;
    case 2 * 8 + 4: TS_VEC(2, 16); break;
    case 3 * 8 + 1: TS_VEC(3, 2); break;
    case 3 * 8 + 2: TS_VEC(3, 4); break;
    case 3 * 8 + 3: TS_VEC(3, 8); break;
    case 3 * 8 + 4: TS_VEC(3, 16); break;
    default: break;
  }
#undef TS_VEC
}

// Boolean step over out indices o = first + k * stride (first = block * 256 + tid and
// stride = blocks * 256: the callers' grid-stride launch shape).
template <int NQ>
__device__ inline void step_u8(const ts_step& st, const LutT<NQ>& L, const uint8_t* __restrict__ lhs,
                               const uint8_t* __restrict__ rhs, uint8_t* __restrict__ out,
                               uint64_t first, uint64_t stride, int mode) {
  const int ro = st.out_rank, nc = st.sum_rank;
  const bool norhs = st.flags & TS_STEP_NORHS;
  const bool kinner = st.flags & TS_STEP_KINNER;
  const uint64_t n = 1ull << ro;
  const int nqo = nchunks(ro), nqc = nchunks(nc);
  if (first >= n) return;
  const uint32_t lo_l = L.ol[0][first & 255];
  const uint32_t lo_r = L.orr[0][first & 255];
  if (kinner && (nc == 2 || nc == 3)) {
    // Rows of 4 or 8 bytes: the workgroup takes U adjacent blocks of 256 outputs per
    // trip (thread t: output t of each), every load of the trip issued before any is
    // used -- U row pairs in flight per thread, all in one neighbourhood of memory
    // (measured faster than one output per trip for these rows, slower for 1-2 bytes).
    constexpr int U = 8;
    const uint64_t lane = first % WG, blk = first / WG, nblk = stride / WG;
    const uint64_t nb = (n + WG - 1) / WG;
    for (uint64_t b0 = blk * U; b0 < nb; b0 += nblk * U) {
      uint64_t a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t o = (b0 + u) * WG + lane;
        a[u] = 0;
        b[u] = ~0ull;
        if (o < n) {
          const uint32_t hi = (uint32_t)(o >> 8);
          a[u] = load_row(lhs + lo_l + hi_off<NQ>(L.ol, hi, nqo), nc);
          if (!norhs) b[u] = load_row(rhs + lo_r + hi_off<NQ>(L.orr, hi, nqo), nc);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t o = (b0 + u) * WG + lane;
        if (o < n) out[o] = (uint8_t)((a[u] & b[u]) != 0);  // at most 8 terms: both epilogues agree
      }
    }
    return;
  }
  for (uint64_t o = first; o < n; o += stride) {
    const uint32_t hi = (uint32_t)(o >> 8);
    const uint32_t offl = lo_l + hi_off<NQ>(L.ol, hi, nqo);
    const uint32_t offr = lo_r + hi_off<NQ>(L.orr, hi, nqo);
    uint32_t acc;
    if (nc == 0) {
      acc = lhs[offl];
      if (!norhs) acc &= rhs[offr];
    } else if (kinner) {
      acc = popc_and(lhs + offl, rhs + offr, nc, norhs);
    } else {
      acc = 0;
      const uint32_t nterm = 1u << nc;
      for (uint32_t c = 0; c < nterm; ++c) {
        const uint8_t a = lhs[offl + c_off<NQ>(L.cl, c, nqc)];
        const uint8_t b = norhs ? (uint8_t)1 : rhs[offr + c_off<NQ>(L.cr, c, nqc)];
        acc += a & b;
      }
    }
    out[o] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
  }
}

// Counting step (uint64 exact); returns true if any value of this thread overflowed.
__device__ inline bool step_u64(const ts_step& st, const Lut& L, const uint64_t* __restrict__ lhs,
                                const uint64_t* __restrict__ rhs, uint64_t* __restrict__ out,
                                uint64_t first, uint64_t stride) {
  const int ro = st.out_rank, nc = st.sum_rank;
  const bool norhs = st.flags & TS_STEP_NORHS;
  const bool kinner = st.flags & TS_STEP_KINNER;
  const uint64_t n = 1ull << ro;
  const int nqo = nchunks(ro), nqc = nchunks(nc);
  bool ovf = false;
  if (first >= n) return false;
  const uint32_t lo_l = L.ol[0][first & 255];
  const uint32_t lo_r = L.orr[0][first & 255];
  for (uint64_t o = first; o < n; o += stride) {
    const uint32_t hi = (uint32_t)(o >> 8);
    const uint32_t offl = lo_l + hi_off<4>(L.ol, hi, nqo);
    const uint32_t offr = lo_r + hi_off<4>(L.orr, hi, nqo);
    uint64_t acc = 0;
    const uint32_t nterm = 1u << nc;
    for (uint32_t c = 0; c < nterm; ++c) {
      const uint32_t cl = kinner ? c : c_off<4>(L.cl, c, nqc);
      const uint32_t cr = kinner ? c : c_off<4>(L.cr, c, nqc);
      const uint64_t a = lhs[offl + cl];
      const uint64_t b = norhs ? 1ull : rhs[offr + cr];
      uint64_t p;
      ovf |= __builtin_mul_overflow(a, b, &p);
      ovf |= __builtin_add_overflow(acc, p, &acc);
    }
    out[o] = acc;
  }
  return ovf;
}

}  // namespace tsd
