// Dense Boolean contraction steps on int8 MFMA (gfx950).
//
// A step out[o] = epilogue(sum_c lhs[..] * rhs[..]) whose summed bits are the lowest
// sum_rank bits of both operands (the K-innermost layout the compiler produces) is a
// batched matrix product with K contiguous in both operands:
//
//   rows    m : the out bits held by lhs only   (lhs offset, out offset per row)
//   columns n : the out bits held by rhs only   (rhs offset, out offset per column)
//   batch   b : the out bits held by both        (one offset triple per batch index)
//   count[b, m, n] = sum_k lhs[b, m, k] * rhs[b, n, k]      (0/1 int8, exact int32)
//
// Each workgroup computes a 64 x 64 tile of (m, n) for one batch index with 4 waves of
// 2 x 2 v_mfma_i32_16x16x64_i8; K is staged through LDS in 64-byte chunks (16-byte vector
// loads, four lanes per 64-byte row segment).  Row/column offsets are scattered by bit
// position (any interleaving of m, n and b bits in the operands and the output), so they
// are computed once per tile into LDS.  The epilogue is the step's: count > 0 (GT0) or
// (count mod 256) != 0 (MOD256); counts are exact because sum_rank <= 27 < 31.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>

#include "../../include/tensorsat_amd.h"
#include "dense.h"

namespace tsd {
namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 64;   // rows and columns per workgroup
constexpr int KC = 64;     // bytes of K per LDS stage (one MFMA K step)
constexpr int LDSW = KC + 16;  // padded LDS row (bytes)

struct DenseArgs {
  int32_t mbits, nbits, bbits, kbits;
  // Bit positions (LSB-first) of the m / n / b index bits in lhs, rhs and out.
  uint8_t m_lpos[TS_MAX_BITS], m_opos[TS_MAX_BITS];
  uint8_t n_rpos[TS_MAX_BITS], n_opos[TS_MAX_BITS];
  uint8_t b_lpos[TS_MAX_BITS], b_rpos[TS_MAX_BITS], b_opos[TS_MAX_BITS];
  int32_t mode;
  int64_t mtiles, ntiles;  // tiles along m and n
};

__device__ __forceinline__ uint64_t scatter(uint64_t idx, int nb, const uint8_t* pos) {
  uint64_t off = 0;
  for (int i = 0; i < nb; ++i) off |= ((idx >> i) & 1ull) << pos[i];
  return off;
}

__global__ __launch_bounds__(256) void k_step_mfma(const uint8_t* __restrict__ lhs, const uint8_t* __restrict__ rhs,
                                                   uint8_t* __restrict__ out, DenseArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t sA[TILE][LDSW];
  __shared__ __attribute__((aligned(16))) uint8_t sB[TILE][LDSW];
  __shared__ uint64_t rowL[TILE], rowO[TILE], colR[TILE], colO[TILE];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Tile order: n fastest, then m, then batch.  Consecutive workgroups land on different
  // XCDs, so regroup them: workgroup w runs on XCD w % 8; give each XCD a contiguous run
  // of tiles (shared rows of lhs stay in that XCD's L2).
  const uint64_t nwg = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  const uint64_t nt = w % a.ntiles;
  const uint64_t mt = (w / a.ntiles) % a.mtiles;
  const uint64_t b = w / (a.ntiles * a.mtiles);
  const uint64_t m0 = mt * TILE, n0 = nt * TILE;
  const uint64_t mcount = 1ull << a.mbits, ncount = 1ull << a.nbits;
  if (tid < TILE) {
    const uint64_t m = m0 + tid;
    rowL[tid] = m < mcount ? scatter(m, a.mbits, a.m_lpos) : 0;
    rowO[tid] = m < mcount ? scatter(m, a.mbits, a.m_opos) : 0;
  } else if (tid < 2 * TILE) {
    const uint64_t n = n0 + (tid - TILE);
    colR[tid - TILE] = n < ncount ? scatter(n, a.nbits, a.n_rpos) : 0;
    colO[tid - TILE] = n < ncount ? scatter(n, a.nbits, a.n_opos) : 0;
  }
  const uint64_t bL = scatter(b, a.bbits, a.b_lpos), bR = scatter(b, a.bbits, a.b_rpos);
  const uint64_t bO = scatter(b, a.bbits, a.b_opos);
  __syncthreads();
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  i4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (i4){0, 0, 0, 0};
  // Loader: thread t moves 16 bytes of row t/4 (segment t%4) of both tiles.
  const int lr = tid >> 2, ls = (tid & 3) * 16;
  const bool lrow_ok = m0 + lr < mcount, lcol_ok = n0 + lr < ncount;
  const uint8_t* pa = lhs + bL + rowL[lr] + ls;
  const uint8_t* pb = rhs + bR + colR[lr] + ls;
  const uint64_t K = 1ull << a.kbits;
  for (uint64_t k0 = 0; k0 < K; k0 += KC) {
    const i4 va = lrow_ok ? *reinterpret_cast<const i4*>(pa + k0) : (i4){0, 0, 0, 0};
    const i4 vb = lcol_ok ? *reinterpret_cast<const i4*>(pb + k0) : (i4){0, 0, 0, 0};
    *reinterpret_cast<i4*>(&sA[lr][ls]) = va;
    *reinterpret_cast<i4*>(&sB[lr][ls]) = vb;
    __syncthreads();
    // Lane l feeds row / column l % 16 with 16 bytes of K at 16 * (l / 16); A and B use
    // the same K assignment, so the sum over K does not depend on the hardware's order.
    const int rr = lane & 15, kk = (lane >> 4) * 16;
    i4 fa[2], fb[2];
    for (int i = 0; i < 2; ++i) fa[i] = *reinterpret_cast<const i4*>(&sA[wm + 16 * i + rr][kk]);
    for (int j = 0; j < 2; ++j) fb[j] = *reinterpret_cast<const i4*>(&sB[wn + 16 * j + rr][kk]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // Accumulator layout: lane l holds rows 4 * (l / 16) + t, column l % 16.
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 4; ++t) {
        const int r = wm + 16 * i + 4 * (lane >> 4) + t;
        const int c = wn + 16 * j + (lane & 15);
        if (m0 + r >= mcount || n0 + c >= ncount) continue;
        const int v = acc[i][j][t];
        const uint8_t bit = a.mode == TS_MODE_MOD256 ? (uint8_t)((v & 255) != 0) : (uint8_t)(v > 0);
        out[bO + rowO[r] + colO[c]] = bit;
      }
}

bool g_enabled = true;
int g_min_kbits = 6;

}  // namespace

bool dense_eligible(const ts_step& st, const void* lhs, const void* rhs) {
  if (!g_enabled) return false;
  if ((st.flags & TS_STEP_NORHS) || !(st.flags & TS_STEP_KINNER)) return false;
  if (st.sum_rank < g_min_kbits || st.sum_rank < 6) return false;
  if ((reinterpret_cast<uintptr_t>(lhs) | reinterpret_cast<uintptr_t>(rhs)) & 15) return false;
  int mb = 0, nb = 0;
  for (int j = 0; j < st.out_rank; ++j) {
    const bool l = st.o_lpos[j] != TS_NOPOS, r = st.o_rpos[j] != TS_NOPOS;
    mb += l && !r;
    nb += r && !l;
  }
  return mb >= 4 && nb >= 4;
}

int launch_dense(const uint8_t* lhs, const uint8_t* rhs, uint8_t* out, const ts_step& st, int mode,
                 hipStream_t s) {
  DenseArgs a;
  std::memset(&a, 0, sizeof(a));
  a.kbits = st.sum_rank;
  a.mode = mode;
  for (int j = 0; j < st.out_rank; ++j) {
    const bool l = st.o_lpos[j] != TS_NOPOS, r = st.o_rpos[j] != TS_NOPOS;
    if (l && r) {
      a.b_lpos[a.bbits] = st.o_lpos[j];
      a.b_rpos[a.bbits] = st.o_rpos[j];
      a.b_opos[a.bbits++] = (uint8_t)j;
    } else if (l) {
      a.m_lpos[a.mbits] = st.o_lpos[j];
      a.m_opos[a.mbits++] = (uint8_t)j;
    } else {
      a.n_rpos[a.nbits] = st.o_rpos[j];
      a.n_opos[a.nbits++] = (uint8_t)j;
    }
  }
  a.mtiles = ((1ll << a.mbits) + TILE - 1) / TILE;
  a.ntiles = ((1ll << a.nbits) + TILE - 1) / TILE;
  const int64_t blocks = a.mtiles * a.ntiles * (1ll << a.bbits);
  if (blocks > (1ll << 24)) return TS_EUNSUPPORTED;  // grid.x * 256 threads < 2^32
  hipLaunchKernelGGL(k_step_mfma, dim3((unsigned)blocks), dim3(256), 0, s, lhs, rhs, out, a);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // namespace tsd

extern "C" int ts_dense_config(int32_t enable, int32_t min_sum_bits) {
  if (min_sum_bits < 6 || min_sum_bits > TS_MAX_BITS) return TS_EINVAL;
  tsd::g_enabled = enable != 0;
  tsd::g_min_kbits = min_sum_bits;
  return TS_OK;
}
