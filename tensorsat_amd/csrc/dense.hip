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
// Each workgroup computes a (32 TM) x (32 TN) tile of (m, n) for one batch index with 4
// waves of TM x TN v_mfma_i32_16x16x64_i8 (TM, TN in {2, 4, 8}: 64 x 64 up to 256 x 128
// and 128 x 256; larger wave tiles read fewer LDS bytes per MFMA); K is staged
// through three LDS buffers in 64-byte chunks loaded straight into LDS
// (global_load_lds_dwordx4, four lanes per 64-byte row, XOR-swizzled segments), each
// chunk's loads issued two stages before its MFMAs.  Row/column offsets are scattered by
// bit position (any interleaving of m, n and b bits in the operands and the output), so
// they are computed once per tile into LDS.  The epilogue is the step's: count > 0 (GT0) or
// (count mod 256) != 0 (MOD256); counts are exact because sum_rank <= 27 < 31.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <map>
#include <mutex>

#include "../../include/tensorsat_amd.h"
#include "dense.h"

namespace tsd {
namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int KC = 64;    // bytes of K per LDS stage (one MFMA K step)
constexpr int LDSW = KC;  // LDS row (bytes): unpadded, 16-byte segments XOR-swizzled

// LDS segment swizzle: K segment s (16 bytes) of tile row r sits at segment s ^ seg_swz(r),
// which depends on r mod 16 only.
// The MFMA operand reads take 16 consecutive rows at one K segment; with this swizzle any
// 8 consecutive rows of equal parity (32 banks x 4 bytes per pass) and any 16 consecutive
// rows of equal r mod 4 (64 banks) land on distinct 16-byte bank groups, and the loader's
// writes (4 lanes per row, 16 consecutive rows per wave) cover whole 64-byte rows, so
// neither side has bank conflicts.
// One 16-byte direct load per lane from global memory into LDS (global_load_lds_dwordx4):
// the wave's 64 x 16 bytes land contiguously from the wave-uniform LDS address `l`.
__device__ __forceinline__ void load_lds16(const uint8_t* g, uint8_t* l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

// Workgroup barrier between K stages without the memory fence of __syncthreads, which
// would wait for every direct load in flight (including the next chunk's).  What the
// stage needs is explicit: each wave has waited (s_waitcnt) for the loads of the chunk
// read next and has consumed its own LDS reads (the MFMAs use them) before it arrives;
// the empty asm keeps the compiler from moving LDS accesses across the barrier.
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

__device__ __forceinline__ int seg_swz(int r) { return ((r >> 1) ^ ((r & 15) >> 3)) & 3; }

struct DenseArgs {
  int32_t mbits, nbits, bbits, kbits;
  // Bit positions (LSB-first) of the m / n / b index bits in lhs, rhs and out.
  uint8_t m_lpos[TS_MAX_BITS], m_opos[TS_MAX_BITS];
  uint8_t n_rpos[TS_MAX_BITS], n_opos[TS_MAX_BITS];
  uint8_t b_lpos[TS_MAX_BITS], b_rpos[TS_MAX_BITS], b_opos[TS_MAX_BITS];
  int32_t mode;
  int64_t mtiles, ntiles;  // tiles along m and n
};

__device__ __forceinline__ uint32_t scatter(uint64_t idx, int nb, const uint8_t* pos) {
  uint32_t off = 0;
  for (int i = 0; i < nb; ++i) off |= (uint32_t)((idx >> i) & 1ull) << pos[i];
  return off;
}

// Workgroup tile BM x BN = (32 TM) x (32 TN): 2 x 2 waves, each TM x TN MFMA tiles of
// 16 x 16.  K advances 64 bytes per stage through three LDS buffers: stage c's MFMAs run
// while chunks c + 1 and c + 2 land, and one partial wait plus barrier per stage hands
// chunk c + 1 to every wave.
// At least two waves per SIMD: the 256-entry tiles hold 128 accumulator registers per lane
// and would otherwise be register-allocated for one.
template <int TM, int TN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_step_mfma(const uint8_t* __restrict__ lhs, const uint8_t* __restrict__ rhs,
                                                   uint8_t* __restrict__ out, DenseArgs a) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  constexpr int LA = BM * KC / 16 / 256, LB = BN * KC / 16 / 256;  // 16-byte loads per thread
  // Three K buffers per operand, as separate LDS objects: a direct load into one buffer
  // then does not hold up the reads of another (the compiler tracks them apart).
  __shared__ __attribute__((aligned(16))) uint8_t sA0[BM][LDSW], sA1[BM][LDSW], sA2[BM][LDSW];
  __shared__ __attribute__((aligned(16))) uint8_t sB0[BN][LDSW], sB1[BN][LDSW], sB2[BN][LDSW];
  __shared__ uint32_t rowL[BM], rowO[BM], colR[BN], colO[BN];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Tile order.  Consecutive workgroups land on different XCDs (w % 8), so first give
  // each XCD a contiguous run of tile ids; within a batch index, ids then walk groups of
  // 8 tile rows column by column, so the workgroups resident on one XCD at a time cover
  // a compact block of rows and columns and share their K chunks through that XCD's L2
  // (row-major order would stream the whole rhs once per tile row).
  const uint64_t nwg = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  const uint64_t tiles = (uint64_t)a.mtiles * (uint64_t)a.ntiles;
  const uint64_t b = w / tiles;
  const uint64_t t = w % tiles;
  constexpr uint64_t G = 8;
  const uint64_t gspan = G * (uint64_t)a.ntiles;
  const uint64_t first_m = (t / gspan) * G;
  const uint64_t gsize = ((uint64_t)a.mtiles - first_m) < G ? ((uint64_t)a.mtiles - first_m) : G;
  const uint64_t mt = first_m + (t % gspan) % gsize;
  const uint64_t nt = (t % gspan) / gsize;
  const uint64_t m0 = mt * BM, n0 = nt * BN;
  const uint64_t mcount = 1ull << a.mbits, ncount = 1ull << a.nbits;
  for (int i = tid; i < BM; i += 256) {
    const uint64_t m = m0 + i;
    rowL[i] = m < mcount ? scatter(m, a.mbits, a.m_lpos) : 0u;
    rowO[i] = m < mcount ? scatter(m, a.mbits, a.m_opos) : 0u;
  }
  for (int i = tid; i < BN; i += 256) {
    const uint64_t n = n0 + i;
    colR[i] = n < ncount ? scatter(n, a.nbits, a.n_rpos) : 0u;
    colO[i] = n < ncount ? scatter(n, a.nbits, a.n_opos) : 0u;
  }
  const uint32_t bL = scatter(b, a.bbits, a.b_lpos), bR = scatter(b, a.bbits, a.b_rpos);
  const uint32_t bO = scatter(b, a.bbits, a.b_opos);
  __syncthreads();
  // Loader: direct global -> LDS loads (global_load_lds_dwordx4).  Load j of a wave fills
  // 1 KiB of LDS, rows 64 j + 16 wave .. + 15: lane l lands at row + l / 4, physical
  // segment l % 4, so it fetches the logical K segment the swizzle stores there (rows 64 j
  // apart share the swizzle).  Rows past the operand's end read its first row instead:
  // they only feed outputs the epilogue discards.
  const int lsw = ((tid & 3) ^ seg_swz(tid >> 2)) * 16;
  const uint8_t* pa[LA];
  const uint8_t* pb[LB];
  for (int j = 0; j < LA; ++j) pa[j] = lhs + bL + rowL[(tid >> 2) + 64 * j] + lsw;
  for (int j = 0; j < LB; ++j) pb[j] = rhs + bR + colR[(tid >> 2) + 64 * j] + lsw;
  using Q0 = std::integral_constant<int, 0>;
  using Q1 = std::integral_constant<int, 1>;
  using Q2 = std::integral_constant<int, 2>;
  auto bufA = [&](auto Q) -> uint8_t(*)[LDSW] {
    if constexpr (decltype(Q)::value == 0) return sA0;
    else if constexpr (decltype(Q)::value == 1) return sA1;
    else return sA2;
  };
  auto bufB = [&](auto Q) -> uint8_t(*)[LDSW] {
    if constexpr (decltype(Q)::value == 0) return sB0;
    else if constexpr (decltype(Q)::value == 1) return sB1;
    else return sB2;
  };
  auto dload = [&](auto Q, uint64_t k0) {
    uint8_t(*A)[LDSW] = bufA(Q);
    uint8_t(*B)[LDSW] = bufB(Q);
    for (int j = 0; j < LA; ++j) load_lds16(pa[j] + k0, &A[64 * j + 16 * wave][0]);
    for (int j = 0; j < LB; ++j) load_lds16(pb[j] + k0, &B[64 * j + 16 * wave][0]);
  };
  // vmcnt(LA + LB), nothing else: every direct load but the LA + LB issued last has landed.
  constexpr int kWaitAllButOneChunk = (LA + LB) | 0xF70;
  static_assert(LA + LB <= 15, "vmcnt field");
  const int wm = (wave >> 1) * 16 * TM, wn = (wave & 1) * 16 * TN;
  i4 acc[TM][TN];
  for (int i = 0; i < TM; ++i)
    for (int j = 0; j < TN; ++j) acc[i][j] = (i4){0, 0, 0, 0};
  const uint64_t nstages = (1ull << a.kbits) / KC;
  dload(Q0{}, 0);
  if (nstages > 1) {
    dload(Q1{}, KC);
    __builtin_amdgcn_s_waitcnt(kWaitAllButOneChunk);  // chunk 0 landed, chunk 1 in flight
  } else {
    __builtin_amdgcn_s_waitcnt(0);
  }
  stage_barrier();
  // Lane l feeds row / column l % 16 with 16 bytes of K at 16 * (l / 16); A and B use the
  // same K assignment, so the sum over K does not depend on the hardware's order.  Operand
  // rows are 16-aligned plus l % 16, so the swizzled segment is the same for every tile.
  // Stage c (buffer q = c % 3) reads its fragments, then issues chunk c + 2 into buffer
  // (c + 2) % 3 (last read in stage c - 1, before its barrier), so every chunk has two
  // stages of MFMAs to land; the stage ends once chunk c + 1 has landed (a wait that leaves
  // chunk c + 2 in flight) and every wave is past its reads.  The loads are issued after the
  // fragment reads: LDS reads behind an in-flight direct load into the same buffer wait.
  const int rr = lane & 15, kk = ((lane >> 4) ^ seg_swz(rr)) * 16;
  // Steady: the stage is known to load ahead (no branch, so the compiler's own wait
  // tracking stays exact across the loop); otherwise it checks.
  auto stage = [&](auto Q, uint64_t c, auto Steady) {
    constexpr int q = decltype(Q)::value;
    uint8_t(*A)[LDSW] = bufA(Q);
    uint8_t(*B)[LDSW] = bufB(Q);
    i4 fa[TM], fb[TN];
    for (int i = 0; i < TM; ++i) fa[i] = *reinterpret_cast<const i4*>(&A[wm + 16 * i + rr][kk]);
    for (int j = 0; j < TN; ++j) fb[j] = *reinterpret_cast<const i4*>(&B[wn + 16 * j + rr][kk]);
    const bool ahead = decltype(Steady)::value || c + 2 < nstages;
    if (ahead) dload(std::integral_constant<int, (q + 2) % 3>{}, (c + 2) * KC);
    for (int i = 0; i < TM; ++i)
      for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    if (ahead) __builtin_amdgcn_s_waitcnt(kWaitAllButOneChunk);
    else __builtin_amdgcn_s_waitcnt(0);
    stage_barrier();
  };
  // Steady state without branches (every stage loads ahead), six stages per trip (the
  // compiler drains the loads in flight once per trip, at the loop head), then at most
  // seven tail stages.
  uint64_t c = 0;
  for (; c + 7 < nstages; c += 6) {
    stage(Q0{}, c, std::true_type{});
    stage(Q1{}, c + 1, std::true_type{});
    stage(Q2{}, c + 2, std::true_type{});
    stage(Q0{}, c + 3, std::true_type{});
    stage(Q1{}, c + 4, std::true_type{});
    stage(Q2{}, c + 5, std::true_type{});
  }
  if (c < nstages) stage(Q0{}, c, std::false_type{});
  if (c + 1 < nstages) stage(Q1{}, c + 1, std::false_type{});
  if (c + 2 < nstages) stage(Q2{}, c + 2, std::false_type{});
  if (c + 3 < nstages) stage(Q0{}, c + 3, std::false_type{});
  if (c + 4 < nstages) stage(Q1{}, c + 4, std::false_type{});
  if (c + 5 < nstages) stage(Q2{}, c + 5, std::false_type{});
  if (c + 6 < nstages) stage(Q0{}, c + 6, std::false_type{});
  // Accumulator layout: lane l holds rows 4 * (l / 16) + t, column l % 16.
  for (int i = 0; i < TM; ++i)
    for (int j = 0; j < TN; ++j)
      for (int t = 0; t < 4; ++t) {
        const int r = wm + 16 * i + 4 * (lane >> 4) + t;
        const int col = wn + 16 * j + (lane & 15);
        if (m0 + r >= mcount || n0 + col >= ncount) continue;
        const int v = acc[i][j][t];
        const uint8_t bit = a.mode == TS_MODE_MOD256 ? (uint8_t)((v & 255) != 0) : (uint8_t)(v > 0);
        out[(uint64_t)bO + rowO[r] + colO[col]] = bit;
      }
}

// Short-K steps (0..5 summed bits, K = 1..32 bytes per row; opt-in, ts_dense_config):
// one K stage, K zero-padded to 64 in LDS.  Such products are output-bound: one
// 16 x 16 x 64 int8 MFMA gives 256 outputs every 16 cycles of a SIMD, about five times
// the HBM write rate of the whole device, so the padding costs nothing that shows.  What
// does show is the output stores: a tile writes runs of 64 bytes per row, where the step
// engine's consecutive lanes write whole lines, and on the measured workloads (the
// reference-style trees of n=40 instances, the n=120 cube programs) the step engine was
// as fast or faster -- hence off by default.  Row r's K
// bytes are read by thread r (rows are K-aligned: the K bits are the operand's lowest),
// written to LDS at the swizzled segment position with zeros after them.
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_step_mfma_short(const uint8_t* __restrict__ lhs, const uint8_t* __restrict__ rhs,
                                                         uint8_t* __restrict__ out, DenseArgs a) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  __shared__ __attribute__((aligned(16))) uint8_t sA[BM][LDSW];
  __shared__ __attribute__((aligned(16))) uint8_t sB[BN][LDSW];
  __shared__ uint32_t rowO[BM], colO[BN];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Tile order as in k_step_mfma: a contiguous run of tile ids per XCD, groups of 8 tile
  // rows walked column by column within a batch index.
  const uint64_t nwg = gridDim.x;
  uint64_t w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  const uint64_t tiles = (uint64_t)a.mtiles * (uint64_t)a.ntiles;
  const uint64_t b = w / tiles;
  const uint64_t t = w % tiles;
  constexpr uint64_t G = 8;
  const uint64_t gspan = G * (uint64_t)a.ntiles;
  const uint64_t first_m = (t / gspan) * G;
  const uint64_t gsize = ((uint64_t)a.mtiles - first_m) < G ? ((uint64_t)a.mtiles - first_m) : G;
  const uint64_t mt = first_m + (t % gspan) % gsize;
  const uint64_t nt = (t % gspan) / gsize;
  const uint64_t m0 = mt * BM, n0 = nt * BN;
  const uint64_t mcount = 1ull << a.mbits, ncount = 1ull << a.nbits;
  const uint32_t bL = scatter(b, a.bbits, a.b_lpos), bR = scatter(b, a.bbits, a.b_rpos);
  const uint32_t bO = scatter(b, a.bbits, a.b_opos);
  const int K = 1 << a.kbits;
  // One row (of A, then of B) per thread and pass: its K bytes into registers (rows past
  // the operand's end read its first row; their outputs are discarded), then the whole
  // 64-byte LDS row, zeros after K.
  for (int i = tid; i < BM + BN; i += 256) {
    const bool isA = i < BM;
    const int r = isA ? i : i - BM;
    const uint64_t idx = (isA ? m0 : n0) + r;
    const bool live = idx < (isA ? mcount : ncount);
    uint32_t off = 0;
    if (live) off = isA ? scatter(idx, a.mbits, a.m_lpos) : scatter(idx, a.nbits, a.n_rpos);
    const uint8_t* src = (isA ? lhs + bL : rhs + bR) + off;
    if (isA) rowO[r] = live ? scatter(idx, a.mbits, a.m_opos) : 0u;
    else colO[r] = live ? scatter(idx, a.nbits, a.n_opos) : 0u;
    i4 seg0 = (i4){0, 0, 0, 0}, seg1 = (i4){0, 0, 0, 0};
    if (K >= 16) {
      seg0 = *reinterpret_cast<const i4*>(src);
      if (K == 32) seg1 = *reinterpret_cast<const i4*>(src + 16);
    } else if (K == 8) {
      const uint2 v = *reinterpret_cast<const uint2*>(src);
      seg0 = (i4){(int)v.x, (int)v.y, 0, 0};
    } else if (K == 4) {
      seg0 = (i4){(int)*reinterpret_cast<const uint32_t*>(src), 0, 0, 0};
    } else if (K == 2) {
      seg0 = (i4){(int)*reinterpret_cast<const uint16_t*>(src), 0, 0, 0};
    } else {
      seg0 = (i4){(int)*src, 0, 0, 0};
    }
    uint8_t* row = isA ? &sA[r][0] : &sB[r][0];
    const int sw = seg_swz(r);
    *reinterpret_cast<i4*>(row + 16 * (0 ^ sw)) = seg0;
    *reinterpret_cast<i4*>(row + 16 * (1 ^ sw)) = seg1;
    *reinterpret_cast<i4*>(row + 16 * (2 ^ sw)) = (i4){0, 0, 0, 0};
    *reinterpret_cast<i4*>(row + 16 * (3 ^ sw)) = (i4){0, 0, 0, 0};
  }
  __syncthreads();
  const int wm = (wave >> 1) * 16 * TM, wn = (wave & 1) * 16 * TN;
  const int rr = lane & 15, kk = ((lane >> 4) ^ seg_swz(rr)) * 16;
  i4 fa[TM], fb[TN];
  for (int i = 0; i < TM; ++i) fa[i] = *reinterpret_cast<const i4*>(&sA[wm + 16 * i + rr][kk]);
  for (int j = 0; j < TN; ++j) fb[j] = *reinterpret_cast<const i4*>(&sB[wn + 16 * j + rr][kk]);
  for (int i = 0; i < TM; ++i)
    for (int j = 0; j < TN; ++j) {
      const i4 acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], (i4){0, 0, 0, 0}, 0, 0, 0);
      // Accumulator layout: lane l holds rows 4 * (l / 16) + q, column l % 16.
      for (int q = 0; q < 4; ++q) {
        const int r = wm + 16 * i + 4 * (lane >> 4) + q;
        const int col = wn + 16 * j + (lane & 15);
        if (m0 + r >= mcount || n0 + col >= ncount) continue;
        const int v = acc[q];
        const uint8_t bit = a.mode == TS_MODE_MOD256 ? (uint8_t)((v & 255) != 0) : (uint8_t)(v > 0);
        out[(uint64_t)bO + rowO[r] + colO[col]] = bit;
      }
    }
}

bool g_enabled = true;
int g_min_kbits = 6;
int g_short_min = 6;  // steps of g_short_min..5 summed bits take k_step_mfma_short (6: none)

// Tile per step shape (mbits, nbits, bbits, kbits): 0 the built-in rule, 1 64x64,
// 2 128x64, 3 64x128, 4 128x128, 5 256x128, 6 128x256 (rows x columns); set by the tile
// autotuner.
std::mutex g_tiles_mu;
std::map<uint32_t, int> g_tiles;

uint32_t shape_key(int m, int n, int b, int k) {
  return (uint32_t)m | ((uint32_t)n << 5) | ((uint32_t)b << 10) | ((uint32_t)k << 15);
}

int tile_rule(int mbits, int nbits) {
  if (mbits >= 10 && nbits >= 10) return 6;
  return mbits >= 7 && nbits >= 7 ? 4 : 1;
}

int tile_for(int m, int n, int b, int k) {
  std::lock_guard<std::mutex> g(g_tiles_mu);
  const auto it = g_tiles.find(shape_key(m, n, b, k));
  return it != g_tiles.end() && it->second != 0 ? it->second : tile_rule(m, n);
}

}  // namespace

bool dense_eligible(const ts_step& st, const void* lhs, const void* rhs) {
  if (!g_enabled) return false;
  if ((st.flags & TS_STEP_NORHS) || !(st.flags & TS_STEP_KINNER)) return false;
  if (st.sum_rank < (st.sum_rank < 6 ? g_short_min : g_min_kbits)) return false;
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
  if (a.kbits < 6) {
    // one K stage: tile by the built-in rule
    const int rule = tile_rule(a.mbits, a.nbits);
    const int64_t BM = rule == 6 ? 128 : rule == 4 ? 128 : 64, BN = rule == 6 ? 256 : rule == 4 ? 128 : 64;
    a.mtiles = ((1ll << a.mbits) + BM - 1) / BM;
    a.ntiles = ((1ll << a.nbits) + BN - 1) / BN;
    const int64_t blocks = a.mtiles * a.ntiles * (1ll << a.bbits);
    if (blocks > (1ll << 24)) return TS_EUNSUPPORTED;
    const dim3 grid((unsigned)blocks), block(256);
    if (rule == 6)
      hipLaunchKernelGGL((k_step_mfma_short<4, 8>), grid, block, 0, s, lhs, rhs, out, a);
    else if (rule == 4)
      hipLaunchKernelGGL((k_step_mfma_short<4, 4>), grid, block, 0, s, lhs, rhs, out, a);
    else
      hipLaunchKernelGGL((k_step_mfma_short<2, 2>), grid, block, 0, s, lhs, rhs, out, a);
    return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
  }
  // Tile from the autotuned table, else the built-in rule (tile_rule)
  // entries and 64 x 64 otherwise.
  const int tile = tile_for(a.mbits, a.nbits, a.bbits, a.kbits);
  static constexpr int kTM[7] = {2, 2, 4, 2, 4, 8, 4}, kTN[7] = {2, 2, 2, 4, 4, 4, 8};
  const int64_t BM = 32 * kTM[tile], BN = 32 * kTN[tile];
  a.mtiles = ((1ll << a.mbits) + BM - 1) / BM;
  a.ntiles = ((1ll << a.nbits) + BN - 1) / BN;
  const int64_t blocks = a.mtiles * a.ntiles * (1ll << a.bbits);
  if (blocks > (1ll << 24)) return TS_EUNSUPPORTED;  // grid.x * 256 threads < 2^32
  const dim3 grid((unsigned)blocks), block(256);
  switch (tile) {
    case 6: hipLaunchKernelGGL((k_step_mfma<4, 8>), grid, block, 0, s, lhs, rhs, out, a); break;
    case 5: hipLaunchKernelGGL((k_step_mfma<8, 4>), grid, block, 0, s, lhs, rhs, out, a); break;
    case 4: hipLaunchKernelGGL((k_step_mfma<4, 4>), grid, block, 0, s, lhs, rhs, out, a); break;
    case 3: hipLaunchKernelGGL((k_step_mfma<2, 4>), grid, block, 0, s, lhs, rhs, out, a); break;
    case 2: hipLaunchKernelGGL((k_step_mfma<4, 2>), grid, block, 0, s, lhs, rhs, out, a); break;
    default: hipLaunchKernelGGL((k_step_mfma<2, 2>), grid, block, 0, s, lhs, rhs, out, a); break;
  }
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // namespace tsd

extern "C" int ts_dense_tiles(int32_t mbits, int32_t nbits, int32_t bbits, int32_t kbits, int32_t tile) {
  if (mbits < 0 || nbits < 0 || bbits < 0 || kbits < 0 || mbits > TS_MAX_BITS || nbits > TS_MAX_BITS ||
      bbits > TS_MAX_BITS || kbits > TS_MAX_BITS || tile < 0 || tile > 6)
    return TS_EINVAL;
  std::lock_guard<std::mutex> g(tsd::g_tiles_mu);
  if (tile == 0)
    tsd::g_tiles.erase(tsd::shape_key(mbits, nbits, bbits, kbits));
  else
    tsd::g_tiles[tsd::shape_key(mbits, nbits, bbits, kbits)] = tile;
  return TS_OK;
}

extern "C" int ts_dense_tile_of(int32_t mbits, int32_t nbits, int32_t bbits, int32_t kbits, int32_t* tile) {
  if (!tile || mbits < 0 || nbits < 0 || bbits < 0 || kbits < 0 || mbits > TS_MAX_BITS || nbits > TS_MAX_BITS ||
      bbits > TS_MAX_BITS || kbits > TS_MAX_BITS)
    return TS_EINVAL;
  *tile = tsd::tile_for(mbits, nbits, bbits, kbits);
  return TS_OK;
}

extern "C" int ts_dense_config(int32_t enable, int32_t min_sum_bits) {
  if (min_sum_bits < 0 || min_sum_bits > TS_MAX_BITS) return TS_EINVAL;
  tsd::g_enabled = enable != 0;
  tsd::g_short_min = min_sum_bits < 6 ? min_sum_bits : 6;
  tsd::g_min_kbits = min_sum_bits < 6 ? 6 : min_sum_bits;
  return TS_OK;
}
