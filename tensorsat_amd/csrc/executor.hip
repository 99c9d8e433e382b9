// Grid-wide kernels and C-ABI entry points of libtensorsat_amd.so (gfx950):
// Box algebra (contract2, rewire, spider, union, stack, select, diagonal, as_scalar),
// the counting semiring, and single-instance program execution.
//
// Reference anchors: tensorsat/lang/fin_rel.py:255-393,472 (FinRel box semantics),
// tensorsat/_utils/misc.py:104-123 (rewire_array), tensorsat/_autoray.py:31-80 (the
// einsum/transpose/stack seam), tensorsat/contractions/simple.py:251-292 (step loop).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "dense.h"
#include "device.h"

using namespace tsd;

namespace {

inline hipStream_t S(ts_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline int hip_rc(hipError_t e) { return e == hipSuccess ? TS_OK : TS_EHIP; }

int grid_for(uint64_t n) {
  uint64_t b = (n + WG - 1) / WG;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}

// ---- kernels -------------------------------------------------------------------------

__global__ __launch_bounds__(WG) void k_step_u8(const uint8_t* __restrict__ lhs,
                                                const uint8_t* __restrict__ rhs,
                                                uint8_t* __restrict__ out, ts_step st, int mode) {
  __shared__ Lut L;
  const bool need_c = !(st.flags & TS_STEP_KINNER) && st.sum_rank > 0;
  build_lut(st, L, need_c, threadIdx.x);
  __syncthreads();
  uint32_t ml, mr;
  const int t = vec_plan(st, ml, mr);
  if (t > 0)
    step_u8_vec(st, L, lhs, rhs, out, (uint64_t)blockIdx.x * WG + threadIdx.x, (uint64_t)gridDim.x * WG, t, ml, mr);
  else
    step_u8(st, L, lhs, rhs, out, (uint64_t)blockIdx.x * WG + threadIdx.x, (uint64_t)gridDim.x * WG, mode);
}

// Steps with few outputs and many summed terms (a scalar closing a contraction tree, a
// reduction to a few bits): one workgroup per output, its 2^sum_rank terms split over
// the threads (16-byte runs when the summed bits are innermost, chunk tables otherwise:
// L.cl / L.cr built by the caller), the partial counts added across the workgroup.
// Counts are exact in uint32 (at most 2^27 terms) before the epilogue.
__device__ void dot_block(const ts_step& st, const Lut& L, const uint8_t* __restrict__ lhs,
                          const uint8_t* __restrict__ rhs, uint8_t* __restrict__ out, uint32_t o, int mode,
                          uint32_t* part) {
  const int tid = threadIdx.x;
  const bool kinner = st.flags & TS_STEP_KINNER, norhs = st.flags & TS_STEP_NORHS;
  uint32_t offl = 0, offr = 0;
  for (int j = 0; j < st.out_rank; ++j)
    if ((o >> j) & 1u) {
      if (st.o_lpos[j] != TS_NOPOS) offl |= 1u << st.o_lpos[j];
      if (st.o_rpos[j] != TS_NOPOS) offr |= 1u << st.o_rpos[j];
    }
  const uint32_t nterm = 1u << st.sum_rank;
  uint32_t acc = 0;
  if (kinner) {
    const uint4* a = reinterpret_cast<const uint4*>(lhs + offl);
    const uint4* b = reinterpret_cast<const uint4*>(rhs + offr);
    for (uint32_t i = tid; i < (nterm >> 4); i += WG) {
      uint4 x = a[i];
      if (!norhs) {
        const uint4 y = b[i];
        x.x &= y.x;
        x.y &= y.y;
        x.z &= y.z;
        x.w &= y.w;
      }
      acc += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
    }
  } else {
    const int nqc = nchunks(st.sum_rank);
    for (uint32_t c = tid; c < nterm; c += WG) {
      const uint8_t x = lhs[offl + c_off<4>(L.cl, c, nqc)];
      const uint8_t y = norhs ? (uint8_t)1 : rhs[offr + c_off<4>(L.cr, c, nqc)];
      acc += x & y;
    }
  }
  for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
  if ((tid & 63) == 0) part[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
    for (int w = 0; w < WG / 64; ++w) t += part[w];
    out[o] = (uint8_t)(mode == TS_MODE_MOD256 ? ((t & 255u) != 0) : (t != 0));
  }
}

// Large steps whose rows (2^4..2^10 bytes, summed bits innermost in both operands) are
// read again and again (at least two out bits held by one operand only): packed into bit
// masks first by gang execution when the job's arena stride leaves room for them.  On
// the masks the step has 3 summed bits fewer; rows of at most 8 bytes (at most 64 terms)
// may take the != 0 shortcut, longer ones keep exact popcounts for the MOD256 epilogue.
__host__ __device__ inline bool packable(const ts_step& st) {
  if (!(st.flags & TS_STEP_KINNER) || (st.flags & TS_STEP_NORHS)) return false;
  if (st.sum_rank < 4 || st.sum_rank > 10 || st.out_rank < 12) return false;
  int one_side = 0;
  for (int q = 0; q < st.out_rank; ++q) one_side += (st.o_lpos[q] == TS_NOPOS) != (st.o_rpos[q] == TS_NOPOS);
  return one_side >= 2;
}

// Offset of the packed rhs after the packed lhs (16-byte aligned), and the bytes both take.
__host__ __device__ inline uint32_t packed_rhs_at(const ts_step& st) {
  return ((1u << (st.lhs_rank - 3)) + 15u) & ~15u;
}
__host__ __device__ inline uint64_t packed_bytes(const ts_step& st) {
  return packed_rhs_at(st) + (1ull << (st.rhs_rank - 3));
}

// Rule shared by the host dispatch: at least 2^8 terms per output, at most 2^12 outputs,
// summed bits innermost (16-byte runs) or not.
inline bool dot_shaped(const ts_step& st) { return st.sum_rank >= 8 && st.out_rank <= 12; }

__global__ __launch_bounds__(WG) void k_step_dot(const uint8_t* __restrict__ lhs, const uint8_t* __restrict__ rhs,
                                                 uint8_t* __restrict__ out, ts_step st, int mode) {
  __shared__ Lut L;
  __shared__ uint32_t part[WG / 64];
  if (!(st.flags & TS_STEP_KINNER)) {
    build_lut(st, L, true, threadIdx.x);
    __syncthreads();
  }
  dot_block(st, L, lhs, rhs, out, blockIdx.x, mode, part);
}

// Descriptor structs passed by value are read through the kernel-argument segment
// (uniform scalar loads, no private copy of their arrays).
template <class T>
__device__ __forceinline__ const T& kernarg() {
  return *(const T*)__builtin_amdgcn_kernarg_segment_ptr();  // constant to generic address space
}

// A run of consecutive small steps of one job on one workgroup: each step's record is
// staged in LDS, its chunk tables built, and its outputs written to the job arena before
// the next step reads them (workgroup barriers between steps).  Saves one launch per
// step where steps are too small to fill the device.
// The next record is loaded while the current step runs.
// Step records are double-buffered in LDS: record i + 1 is stored while step i runs, so
// the barrier that ends step i also publishes the next record.  Steps of at most 8 out
// bits whose summed bits are innermost (or absent) compute their offsets directly, one
// output per thread, without chunk tables: one barrier per step.  Larger steps build
// their tables first (one more barrier).
__device__ void seq_run(uint8_t* __restrict__ arena, const ts_step* __restrict__ steps, int n, int mode, Lut& L,
                        ts_step* s_st /* [2] in LDS */) {
  constexpr int RW = (int)(sizeof(ts_step) / 16);
  const int tid = threadIdx.x;
  uint4 r = {0u, 0u, 0u, 0u};
  if (tid < RW && n > 0) reinterpret_cast<uint4*>(&s_st[0])[tid] = reinterpret_cast<const uint4*>(steps)[tid];
  if (tid < RW && n > 1) r = reinterpret_cast<const uint4*>(steps + 1)[tid];
  __syncthreads();
  for (int i = 0; i < n; ++i) {
    if (tid < RW && i + 1 < n) {
      reinterpret_cast<uint4*>(&s_st[(i + 1) & 1])[tid] = r;
      if (i + 2 < n) r = reinterpret_cast<const uint4*>(steps + i + 2)[tid];
    }
    const ts_step& st = s_st[i & 1];
    const bool norhs = st.flags & TS_STEP_NORHS, kinner = st.flags & TS_STEP_KINNER;
    const int ro = st.out_rank, nc = st.sum_rank;
    const uint8_t* lhs = arena + st.lhs_off;
    const uint8_t* rhs = arena + (norhs ? st.lhs_off : st.rhs_off);
    if (ro <= 8 && (kinner || nc == 0)) {
      if (tid < (1 << ro)) {
        uint32_t offl = 0, offr = 0;
        for (int j = 0; j < ro; ++j)
          if ((tid >> j) & 1) {
            if (st.o_lpos[j] != TS_NOPOS) offl |= 1u << st.o_lpos[j];
            if (st.o_rpos[j] != TS_NOPOS) offr |= 1u << st.o_rpos[j];
          }
        uint32_t acc;
        if (nc == 0) {
          acc = lhs[offl];
          if (!norhs) acc &= rhs[offr];
        } else {
          acc = popc_and(lhs + offl, rhs + offr, nc, norhs);
        }
        arena[st.out_off + tid] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
      }
    } else {
      build_lut(st, L, !kinner && nc > 0, tid);
      __syncthreads();
      step_u8(st, L, lhs, rhs, arena + st.out_off, (uint64_t)tid, (uint64_t)WG, mode);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(WG) void k_step_seq(uint8_t* __restrict__ arena, const ts_step* __restrict__ steps,
                                                 int begin, int end, int mode) {
  __shared__ Lut L;
  __shared__ ts_step s_st[2];
  seq_run(arena, steps + begin, end - begin, mode, L, s_st);
}

// ---- gangs: several jobs of one program advanced together --------------------------
//
// Each job has its own arena; a launch serves up to GANG jobs, its descriptors passed by
// value (kernarg above).

constexpr int GANG = 64;

struct GangSeq {  // one workgroup per job: a run of small steps
  const ts_step* steps[GANG];
  uint8_t* arena[GANG];
  int32_t count[GANG];
  int32_t mode;
};

// kind[j]: 0 = step_u8 over the job's blocks (grid-stride), 1 = one block per output
// (few outputs, many terms).
// kind 2: the step's rows (2^4..2^10 bytes, summed bits innermost) were packed into bit
// masks at arena offset pk[j] (lhs) and pk[j] + packed_rhs_at(record) (rhs): the step
// runs on them as on rows of 2..128 bytes, each byte holding 8 terms (popcount of the
// AND is still the exact count).
struct GangStep {
  const ts_step* rec[GANG];
  uint8_t* arena[GANG];
  uint32_t blk[GANG + 1];
  uint32_t pk[GANG];
  uint8_t kind[GANG];
  int32_t n, mode;
};

struct GangPack {  // pack tasks: rows of 2^k bytes -> 2^k-bit masks (k = 4..10)
  const uint8_t* src[2 * GANG];
  uint8_t* dst[2 * GANG];
  uint32_t rows[2 * GANG];
  uint8_t k[2 * GANG];
};

struct GangCopy {
  const uint8_t* src[GANG];
  uint8_t* dst[GANG];
  int64_t bytes[GANG];
};

struct GangAny {
  const uint8_t* res[GANG];
  int64_t bytes[GANG];
  uint8_t* out;
};

__global__ __launch_bounds__(WG) void k_gang_seq(GangSeq) {
  __shared__ Lut L;
  __shared__ ts_step s_st[2];
  const GangSeq& g = kernarg<GangSeq>();
  const int j = blockIdx.x;
  seq_run(g.arena[j], g.steps[j], g.count[j], g.mode, L, s_st);
}

__global__ __launch_bounds__(WG) void k_gang_step(GangStep) {
  __shared__ Lut L;
  __shared__ ts_step s_st;
  __shared__ uint32_t part[WG / 64];
  const GangStep& g = kernarg<GangStep>();
  const uint32_t b = blockIdx.x;
  int j = 0;
  while (j + 1 < g.n && g.blk[j + 1] <= b) ++j;
  const int tid = threadIdx.x;
  if (tid < (int)(sizeof(ts_step) / 16))
    reinterpret_cast<uint4*>(&s_st)[tid] = reinterpret_cast<const uint4*>(g.rec[j])[tid];
  __syncthreads();
  if (g.kind[j] == 2) {  // the same step on the packed rows
    if (tid == 0) {
      const uint32_t pk = g.pk[j];
      s_st.rhs_off = pk + packed_rhs_at(s_st);
      s_st.lhs_off = pk;
      s_st.lhs_rank -= 3;
      s_st.rhs_rank -= 3;
      s_st.sum_rank -= 3;
      for (int q = 0; q < s_st.out_rank; ++q) {
        if (s_st.o_lpos[q] != TS_NOPOS) s_st.o_lpos[q] -= 3;
        if (s_st.o_rpos[q] != TS_NOPOS) s_st.o_rpos[q] -= 3;
      }
    }
    __syncthreads();
  }
  uint8_t* arena = g.arena[j];
  const bool kinner = s_st.flags & TS_STEP_KINNER;
  const uint8_t* rhs = arena + ((s_st.flags & TS_STEP_NORHS) ? s_st.lhs_off : s_st.rhs_off);
  const uint32_t first = b - g.blk[j], nb = g.blk[j + 1] - g.blk[j];
  if (g.kind[j] == 1) {
    if (!kinner) {
      build_lut(s_st, L, true, tid);
      __syncthreads();
    }
    dot_block(s_st, L, arena + s_st.lhs_off, rhs, arena + s_st.out_off, first, g.mode, part);
    return;
  }
  build_lut(s_st, L, !kinner && s_st.sum_rank > 0, tid);
  __syncthreads();
  uint32_t ml, mr;
  const int t = vec_plan(s_st, ml, mr);
  if (t > 0)
    step_u8_vec(s_st, L, arena + s_st.lhs_off, rhs, arena + s_st.out_off, (uint64_t)first * WG + tid,
                (uint64_t)nb * WG, t, ml, mr);
  else
    step_u8(s_st, L, arena + s_st.lhs_off, rhs, arena + s_st.out_off, (uint64_t)first * WG + tid, (uint64_t)nb * WG,
            g.mode);
}

// Nonzero bytes of a 32-bit word as 4 bits.
__device__ __forceinline__ uint32_t nz_bits4(uint32_t w) {
  const uint32_t m = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
  return ((m >> 7) | (m >> 14) | (m >> 21) | (m >> 28)) & 0xFu;
}

// grid (x: row chunks, y: task); one row (2^k bytes, k = 4..10) per thread, packed to a
// 2^k-bit mask (bit i = byte i nonzero) stored little-endian.
__global__ __launch_bounds__(WG) void k_gang_pack(GangPack) {
  const GangPack& g = kernarg<GangPack>();
  const int t = blockIdx.y;
  const uint32_t rows = g.rows[t];
  const int k = g.k[t];
  const uint4* src = reinterpret_cast<const uint4*>(g.src[t]);
  uint8_t* dst = g.dst[t];
  const int nv = 1 << (k - 4);               // 16-byte vectors per row
  const int nw = k >= 6 ? 1 << (k - 6) : 1;  // 64-bit mask words per row
  for (uint32_t r = blockIdx.x * WG + threadIdx.x; r < rows; r += gridDim.x * WG) {
    for (int w = 0; w < nw; ++w) {
      uint64_t bits = 0;
      for (int v = 0; v < (nv < 4 ? nv : 4); ++v) {
        const uint4 x = src[(uint64_t)r * nv + 4 * w + v];
        const uint64_t b16 = nz_bits4(x.x) | (nz_bits4(x.y) << 4) | (nz_bits4(x.z) << 8) | (nz_bits4(x.w) << 12);
        bits |= b16 << (16 * v);
      }
      if (k == 4) reinterpret_cast<uint16_t*>(dst)[r] = (uint16_t)bits;
      else if (k == 5) reinterpret_cast<uint32_t*>(dst)[r] = (uint32_t)bits;
      else reinterpret_cast<uint64_t*>(dst)[(uint64_t)r * nw + w] = bits;
    }
  }
}

// grid (x: 4 KiB chunks, y: job)
__global__ __launch_bounds__(WG) void k_gang_copy(GangCopy) {
  const GangCopy& g = kernarg<GangCopy>();
  const int j = blockIdx.y;
  const int64_t n = g.bytes[j];
  const uint8_t* src = g.src[j];
  uint8_t* dst = g.dst[j];
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) dst[i] = src[i];
}

// One workgroup per job: out[j] = 1 if any byte of the job's result is nonzero.
__global__ __launch_bounds__(WG) void k_gang_any(GangAny) {
  __shared__ int any;
  const GangAny& g = kernarg<GangAny>();
  const int j = blockIdx.x;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const uint8_t* r = g.res[j];
  int v = 0;
  for (int64_t i = threadIdx.x; i < g.bytes[j]; i += WG) v |= r[i];
  if (v) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) g.out[j] = (uint8_t)(any != 0);
}

__global__ __launch_bounds__(WG) void k_step_u64(const uint64_t* __restrict__ lhs,
                                                 const uint64_t* __restrict__ rhs,
                                                 uint64_t* __restrict__ out, ts_step st,
                                                 int32_t* __restrict__ overflow) {
  __shared__ Lut L;
  const bool need_c = !(st.flags & TS_STEP_KINNER) && st.sum_rank > 0;
  build_lut(st, L, need_c, threadIdx.x);
  __syncthreads();
  bool ovf = step_u64(st, L, lhs, rhs, out, (uint64_t)blockIdx.x * WG + threadIdx.x, (uint64_t)gridDim.x * WG);
  if (ovf) atomicOr(overflow, 1);
}

// General rewire (permute + duplicate + discard): out axis j holds input port ports[j].
// Discarded ports are summed (existential OR with the chosen epilogue); output entries
// whose duplicated axes disagree are 0.
struct RewireArgs {
  int32_t in_rank, out_rank, num_disc;
  uint8_t out_port[TS_MAX_BITS];   // input axis for each out axis
  uint8_t disc_port[TS_MAX_BITS];  // discarded input axes
};

__global__ __launch_bounds__(WG) void k_rewire(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                               RewireArgs a, int mode) {
  const uint64_t n = 1ull << a.out_rank;
  for (uint64_t o = (uint64_t)blockIdx.x * WG + threadIdx.x; o < n; o += (uint64_t)gridDim.x * WG) {
    uint32_t base = 0, seen = 0, val = 0;
    bool ok = true;
    for (int j = 0; j < a.out_rank; ++j) {
      const uint32_t bit = (uint32_t)(o >> (a.out_rank - 1 - j)) & 1u;
      const int p = a.out_port[j];
      const uint32_t m = 1u << (a.in_rank - 1 - p);
      if (seen & m) {
        if (((val & m) != 0) != (bit != 0)) ok = false;
      } else {
        seen |= m;
        if (bit) val |= m;
      }
    }
    base = val;
    uint32_t acc = 0;
    if (ok) {
      const uint32_t nt = 1u << a.num_disc;
      for (uint32_t c = 0; c < nt; ++c) {
        uint32_t off = base;
        for (int i = 0; i < a.num_disc; ++i)
          if ((c >> i) & 1) off |= 1u << (a.in_rank - 1 - a.disc_port[i]);
        acc += in[off];
      }
    }
    out[o] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
  }
}

// Gather path of ts_select: out[o] = in[L(o)] over the kept axes, byte lookup tables,
// coalesced stores.
__global__ __launch_bounds__(WG) void k_permute(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                ts_step st) {
  __shared__ Lut L;
  build_lut(st, L, false, threadIdx.x);
  __syncthreads();
  step_u8(st, L, in, nullptr, out, (uint64_t)blockIdx.x * WG + threadIdx.x, (uint64_t)gridDim.x * WG,
          TS_MODE_GT0);
}

// Kernel (a): LDS-staged bit permutation, out[o] = in[pi(o)].  A tile is the set S of out
// bits made of the PB lowest out bits and the out bits that pi sends to the PB lowest in
// bits (|S| <= 2 PB).  A workgroup reads its tile in input order (runs of 2^PB = 128
// consecutive bytes, whole L2 lines), stores it to LDS by tile-local out index, then
// writes it in output order (128-byte runs): both HBM sides are coalesced whatever the
// permutation.
constexpr int PB = 7;  // 2^7 = 128 contiguous bytes: one L2 line per run
struct PermArgs {
  int32_t s, nr;                         // tile bits, remaining (tile-index) bits
  uint8_t load_in[2 * PB], load_lds[2 * PB];  // per tile bit in input order: in pos, LDS bit
  uint8_t store_out[2 * PB];                  // per tile bit in output order: out pos
  uint8_t rest_out[TS_MAX_BITS], rest_in[TS_MAX_BITS];
};

// Bits of v (LSB-first, the first n <= N of them) placed at positions pos[lo + i]; fully
// unrolled so the positions stay in registers.
template <int N>
__device__ __forceinline__ uint32_t spread(uint32_t v, int lo, int n, const uint8_t* pos) {
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (i < n) r |= ((v >> i) & 1u) << pos[lo + i];
  return r;
}

__global__ __launch_bounds__(WG) void k_permute_lds(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                    PermArgs a) {
  __shared__ uint8_t tile[1 << (2 * PB)];
  // Tile element u = tid + 256 k: the low 8 tile bits come from the thread, the rest from
  // k; both parts are fixed for the whole launch.
  const int tid = threadIdx.x;
  const uint32_t ts = 1u << a.s;
  const int nlo = a.s < 8 ? a.s : 8, nhi = a.s - nlo;
  const uint32_t lo_in = spread<8>(tid, 0, nlo, a.load_in), lo_lds = spread<8>(tid, 0, nlo, a.load_lds);
  const uint32_t lo_out = spread<8>(tid, 0, nlo, a.store_out);
  const bool live = (uint32_t)tid < ts;
  const uint32_t nk = ts > WG ? ts / WG : 1;
  const uint64_t ntiles = 1ull << a.nr;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t ib = spread<TS_MAX_BITS>((uint32_t)t, 0, a.nr, a.rest_in);
    const uint32_t ob = spread<TS_MAX_BITS>((uint32_t)t, 0, a.nr, a.rest_out);
    if (live)
      for (uint32_t k = 0; k < nk; ++k)
        tile[lo_lds | spread<2 * PB - 8>(k, 8, nhi, a.load_lds)] = in[ib | lo_in | spread<2 * PB - 8>(k, 8, nhi, a.load_in)];
    __syncthreads();
    if (live)
      for (uint32_t k = 0; k < nk; ++k)
        out[ob | lo_out | spread<2 * PB - 8>(k, 8, nhi, a.store_out)] = tile[(uint32_t)tid + WG * k];
    __syncthreads();
  }
}

// 16-byte variant (tile >= 16 bytes, 16-byte aligned tensors): the 4 lowest tile bits are
// the 4 lowest input bits in load order and the 4 lowest output bits in store order, so
// group g moves bytes 16g..16g+15 of the tile with one uint4 load and one uint4 store;
// only the LDS writes scatter.
__global__ __launch_bounds__(WG) void k_permute_lds16(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                      PermArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t tile[1 << (2 * PB)];
  const uint32_t ngroups = (1u << a.s) >> 4;  // <= 1024
  const int ng = a.s - 4;
  uint32_t lds_j[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) lds_j[j] = spread<4>(j, 0, 4, a.load_lds);
  const uint64_t ntiles = 1ull << a.nr;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t ib = spread<TS_MAX_BITS>((uint32_t)t, 0, a.nr, a.rest_in);
    const uint32_t ob = spread<TS_MAX_BITS>((uint32_t)t, 0, a.nr, a.rest_out);
    for (uint32_t g = threadIdx.x; g < ngroups; g += WG) {
      const uint4 v = *reinterpret_cast<const uint4*>(in + (ib | spread<2 * PB - 4>(g, 4, ng, a.load_in)));
      const uint32_t lg = spread<2 * PB - 4>(g, 4, ng, a.load_lds);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 16; ++j) tile[lg | lds_j[j]] = (uint8_t)(w[j >> 2] >> (8 * (j & 3)));
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < ngroups; g += WG)
      *reinterpret_cast<uint4*>(out + (ob | spread<2 * PB - 4>(g, 4, ng, a.store_out))) =
          *reinterpret_cast<const uint4*>(&tile[16 * g]);
    __syncthreads();
  }
}

// Existential reduction (GT0) over the 7 lowest input bits: lane l reads the two bytes at
// 2l and 2l+1 (in bit 0 inside the pair, bits 1..6 on the lanes), so each load covers 128
// contiguous bytes whichever axes are discarded.  A wave handles one value of the
// remaining kept bits: it ORs the ballot masks of both bytes over all values of the
// remaining discarded bits, and the out entry of each value of the low kept bits is set
// when its lanes' mask bits are.
struct ReduceLowArgs {
  int32_t nkl;                     // kept bits among in bits 1..6
  int32_t nkh, ndh;                // kept / discarded bits above bit 6
  int32_t bit0_out;                // out bit of in bit 0, or -1 when it is discarded
  uint8_t kl_out[6];               // out bit position of each low kept bit
  uint8_t kl_lane[6];              // lane bit of each low kept bit
  uint8_t kh_in[TS_MAX_BITS], kh_out[TS_MAX_BITS];  // high kept bits: in / out positions
  uint8_t dh_in[TS_MAX_BITS];      // high discarded bits: in positions
};

__global__ __launch_bounds__(WG) void k_reduce_low(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                   ReduceLowArgs a) {
  const int lane = threadIdx.x & 63;
  const uint64_t ngroups = 1ull << a.nkh;
  const uint64_t waves = (uint64_t)gridDim.x * (WG / 64);
  const uint32_t nterm = 1u << a.ndh;
  // Lanes of low kept value kappa = lane (for lane < 2^nkl), and kappa's out offset.
  uint64_t my_mask = 0;
  for (int l = 0; l < 64; ++l) {
    uint32_t kl = 0;
    for (int i = 0; i < a.nkl; ++i) kl |= ((uint32_t)(l >> a.kl_lane[i]) & 1u) << i;
    if (kl == (uint32_t)lane) my_mask |= 1ull << l;
  }
  uint32_t kout = 0;
  for (int i = 0; i < a.nkl; ++i) kout |= ((uint32_t)(lane >> i) & 1u) << a.kl_out[i];
  // Discarded high bits: the 3 lowest from an 8-entry table, the rest once per group of 8
  // terms; 8 loads in flight per group.
  const int n3 = a.ndh < 3 ? a.ndh : 3;
  uint32_t low3[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) low3[j] = spread<3>(j, 0, n3, a.dh_in);
  const uint16_t* in2 = reinterpret_cast<const uint16_t*>(in);
  for (uint64_t gidx = (uint64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6); gidx < ngroups; gidx += waves) {
    const uint32_t ibase = spread<TS_MAX_BITS>((uint32_t)gidx, 0, a.nkh, a.kh_in) | ((uint32_t)lane << 1);
    uint64_t acc0 = 0, acc1 = 0;
    if (nterm >= 8) {
      for (uint32_t kg = 0; kg < (nterm >> 3); ++kg) {
        const uint32_t hb = ibase | spread<TS_MAX_BITS - 3>(kg, 3, a.ndh - 3, a.dh_in);
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = in2[(hb | low3[j]) >> 1];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          acc0 |= __ballot((v[j] & 0xffu) != 0);
          acc1 |= __ballot((v[j] >> 8) != 0);
        }
      }
    } else {
      for (uint32_t k = 0; k < nterm; ++k) {
        const uint32_t v = in2[(ibase | low3[k]) >> 1];
        acc0 |= __ballot((v & 0xffu) != 0);
        acc1 |= __ballot((v >> 8) != 0);
      }
    }
    if ((uint32_t)lane < (1u << a.nkl)) {
      const uint32_t obase = spread<TS_MAX_BITS>((uint32_t)gidx, 0, a.nkh, a.kh_out) | kout;
      if (a.bit0_out >= 0) {
        out[obase] = (uint8_t)((acc0 & my_mask) != 0);
        out[obase | (1u << a.bit0_out)] = (uint8_t)((acc1 & my_mask) != 0);
      } else {
        out[obase] = (uint8_t)(((acc0 | acc1) & my_mask) != 0);
      }
    }
  }
}

// Kernel (b), 16-byte lanes: existential reduction (GT0) of a tensor of >= 10 bits.  In
// bits 0..3 are the 16 bytes of a lane's load, bits 4..9 the lanes, higher bits index
// groups (kept) and terms (discarded).  Each load becomes a 16-bit presence mask (byte
// != 0), the discarded in-register bits are folded into the kept positions by
// shift-OR, and every kept in-register combination k collects the lanes holding a set
// byte in a ballot mask; an out entry is then the OR of its lanes' bits (the lanes whose
// kept lane bits spell it).  8 loads of 1 KiB per wave in flight.  A work item is a group
// and one of 2^cb chunks of its terms, so that small groups counts still fill the GPU; the
// output is cleared first and items store only the entries they find set (every writer
// writes 1).
struct ReduceWideArgs {
  int32_t nkl;               // kept lane bits (in bits 4..9)
  int32_t nkh, ndh;          // kept / discarded bits above bit 9
  int32_t ndb;               // discarded in-register bits (in bits 0..3)
  uint8_t kb_in[4], kb_out[4];  // kept in-register bits: in position (0..3), out position
  uint8_t db_in[4];          // discarded in-register bits (0..3)
  uint8_t kl_out[6], kl_lane[6];  // kept lane bits: out position, lane bit
  uint8_t kh_in[TS_MAX_BITS], kh_out[TS_MAX_BITS];
  uint8_t dh_in[TS_MAX_BITS];
  int32_t cb;                // log2 of the term chunks per group
  uint64_t lanes0;           // lanes whose kept lane bits are all 0
  uint32_t fold;             // bit d set: in-register bit d is discarded
};

__global__ __launch_bounds__(WG) void k_zero16(uint4* __restrict__ p, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * WG)
    p[i] = make_uint4(0, 0, 0, 0);
}

// Bit 7 of each byte set iff that byte of x is nonzero (the other bits clear).
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t x) {
  return (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

template <int NKB>
__global__ __launch_bounds__(WG) void k_reduce_w16(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                   ReduceWideArgs a) {
  constexpr int NK = 1 << NKB;
  const int lane = threadIdx.x & 63;
  const uint64_t nitems = 1ull << (a.nkh + a.cb);
  const uint64_t waves = (uint64_t)gridDim.x * (WG / 64);
  const uint32_t nterm = 1u << a.ndh;
  const uint32_t ngk = (nterm + 7) >> 3;              // groups of 8 terms
  const uint32_t per = ngk >> a.cb;                   // per chunk (cb <= log2 ngk)
  // Lanes whose kept lane bits spell this lane's index: the lanes with those bits all 0,
  // shifted by the index's kept bits (disjoint from the discarded ones: no carries).
  uint32_t kin = 0, kout = 0;
  for (int i = 0; i < a.nkl; ++i) {
    kin |= ((uint32_t)(lane >> i) & 1u) << a.kl_lane[i];
    kout |= ((uint32_t)(lane >> i) & 1u) << a.kl_out[i];
  }
  const uint64_t my_mask = a.lanes0 << kin;
  uint32_t kpos[NK], kofs[NK];  // in-register position and out offset of kept combination k
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    kpos[k] = 0;
    kofs[k] = 0;
#pragma unroll
    for (int i = 0; i < NKB; ++i) {
      kpos[k] |= ((uint32_t)(k >> i) & 1u) << a.kb_in[i];
      kofs[k] |= ((uint32_t)(k >> i) & 1u) << a.kb_out[i];
    }
  }
  const int n3 = a.ndh < 3 ? a.ndh : 3;
  uint32_t low3[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) low3[j] = spread<3>(j, 0, n3, a.dh_in);
  for (uint64_t item = (uint64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6); item < nitems; item += waves) {
    const uint32_t gidx = (uint32_t)(item >> a.cb), chunk = (uint32_t)(item & ((1u << a.cb) - 1u));
    const uint32_t ibase = spread<TS_MAX_BITS>(gidx, 0, a.nkh, a.kh_in) | ((uint32_t)lane << 4);
    uint64_t acc[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) acc[k] = 0;
    const uint32_t nj = nterm < 8 ? nterm : 8;
    for (uint32_t kg = chunk * per; kg < (chunk + 1) * per; ++kg) {
      const uint32_t hb = ibase | (a.ndh > 3 ? spread<TS_MAX_BITS - 3>(kg, 3, a.ndh - 3, a.dh_in) : 0u);
      uint4 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if ((uint32_t)j < nj) v[j] = *reinterpret_cast<const uint4*>(in + (hb | low3[j]));
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if ((uint32_t)j >= nj) break;
        // Flags of the 16 bytes (byte i = in bits 0..3 = i), discarded in-register bits
        // folded into the positions where they are 0.
        uint32_t f[4] = {nonzero_bytes(v[j].x), nonzero_bytes(v[j].y), nonzero_bytes(v[j].z), nonzero_bytes(v[j].w)};
        if (a.fold & 1u) {
#pragma unroll
          for (int q = 0; q < 4; ++q) f[q] |= f[q] >> 8;
        }
        if (a.fold & 2u) {
#pragma unroll
          for (int q = 0; q < 4; ++q) f[q] |= f[q] >> 16;
        }
        if (a.fold & 4u) {
          f[0] |= f[1];
          f[2] |= f[3];
        }
        if (a.fold & 8u) {
          f[0] |= f[2];
          f[1] |= f[3];
        }
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const uint32_t q = kpos[k] >> 2, sh = ((kpos[k] & 3u) << 3) + 7u;
          const uint32_t word = q == 0 ? f[0] : q == 1 ? f[1] : q == 2 ? f[2] : f[3];
          acc[k] |= __ballot((word >> sh) & 1u);
        }
      }
    }
    if ((uint32_t)lane < (1u << a.nkl)) {
      const uint32_t obase = spread<TS_MAX_BITS>(gidx, 0, a.nkh, a.kh_out) | kout;
#pragma unroll
      for (int k = 0; k < NK; ++k)
        if (acc[k] & my_mask) out[obase | kofs[k]] = 1;
    }
  }
}

// Kernel (b): reduction over discarded axes with many terms per output: one wave per out
// entry, lanes sweep the 2^nd terms (lowest discarded bits on the lanes), then a
// wave-level OR through ballot (GT0) or a wave sum (MOD256, count mod 256).
struct ReduceArgs {
  int32_t out_rank, nd;
  uint8_t keep_in[TS_MAX_BITS];  // in bit position of out bit q (LSB-first)
  uint8_t disc_in[TS_MAX_BITS];  // in bit positions of the discarded bits, ascending
};

__global__ __launch_bounds__(WG) void k_reduce(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                               ReduceArgs a, int mode) {
  const int lane = threadIdx.x & 63;
  const uint64_t nout = 1ull << a.out_rank;
  const uint64_t waves = (uint64_t)gridDim.x * (WG / 64);
  // term c = lane + 64 k (nd >= 6): the lane part is fixed for the launch.
  const uint32_t lane_off = spread<6>(lane, 0, 6, a.disc_in);
  const uint32_t nk = 1u << (a.nd - 6);
  for (uint64_t o = (uint64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6); o < nout; o += waves) {
    const uint32_t base = spread<TS_MAX_BITS>((uint32_t)o, 0, a.out_rank, a.keep_in) | lane_off;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < nk; ++k) {
      const uint32_t v = in[base | spread<TS_MAX_BITS - 6>(k, 6, a.nd - 6, a.disc_in)];
      acc = mode == TS_MODE_MOD256 ? acc + v : (acc | v);
    }
    uint32_t bit;
    if (mode == TS_MODE_MOD256) {
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
      bit = (acc & 255u) != 0;
    } else {
      bit = __ballot(acc != 0) != 0;
    }
    if (lane == 0) out[o] = (uint8_t)bit;
  }
}

__global__ __launch_bounds__(WG) void k_diagonal(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                 int in_rank, int out_rank, RewireArgs a) {
  // a.out_port[i] = out axis feeding input axis i
  const uint64_t n = 1ull << out_rank;
  for (uint64_t o = (uint64_t)blockIdx.x * WG + threadIdx.x; o < n; o += (uint64_t)gridDim.x * WG) {
    uint32_t off = 0;
    for (int i = 0; i < in_rank; ++i) {
      const int j = a.out_port[i];
      if ((o >> (out_rank - 1 - j)) & 1) off |= 1u << (in_rank - 1 - i);
    }
    out[o] = in[off];
  }
}

__global__ __launch_bounds__(WG) void k_spider(uint8_t* __restrict__ out, int num_ports, int bits, int dim) {
  const int rank = num_ports * bits;
  const uint64_t n = 1ull << rank;
  const uint32_t mask = (1u << bits) - 1u;
  for (uint64_t o = (uint64_t)blockIdx.x * WG + threadIdx.x; o < n; o += (uint64_t)gridDim.x * WG) {
    const uint32_t v0 = (uint32_t)(o >> (rank - bits)) & mask;
    bool eq = v0 < (uint32_t)dim;
    for (int p = 1; p < num_ports && eq; ++p)
      eq = (((uint32_t)(o >> (rank - (p + 1) * bits)) & mask) == v0);
    out[o] = eq ? 1 : 0;
  }
}

__global__ __launch_bounds__(WG) void k_union(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                              uint8_t* __restrict__ out, int64_t n) {
  const int64_t n16 = n / 16;
  const uint4* a4 = reinterpret_cast<const uint4*>(a);
  const uint4* b4 = reinterpret_cast<const uint4*>(b);
  uint4* o4 = reinterpret_cast<uint4*>(out);
  const int64_t tid = (int64_t)blockIdx.x * WG + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * WG;
  for (int64_t i = tid; i < n16; i += stride) {
    uint4 x = a4[i], y = b4[i];
    x.x |= y.x;
    x.y |= y.y;
    x.z |= y.z;
    x.w |= y.w;
    o4[i] = x;
  }
  for (int64_t i = n16 * 16 + tid; i < n; i += stride) out[i] = a[i] | b[i];
}

// Copy `in` (rank r) into out (rank r + bits) at new-axis value k placed at axis `port`.
__global__ __launch_bounds__(WG) void k_embed(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                              int in_rank, int port, int bits, uint32_t k) {
  const uint64_t n = 1ull << in_rank;
  const int low = in_rank - port;  // input axes after the port stay the low bits
  for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
    const uint64_t lo = i & ((1ull << low) - 1);
    const uint64_t hi = i >> low;
    const uint64_t o = (((hi << bits) | k) << low) | lo;
    out[o] = in[i];
  }
}

__global__ __launch_bounds__(WG) void k_fill(uint8_t* __restrict__ out, uint8_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) out[i] = v;
}

__global__ __launch_bounds__(WG) void k_widen(const uint8_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) out[i] = in[i];
}

__global__ void k_scale_add(const uint64_t* __restrict__ val, int shift, uint64_t* __restrict__ acc,
                            int32_t* __restrict__ overflow) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint64_t v = *val;
    if (shift >= 64 ? v != 0 : (shift > 0 && (v >> (64 - shift)) != 0)) {
      atomicOr(overflow, 1);
      return;
    }
    const uint64_t s = shift >= 64 ? 0 : (v << shift);
    const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(acc), (unsigned long long)s);
    if (old + s < old) atomicOr(overflow, 1);
  }
}

// ---- host helpers ----------------------------------------------------------------------

// Build a step record from wire lists; summed wires ordered LSB-first by their lhs
// position, so K-innermost layouts get the vector path.
int make_step(int rl, const int32_t* lw, int rr, const int32_t* rw, int ro, const int32_t* ow,
              bool norhs, ts_step& st) {
  if (rl < 0 || rr < 0 || ro < 0 || rl > TS_MAX_BITS || rr > TS_MAX_BITS || ro > TS_MAX_BITS)
    return TS_EINVAL;
  auto uniq = [](int n, const int32_t* w) {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        if (w[i] == w[j]) return false;
    return true;
  };
  if (!uniq(rl, lw) || (!norhs && !uniq(rr, rw))) return TS_EUNSUPPORTED;
  if (!uniq(ro, ow)) return TS_EUNSUPPORTED;
  auto find = [](int n, const int32_t* w, int32_t x) {
    for (int i = 0; i < n; ++i)
      if (w[i] == x) return n - 1 - i;  // LSB-first bit position
    return -1;
  };
  std::memset(&st, 0, sizeof(st));
  std::memset(st.o_lpos, TS_NOPOS, sizeof(st.o_lpos));
  std::memset(st.o_rpos, TS_NOPOS, sizeof(st.o_rpos));
  std::memset(st.c_lpos, TS_NOPOS, sizeof(st.c_lpos));
  std::memset(st.c_rpos, TS_NOPOS, sizeof(st.c_rpos));
  st.lhs_rank = (uint8_t)rl;
  st.rhs_rank = norhs ? 0 : (uint8_t)rr;
  st.out_rank = (uint8_t)ro;
  for (int j = 0; j < ro; ++j) {
    const int32_t w = ow[ro - 1 - j];
    const int pl = find(rl, lw, w);
    const int pr = norhs ? -1 : find(rr, rw, w);
    if (pl < 0 && pr < 0) return TS_EINVAL;
    st.o_lpos[j] = pl < 0 ? TS_NOPOS : (uint8_t)pl;
    st.o_rpos[j] = pr < 0 ? TS_NOPOS : (uint8_t)pr;
  }
  // Summed wires: lhs wires not in out (LSB-first), then rhs-only wires not in out.
  std::vector<int32_t> sum;
  for (int i = rl - 1; i >= 0; --i)
    if (find(ro, ow, lw[i]) < 0) sum.push_back(lw[i]);
  if (!norhs)
    for (int i = rr - 1; i >= 0; --i)
      if (find(ro, ow, rw[i]) < 0 && find(rl, lw, rw[i]) < 0) sum.push_back(rw[i]);
  if ((int)sum.size() > TS_MAX_BITS) return TS_EUNSUPPORTED;
  st.sum_rank = (uint8_t)sum.size();
  bool kinner = true;
  for (size_t j = 0; j < sum.size(); ++j) {
    const int pl = find(rl, lw, sum[j]);
    const int pr = norhs ? -1 : find(rr, rw, sum[j]);
    st.c_lpos[j] = pl < 0 ? TS_NOPOS : (uint8_t)pl;
    st.c_rpos[j] = pr < 0 ? TS_NOPOS : (uint8_t)pr;
    if (pl != (int)j || (!norhs && pr != (int)j)) kinner = false;
  }
  st.flags = (uint8_t)((kinner ? TS_STEP_KINNER : 0) | (norhs ? TS_STEP_NORHS : 0));
  return TS_OK;
}

int launch_step_u8(const uint8_t* lhs, const uint8_t* rhs, uint8_t* out, const ts_step& st, int mode,
                   hipStream_t s) {
  if (tsd::dense_eligible(st, lhs, rhs)) return tsd::launch_dense(lhs, rhs, out, st, mode, s);
  if (dot_shaped(st)) {
    hipLaunchKernelGGL(k_step_dot, dim3(1u << st.out_rank), dim3(WG), 0, s, lhs, rhs, out, st, mode);
    return hip_rc(hipGetLastError());
  }
  hipLaunchKernelGGL(k_step_u8, dim3(grid_for(1ull << st.out_rank)), dim3(WG), 0, s, lhs, rhs, out, st, mode);
  return hip_rc(hipGetLastError());
}

int launch_step_u64(const uint64_t* lhs, const uint64_t* rhs, uint64_t* out, const ts_step& st,
                    int32_t* ovf, hipStream_t s) {
  hipLaunchKernelGGL(k_step_u64, dim3(grid_for(1ull << st.out_rank)), dim3(WG), 0, s, lhs, rhs, out, st, ovf);
  return hip_rc(hipGetLastError());
}

}  // namespace

extern "C" {

int ts_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return TS_EHIP;
  if (hipSetDevice(device) != hipSuccess) return TS_EHIP;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return TS_EHIP;
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) return TS_EUNSUPPORTED;
  return TS_OK;
}

const char* ts_version(void) { return "tensorsat_amd 0.1.0 (gfx950)"; }

int ts_contract2(const uint8_t* lhs, int32_t lhs_rank, const int32_t* lhs_wires, const uint8_t* rhs,
                 int32_t rhs_rank, const int32_t* rhs_wires, uint8_t* out, int32_t out_rank,
                 const int32_t* out_wires, int32_t mode, ts_stream stream) {
  if (!lhs || !rhs || !out || (mode != TS_MODE_GT0 && mode != TS_MODE_MOD256)) return TS_EINVAL;
  ts_step st;
  int rc = make_step(lhs_rank, lhs_wires, rhs_rank, rhs_wires, out_rank, out_wires, false, st);
  if (rc) return rc;
  if ((reinterpret_cast<uintptr_t>(lhs) | reinterpret_cast<uintptr_t>(rhs)) & 15) st.flags &= ~TS_STEP_KINNER;
  return launch_step_u8(lhs, rhs, out, st, mode, S(stream));
}

int ts_count_contract2(const uint64_t* lhs, int32_t lhs_rank, const int32_t* lhs_wires, const uint64_t* rhs,
                       int32_t rhs_rank, const int32_t* rhs_wires, uint64_t* out, int32_t out_rank,
                       const int32_t* out_wires, int32_t* overflow, ts_stream stream) {
  if (!lhs || !rhs || !out || !overflow) return TS_EINVAL;
  ts_step st;
  int rc = make_step(lhs_rank, lhs_wires, rhs_rank, rhs_wires, out_rank, out_wires, false, st);
  if (rc) return rc;
  return launch_step_u64(lhs, rhs, out, st, overflow, S(stream));
}

int ts_rewire(const uint8_t* in, int32_t in_rank, uint8_t* out, int32_t out_rank, const int32_t* out_ports,
              int32_t mode, ts_stream stream) {
  if (!in || !out || in_rank < 0 || in_rank > TS_MAX_BITS || out_rank < 0 || out_rank > TS_MAX_BITS)
    return TS_EINVAL;
  if (mode != TS_MODE_GT0 && mode != TS_MODE_MOD256) return TS_EINVAL;
  RewireArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in_rank = in_rank;
  a.out_rank = out_rank;
  std::vector<char> used(in_rank, 0);
  bool dup = false;
  for (int j = 0; j < out_rank; ++j) {
    const int p = out_ports[j];
    if (p < 0 || p >= in_rank) return TS_EINVAL;
    if (used[p]) dup = true;
    used[p] = 1;
    a.out_port[j] = (uint8_t)p;
  }
  for (int p = 0; p < in_rank; ++p)
    if (!used[p]) a.disc_port[a.num_disc++] = (uint8_t)p;
  hipStream_t s = S(stream);
  if (!dup) {
    // Permutation / reduction: a unary step (kernel (a) when nothing is discarded,
    // the reduction path of the step engine otherwise).
    std::vector<int32_t> iw(in_rank), ow(out_rank);
    for (int p = 0; p < in_rank; ++p) iw[p] = p;
    for (int j = 0; j < out_rank; ++j) ow[j] = out_ports[j];
    ts_step st;
    int rc = make_step(in_rank, iw.data(), 0, nullptr, out_rank, ow.data(), true, st);
    if (rc) return rc;
    if (a.num_disc == 0) {
      // Bit positions: out bit q is out axis r-1-q; its in bit is r-1-out_port[r-1-q].
      const int r = out_rank;
      std::vector<int> pi(r);
      for (int q = 0; q < r; ++q) pi[q] = r - 1 - out_ports[r - 1 - q];
      const int b = std::min(PB, r);
      std::vector<char> inS(r, 0);
      for (int q = 0; q < r; ++q)
        if (q < b || pi[q] < b) inS[q] = 1;
      std::vector<int> Sout, Sin, rest;
      for (int q = 0; q < r; ++q) (inS[q] ? Sout : rest).push_back(q);
      Sin = Sout;
      std::sort(Sin.begin(), Sin.end(), [&](int x, int y) { return pi[x] < pi[y]; });
      PermArgs pa;
      std::memset(&pa, 0, sizeof(pa));
      pa.s = (int)Sout.size();
      pa.nr = (int)rest.size();
      for (int k = 0; k < pa.s; ++k) {
        pa.load_in[k] = (uint8_t)pi[Sin[k]];
        pa.load_lds[k] = (uint8_t)(std::find(Sout.begin(), Sout.end(), Sin[k]) - Sout.begin());
        pa.store_out[k] = (uint8_t)Sout[k];
      }
      for (int k = 0; k < pa.nr; ++k) {
        pa.rest_out[k] = (uint8_t)rest[k];
        pa.rest_in[k] = (uint8_t)pi[rest[k]];
      }
      const uint64_t ntiles = 1ull << pa.nr;
      const int grid = (int)std::min<uint64_t>(ntiles, 65536);
      const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
      if (pa.s >= 4 && aligned)
        hipLaunchKernelGGL(k_permute_lds16, dim3(grid), dim3(WG), 0, s, in, out, pa);
      else
        hipLaunchKernelGGL(k_permute_lds, dim3(grid), dim3(WG), 0, s, in, out, pa);
      return hip_rc(hipGetLastError());
    }
    if (a.num_disc >= 6 && mode == TS_MODE_GT0 && in_rank >= 10 && (reinterpret_cast<uintptr_t>(in) & 15) == 0) {
      // Bits 0..3 in each lane's 16 bytes, bits 4..9 on the lanes.
      ReduceWideArgs rw;
      std::memset(&rw, 0, sizeof(rw));
      std::vector<int> out_of_in(in_rank, -1);  // in bit -> out bit (kept)
      for (int q = 0; q < out_rank; ++q) out_of_in[in_rank - 1 - out_ports[out_rank - 1 - q]] = q;
      int nkb = 0;
      for (int p = 0; p < in_rank; ++p) {
        const int o = out_of_in[p];
        if (p < 4) {
          if (o >= 0) {
            rw.kb_in[nkb] = (uint8_t)p;
            rw.kb_out[nkb++] = (uint8_t)o;
          } else {
            rw.db_in[rw.ndb++] = (uint8_t)p;
          }
        } else if (p < 10) {
          if (o >= 0) {
            rw.kl_out[rw.nkl] = (uint8_t)o;
            rw.kl_lane[rw.nkl++] = (uint8_t)(p - 4);
          }
        } else if (o >= 0) {
          rw.kh_in[rw.nkh] = (uint8_t)p;
          rw.kh_out[rw.nkh++] = (uint8_t)o;
        } else {
          rw.dh_in[rw.ndh++] = (uint8_t)p;
        }
      }
      // Term chunks: up to 32k work items (waves), a chunk being at least two groups of 8
      // terms (8 loads of 1 KiB in flight each).
      const int ngk_bits = rw.ndh > 3 ? rw.ndh - 3 : 0;
      rw.cb = 0;
      while (rw.cb < ngk_bits - 1 && rw.nkh + rw.cb < 15) ++rw.cb;
      uint32_t kmask = 0;
      for (int i = 0; i < rw.nkl; ++i) kmask |= 1u << rw.kl_lane[i];
      rw.fold = 0;
      for (int i = 0; i < rw.ndb; ++i) rw.fold |= 1u << rw.db_in[i];
      rw.lanes0 = 0;
      for (int l = 0; l < 64; ++l)
        if ((l & kmask) == 0) rw.lanes0 |= 1ull << l;
      const uint64_t items = 1ull << (rw.nkh + rw.cb);
      const int grid = (int)std::min<uint64_t>((items + 3) / 4, 8192);
      const uint64_t nout = 1ull << out_rank;
      if (nout % 16 == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0) {
        hipLaunchKernelGGL(k_zero16, dim3((int)std::min<uint64_t>((nout / 16 + WG - 1) / WG, 1024)), dim3(WG), 0, s,
                           reinterpret_cast<uint4*>(out), nout / 16);
      } else {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(nout)), dim3(WG), 0, s, out, (uint8_t)0, (int64_t)nout);
      }
      switch (nkb) {
        case 0: hipLaunchKernelGGL(k_reduce_w16<0>, dim3(grid), dim3(WG), 0, s, in, out, rw); break;
        case 1: hipLaunchKernelGGL(k_reduce_w16<1>, dim3(grid), dim3(WG), 0, s, in, out, rw); break;
        case 2: hipLaunchKernelGGL(k_reduce_w16<2>, dim3(grid), dim3(WG), 0, s, in, out, rw); break;
        case 3: hipLaunchKernelGGL(k_reduce_w16<3>, dim3(grid), dim3(WG), 0, s, in, out, rw); break;
        default: hipLaunchKernelGGL(k_reduce_w16<4>, dim3(grid), dim3(WG), 0, s, in, out, rw); break;
      }
      return hip_rc(hipGetLastError());
    }
    if (a.num_disc >= 6 && mode == TS_MODE_GT0 && in_rank >= 7 && (reinterpret_cast<uintptr_t>(in) & 1) == 0) {
      // In bit 0 inside each lane's byte pair, in bits 1..6 on the lanes.
      ReduceLowArgs rl;
      std::memset(&rl, 0, sizeof(rl));
      rl.bit0_out = -1;
      std::vector<int> out_of_in(in_rank, -1);  // in bit -> out bit (kept)
      for (int q = 0; q < out_rank; ++q) out_of_in[in_rank - 1 - out_ports[out_rank - 1 - q]] = q;
      for (int p = 0; p < in_rank; ++p) {
        const int o = out_of_in[p];
        if (p == 0) {
          rl.bit0_out = o;
        } else if (p < 7) {
          if (o >= 0) {
            rl.kl_out[rl.nkl] = (uint8_t)o;
            rl.kl_lane[rl.nkl++] = (uint8_t)(p - 1);
          }
        } else if (o >= 0) {
          rl.kh_in[rl.nkh] = (uint8_t)p;
          rl.kh_out[rl.nkh++] = (uint8_t)o;
        } else {
          rl.dh_in[rl.ndh++] = (uint8_t)p;
        }
      }
      const uint64_t groups = 1ull << rl.nkh;
      const int grid = (int)std::min<uint64_t>((groups + 3) / 4, 65536);
      hipLaunchKernelGGL(k_reduce_low, dim3(grid), dim3(WG), 0, s, in, out, rl);
      return hip_rc(hipGetLastError());
    }
    if (a.num_disc >= 6) {
      ReduceArgs ra;
      std::memset(&ra, 0, sizeof(ra));
      ra.out_rank = out_rank;
      ra.nd = a.num_disc;
      for (int q = 0; q < out_rank; ++q) ra.keep_in[q] = (uint8_t)(in_rank - 1 - out_ports[out_rank - 1 - q]);
      std::vector<int> dpos;
      for (int i = 0; i < a.num_disc; ++i) dpos.push_back(in_rank - 1 - a.disc_port[i]);
      std::sort(dpos.begin(), dpos.end());
      for (int i = 0; i < a.num_disc; ++i) ra.disc_in[i] = (uint8_t)dpos[i];
      const uint64_t waves_needed = 1ull << out_rank;
      const int grid = (int)std::min<uint64_t>((waves_needed + 3) / 4, 65536);
      hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(WG), 0, s, in, out, ra, mode);
      return hip_rc(hipGetLastError());
    }
    return launch_step_u8(in, in, out, st, mode, s);  // unary: rhs is never read
  }
  hipLaunchKernelGGL(k_rewire, dim3(grid_for(1ull << out_rank)), dim3(WG), 0, s, in, out, a, mode);
  return hip_rc(hipGetLastError());
}

int ts_diagonal(const uint8_t* in, int32_t in_rank, const int32_t* axis_map, uint8_t* out, int32_t out_rank,
                ts_stream stream) {
  if (!in || !out || in_rank < 0 || in_rank > TS_MAX_BITS || out_rank < 0 || out_rank > TS_MAX_BITS)
    return TS_EINVAL;
  RewireArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int i = 0; i < in_rank; ++i) {
    if (axis_map[i] < 0 || axis_map[i] >= out_rank) return TS_EINVAL;
    a.out_port[i] = (uint8_t)axis_map[i];
  }
  hipLaunchKernelGGL(k_diagonal, dim3(grid_for(1ull << out_rank)), dim3(WG), 0, S(stream), in, out, in_rank,
                     out_rank, a);
  return hip_rc(hipGetLastError());
}

int ts_spider(uint8_t* out, int32_t num_ports, int32_t bits, int32_t dim, ts_stream stream) {
  if (!out || num_ports <= 0 || bits < 0 || dim <= 0 || (int64_t)num_ports * bits > TS_MAX_BITS ||
      (bits < 31 && dim > (1 << bits)))
    return TS_EINVAL;
  hipLaunchKernelGGL(k_spider, dim3(grid_for(1ull << (num_ports * bits))), dim3(WG), 0, S(stream), out, num_ports,
                     bits, dim);
  return hip_rc(hipGetLastError());
}

int ts_union(const uint8_t* a, const uint8_t* b, uint8_t* out, int64_t n, ts_stream stream) {
  if (!a || !b || !out || n < 0) return TS_EINVAL;
  if (n == 0) return TS_OK;
  if ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(out)) & 15)
    return TS_EINVAL;
  hipLaunchKernelGGL(k_union, dim3(grid_for((uint64_t)(n + 15) / 16)), dim3(WG), 0, S(stream), a, b, out, n);
  return hip_rc(hipGetLastError());
}

int ts_stack(const uint8_t* const* ins, int32_t num, int32_t in_rank, int32_t port, int32_t axis_bits,
             uint8_t* out, ts_stream stream) {
  if (!ins || !out || num <= 0 || in_rank < 0 || port < 0 || port > in_rank || axis_bits < 0 ||
      in_rank + axis_bits > TS_MAX_BITS || num > (1 << axis_bits))
    return TS_EINVAL;
  hipStream_t s = S(stream);
  const int64_t total = int64_t(1) << (in_rank + axis_bits);
  if (num < (1 << axis_bits)) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for(total)), dim3(WG), 0, s, out, (uint8_t)0, total);
  }
  for (int k = 0; k < num; ++k) {
    if (!ins[k]) return TS_EINVAL;
    hipLaunchKernelGGL(k_embed, dim3(grid_for(1ull << in_rank)), dim3(WG), 0, s, ins[k], out, in_rank, port,
                       axis_bits, (uint32_t)k);
  }
  return hip_rc(hipGetLastError());
}

int ts_select(const uint8_t* in, int32_t in_rank, const int32_t* fixed, uint8_t* out, ts_stream stream) {
  if (!in || !out || in_rank < 0 || in_rank > TS_MAX_BITS) return TS_EINVAL;
  uint32_t base = 0;
  std::vector<int32_t> ow;
  int kept = 0;
  for (int i = 0; i < in_rank; ++i) {
    if (fixed[i] < 0) {
      ++kept;
    } else if (fixed[i] > 1) {
      return TS_EINVAL;
    } else if (fixed[i] == 1) {
      base |= 1u << (in_rank - 1 - i);
    }
  }
  // The kept axes, read at the offset of the fixed values: a unary copy.
  for (int i = 0; i < in_rank; ++i)
    if (fixed[i] < 0) ow.push_back(i);
  ts_step st;
  std::memset(&st, 0, sizeof(st));
  std::memset(st.o_lpos, TS_NOPOS, sizeof(st.o_lpos));
  std::memset(st.o_rpos, TS_NOPOS, sizeof(st.o_rpos));
  std::memset(st.c_lpos, TS_NOPOS, sizeof(st.c_lpos));
  std::memset(st.c_rpos, TS_NOPOS, sizeof(st.c_rpos));
  st.lhs_rank = (uint8_t)in_rank;
  st.out_rank = (uint8_t)kept;
  st.flags = TS_STEP_NORHS;
  for (int j = 0; j < kept; ++j) {
    const int axis = ow[kept - 1 - j];
    st.o_lpos[j] = (uint8_t)(in_rank - 1 - axis);
  }
  hipLaunchKernelGGL(k_permute, dim3(grid_for(1ull << kept)), dim3(WG), 0, S(stream), in + base, out, st);
  return hip_rc(hipGetLastError());
}

int ts_as_scalar(const uint8_t* x, int64_t n, int32_t* result, ts_stream stream) {
  if (!x || !result || n < 0) return TS_EINVAL;
  std::vector<uint8_t> h((size_t)n);
  if (n > 0) {
    if (hipMemcpyAsync(h.data(), x, (size_t)n, hipMemcpyDeviceToHost, S(stream)) != hipSuccess) return TS_EHIP;
    if (hipStreamSynchronize(S(stream)) != hipSuccess) return TS_EHIP;
  }
  int all = 1;
  for (uint8_t v : h)
    if (v != 1) all = 0;
  *result = all;
  return TS_OK;
}

int ts_bytes_to_u64(const uint8_t* in, uint64_t* out, int64_t n, ts_stream stream) {
  if (!in || !out || n < 0) return TS_EINVAL;
  if (n == 0) return TS_OK;
  hipLaunchKernelGGL(k_widen, dim3(grid_for(n)), dim3(WG), 0, S(stream), in, out, n);
  return hip_rc(hipGetLastError());
}

int ts_program_run(const ts_job* job, const ts_step* steps, const uint8_t* leaves, uint8_t* arena, int32_t mode,
                   uint8_t* inter, const int64_t* inter_offsets, ts_stream stream) {
  if (!job || !arena || !leaves || (job->num_steps > 0 && !steps)) return TS_EINVAL;
  if (mode != TS_MODE_GT0 && mode != TS_MODE_MOD256) return TS_EINVAL;
  hipStream_t s = S(stream);
  if (job->num_leaf_bytes > job->arena_elems) return TS_EINVAL;
  if (hipMemcpyAsync(arena, leaves + job->leaf_begin, (size_t)job->num_leaf_bytes, hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return TS_EHIP;
  for (int i = 0; i < job->num_steps; ++i) {
    const ts_step& st = steps[job->step_begin + i];
    const uint64_t hi = std::max<uint64_t>({st.lhs_off + (1ull << st.lhs_rank),
                                             (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_off + (1ull << st.rhs_rank),
                                             st.out_off + (1ull << st.out_rank)});
    if (hi > (uint64_t)job->arena_elems) return TS_EINVAL;
    const uint8_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
    int rc = launch_step_u8(arena + st.lhs_off, rhs, arena + st.out_off, st, mode, s);
    if (rc) return rc;
    if (inter) {
      if (hipMemcpyAsync(inter + inter_offsets[i], arena + st.out_off, (size_t)1 << st.out_rank,
                         hipMemcpyDeviceToDevice, s) != hipSuccess)
        return TS_EHIP;
    }
  }
  return TS_OK;
}

int ts_program_run_fused(const ts_job* job, const ts_step* steps, const ts_step* steps_dev, const uint8_t* leaves,
                         uint8_t* arena, int32_t mode, int32_t small_bits, ts_stream stream) {
  if (!job || !arena || !leaves || (job->num_steps > 0 && (!steps || !steps_dev))) return TS_EINVAL;
  if (mode != TS_MODE_GT0 && mode != TS_MODE_MOD256) return TS_EINVAL;
  if (small_bits < 0 || small_bits > TS_MAX_BITS) return TS_EINVAL;
  if (job->num_leaf_bytes > job->arena_elems) return TS_EINVAL;
  const ts_step* rec = steps + job->step_begin;
  for (int i = 0; i < job->num_steps; ++i) {
    const ts_step& st = rec[i];
    const uint64_t hi = std::max<uint64_t>({st.lhs_off + (1ull << st.lhs_rank),
                                             (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_off + (1ull << st.rhs_rank),
                                             st.out_off + (1ull << st.out_rank)});
    if (hi > (uint64_t)job->arena_elems) return TS_EINVAL;
  }
  hipStream_t s = S(stream);
  if (hipMemcpyAsync(arena, leaves + job->leaf_begin, (size_t)job->num_leaf_bytes, hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return TS_EHIP;
  auto small = [&](const ts_step& st) {
    const int r = std::max<int>({st.lhs_rank, (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_rank, st.out_rank});
    return r <= small_bits;
  };
  int i = 0;
  while (i < job->num_steps) {
    if (small(rec[i])) {
      int j = i + 1;
      while (j < job->num_steps && small(rec[j])) ++j;
      hipLaunchKernelGGL(k_step_seq, dim3(1), dim3(WG), 0, s, arena, steps_dev + job->step_begin, i, j, mode);
      if (hipGetLastError() != hipSuccess) return TS_EHIP;
      i = j;
      continue;
    }
    const ts_step& st = rec[i];
    const uint8_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
    const int rc = launch_step_u8(arena + st.lhs_off, rhs, arena + st.out_off, st, mode, s);
    if (rc) return rc;
    ++i;
  }
  return TS_OK;
}

int ts_program_run_gang(const ts_job* jobs, int32_t count, const ts_step* steps, const ts_step* steps_dev,
                        const uint8_t* leaves, uint8_t* arena, int64_t arena_stride, uint8_t* results, int32_t mode,
                        int32_t small_bits, ts_stream stream) {
  if (!jobs || count < 0 || !arena || !leaves || !results) return TS_EINVAL;
  if (mode != TS_MODE_GT0 && mode != TS_MODE_MOD256) return TS_EINVAL;
  if (small_bits < 0 || small_bits > TS_MAX_BITS || arena_stride <= 0 || (arena_stride & 255)) return TS_EINVAL;
  if (count == 0) return TS_OK;
  auto small = [&](const ts_step& st) {
    const int r = std::max<int>({st.lhs_rank, (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_rank, st.out_rank});
    return r <= small_bits;
  };
  // Segments per job: [begin, end) runs of small steps (kind 0) or one large step (kind 1).
  struct Seg {
    int begin, end, kind;
  };
  std::vector<std::vector<Seg>> segs(count);
  size_t phases = 0;
  for (int g = 0; g < count; ++g) {
    const ts_job& job = jobs[g];
    if (job.num_steps > 0 && !(steps && steps_dev)) return TS_EINVAL;
    if (job.arena_elems > arena_stride || job.num_leaf_bytes > job.arena_elems || job.num_leaf_bytes < 0)
      return TS_EINVAL;
    if (job.result_rank < 0 || job.result_rank > TS_MAX_BITS ||
        job.result_off + (1ll << job.result_rank) > job.arena_elems)
      return TS_EINVAL;
    const ts_step* rec = steps + job.step_begin;
    for (int i = 0; i < job.num_steps; ++i) {
      const ts_step& st = rec[i];
      const uint64_t hi = std::max<uint64_t>({st.lhs_off + (1ull << st.lhs_rank),
                                               (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_off + (1ull << st.rhs_rank),
                                               st.out_off + (1ull << st.out_rank)});
      if (hi > (uint64_t)job.arena_elems) return TS_EINVAL;
    }
    int i = 0;
    while (i < job.num_steps) {
      if (small(rec[i])) {
        int k = i + 1;
        while (k < job.num_steps && small(rec[k])) ++k;
        segs[g].push_back({i, k, 0});
        i = k;
      } else {
        segs[g].push_back({i, i + 1, 1});
        ++i;
      }
    }
    phases = std::max(phases, segs[g].size());
  }
  hipStream_t s = S(stream);
  auto job_arena = [&](int g) { return arena + (int64_t)g * arena_stride; };
  for (int g0 = 0; g0 < count; g0 += GANG) {
    const int n = std::min(GANG, count - g0);
    GangCopy c;
    std::memset(&c, 0, sizeof(c));
    int64_t most = 0;
    for (int k = 0; k < n; ++k) {
      c.src[k] = leaves + jobs[g0 + k].leaf_begin;
      c.dst[k] = job_arena(g0 + k);
      c.bytes[k] = jobs[g0 + k].num_leaf_bytes;
      most = std::max(most, c.bytes[k]);
    }
    if (most > 0) {
      const int gx = (int)std::min<int64_t>((most + 4095) / 4096, 1024);
      hipLaunchKernelGGL(k_gang_copy, dim3(gx, n), dim3(WG), 0, s, c);
      if (hipGetLastError() != hipSuccess) return TS_EHIP;
    }
  }
  for (size_t p = 0; p < phases; ++p) {
    // runs of small steps: one workgroup per job
    GangSeq q;
    int nq = 0;
    auto flush_seq = [&]() {
      if (!nq) return TS_OK;
      q.mode = mode;
      hipLaunchKernelGGL(k_gang_seq, dim3(nq), dim3(WG), 0, s, q);
      nq = 0;
      return hip_rc(hipGetLastError());
    };
    // large steps: MFMA-routed ones alone, the others shared out by block ranges
    GangStep t;
    int nt = 0;
    uint64_t want = 0;
    std::vector<uint64_t> wants;
    // row packing for kind-2 steps, launched just before the step launch that uses it
    GangPack pkd;
    int npk = 0;
    uint32_t pk_rows = 0;
    auto flush_pack = [&]() {
      if (!npk) return TS_OK;
      const int gx = (int)std::min<uint32_t>((pk_rows + WG - 1) / WG, 1024);
      hipLaunchKernelGGL(k_gang_pack, dim3(gx, npk), dim3(WG), 0, s, pkd);
      npk = 0;
      pk_rows = 0;
      return hip_rc(hipGetLastError());
    };
    auto flush_step = [&]() {
      if (!nt) return TS_OK;
      const int prc = flush_pack();
      if (prc) return prc;
      // at most 65536 blocks per launch, every job keeping at least one (grid-stride
      // loops take any count); one-block-per-output jobs keep theirs
      const uint64_t cap = 65536;
      uint32_t at = 0;
      for (int k = 0; k < nt; ++k) {
        t.blk[k] = at;
        uint64_t b = wants[k];
        if (t.kind[k] != 1 && want > cap) b = std::max<uint64_t>(1, b * cap / want);
        at += (uint32_t)b;
      }
      t.blk[nt] = at;
      t.n = nt;
      t.mode = mode;
      hipLaunchKernelGGL(k_gang_step, dim3(at), dim3(WG), 0, s, t);
      nt = 0;
      want = 0;
      wants.clear();
      return hip_rc(hipGetLastError());
    };
    for (int g = 0; g < count; ++g) {
      if (p >= segs[g].size()) continue;
      const Seg& sg = segs[g][p];
      const ts_step* rec = steps + jobs[g].step_begin;
      const ts_step* rec_dev = steps_dev + jobs[g].step_begin;
      uint8_t* ar = job_arena(g);
      if (sg.kind == 0) {
        q.steps[nq] = rec_dev + sg.begin;
        q.arena[nq] = ar;
        q.count[nq] = sg.end - sg.begin;
        if (++nq == GANG) {
          const int rc = flush_seq();
          if (rc) return rc;
        }
        continue;
      }
      const ts_step& st = rec[sg.begin];
      const uint8_t* lhs = ar + st.lhs_off;
      const uint8_t* rhs = ar + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
      if (tsd::dense_eligible(st, lhs, rhs)) {
        const int rc = tsd::launch_dense(lhs, rhs, ar + st.out_off, st, mode, s);
        if (rc) return rc;
        continue;
      }
      t.rec[nt] = rec_dev + sg.begin;
      t.arena[nt] = ar;
      t.kind[nt] = dot_shaped(st) ? 1 : 0;
      const uint64_t pk = ((uint64_t)jobs[g].arena_elems + 255) & ~uint64_t(255);
      if (!t.kind[nt] && packable(st) && pk + packed_bytes(st) <= (uint64_t)arena_stride) {
        const int k = st.sum_rank;
        const uint32_t lrows = 1u << (st.lhs_rank - k), rrows = 1u << (st.rhs_rank - k);
        pkd.src[npk] = ar + st.lhs_off;
        pkd.dst[npk] = ar + pk;
        pkd.rows[npk] = lrows;
        pkd.k[npk++] = (uint8_t)k;
        pkd.src[npk] = ar + st.rhs_off;
        pkd.dst[npk] = ar + pk + packed_rhs_at(st);
        pkd.rows[npk] = rrows;
        pkd.k[npk++] = (uint8_t)k;
        pk_rows = std::max(pk_rows, std::max(lrows, rrows));
        t.kind[nt] = 2;
        t.pk[nt] = (uint32_t)pk;
      }
      const uint64_t b = t.kind[nt] == 1 ? (1ull << st.out_rank) : (uint64_t)grid_for(1ull << st.out_rank);
      wants.push_back(b);
      if (t.kind[nt] != 1) want += b;
      if (++nt == GANG) {
        const int rc = flush_step();
        if (rc) return rc;
      }
    }
    int rc = flush_seq();
    if (rc) return rc;
    rc = flush_step();
    if (rc) return rc;
  }
  for (int g0 = 0; g0 < count; g0 += GANG) {
    const int n = std::min(GANG, count - g0);
    GangAny a;
    std::memset(&a, 0, sizeof(a));
    for (int k = 0; k < n; ++k) {
      a.res[k] = job_arena(g0 + k) + jobs[g0 + k].result_off;
      a.bytes[k] = 1ll << jobs[g0 + k].result_rank;
    }
    a.out = results + g0;
    hipLaunchKernelGGL(k_gang_any, dim3(n), dim3(WG), 0, s, a);
    if (hipGetLastError() != hipSuccess) return TS_EHIP;
  }
  return TS_OK;
}

int ts_program_count(const ts_job* job, const ts_step* steps, const uint8_t* leaves, uint64_t* arena,
                     uint64_t* count, int32_t* overflow, ts_stream stream) {
  if (!job || !arena || !leaves || !count || !overflow || (job->num_steps > 0 && !steps)) return TS_EINVAL;
  hipStream_t s = S(stream);
  if (job->num_leaf_bytes > job->arena_elems) return TS_EINVAL;
  if (job->num_leaf_bytes > 0) {
    hipLaunchKernelGGL(k_widen, dim3(grid_for(job->num_leaf_bytes)), dim3(WG), 0, s, leaves + job->leaf_begin,
                       arena, (int64_t)job->num_leaf_bytes);
  }
  for (int i = 0; i < job->num_steps; ++i) {
    const ts_step& st = steps[job->step_begin + i];
    const uint64_t hi = std::max<uint64_t>({st.lhs_off + (1ull << st.lhs_rank),
                                             (st.flags & TS_STEP_NORHS) ? 0 : st.rhs_off + (1ull << st.rhs_rank),
                                             st.out_off + (1ull << st.out_rank)});
    if (hi > (uint64_t)job->arena_elems) return TS_EINVAL;
    const uint64_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
    int rc = launch_step_u64(arena + st.lhs_off, rhs, arena + st.out_off, st, overflow, s);
    if (rc) return rc;
  }
  if (job->result_rank != 0) return TS_EINVAL;
  hipLaunchKernelGGL(k_scale_add, dim3(1), dim3(64), 0, s, arena + job->result_off, job->free_vars, count, overflow);
  return hip_rc(hipGetLastError());
}

}  // extern "C"
