// Batched persistent executor (gfx950): launches that walk a device-resident queue of
// contraction jobs.
//
// Dispatch runs in rounds by cube index (host-chosen boundaries): round 0 holds the first
// cube of every instance; before each later round, k_compact keeps only the jobs whose
// instance is still undecided, and the persistent kernel reads that count from device
// memory, so no host synchronisation is needed between rounds.  Inside a round every
// workgroup takes one queue entry per atomic, re-checks the instance verdict (early
// exit), stages the job's step records in LDS (one coalesced copy per 64 steps), copies
// the leaves into its private arena, runs each step with all 256 threads and ORs the
// scalar result into the instance verdict.  The verdict of an instance is the OR over
// its jobs whether or not later jobs are skipped.
//
// This replaces, for many instances at once, the reference's per-instance Python loop
// (tensorsat/contractions/cotengra.py:120-126 -> _autoray.einsum -> FinRel._contract2,
// fin_rel.py:255-288) and the final bool() (Box.__bool__, diagrams.py:949-958).
//
// Jobs whose arenas fit in 64 KiB (every job of cube widths up to 12 on the batched
// workload) can keep the arena in LDS instead (k_batch<true>): leaves are copied in and
// every intermediate is written and read there.  Steps with at most 64 out entries (most
// steps) then run on one wave with offsets scattered directly from the record, and a
// workgroup barrier is taken only where a larger step (4-wave table engine) follows.
//
// Every loop bound comes from the job/step records or the queue length; the dispatch loop
// ends once the counter passes the queue length, which every workgroup reaches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "device.h"

using namespace tsd;

namespace {

constexpr int SREC = 64;  // step records staged in LDS at a time

__device__ __forceinline__ bool decided(const uint32_t* verdicts, int instance) {
  return __hip_atomic_load(&verdicts[instance], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// The first two rounds of the queue are assigned statically: workgroup w takes position
// w, then position G + w (G = num_workgroups), so no atomic is needed (every workgroup
// starts at once, and thousands of atomics on one address at launch serialise); the
// host orders both rounds so that the workgroups dispatched first, which alone get a
// second job, pair short first jobs with long second ones.  Positions from 2G on are split into QSEG interleaved segments (position
// 2G + k belongs to segment k % QSEG, so each keeps the queue's order) with one counter
// each, TS_BATCH_COUNTER_STRIDE int32s (one 256-byte line) apart; workgroup w takes them
// from segment w % QSEG, i.e. from the segment of its own XCD.  A plain load of the
// counter first keeps workgroups that find their segment exhausted off the atomic.
// `taken` counts the workgroup's takes; returns the next job, or -1 once the
// workgroup's share is exhausted.
constexpr int QSEG = TS_BATCH_SEGMENTS;
constexpr int QSTRIDE = TS_BATCH_COUNTER_STRIDE;

__device__ __forceinline__ int take_job(const ts_batch& b, int num_jobs, int& taken) {
  const int w = (int)blockIdx.x, g = b.num_workgroups;
  while (taken < 2) {
    const int64_t pos = taken == 0 ? (int64_t)w : (int64_t)g + w;
    ++taken;
    if (pos < num_jobs) return b.order ? b.order[pos] : (int)pos;
  }
  const int64_t rest = (int64_t)num_jobs - 2 * (int64_t)g;
  const int x = w % QSEG;
  const int seg_len = rest > x ? (int)((rest - x + QSEG - 1) / QSEG) : 0;
  int32_t* c = b.counter + x * QSTRIDE;
  if (seg_len == 0 || __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= seg_len) return -1;
  const int kl = atomicAdd(c, 1);
  const int64_t pos = 2 * (int64_t)g + x + (int64_t)QSEG * kl;
  return kl < seg_len ? (b.order ? b.order[pos] : (int)pos) : -1;
}

// Offset of the bits of v (LSB-first, nbits <= NB of them) placed at positions pos[i];
// TS_NOPOS entries contribute nothing.
template <int NB>
__device__ __forceinline__ uint32_t scatter_bits(uint32_t v, int nbits, const uint8_t* pos) {
  uint32_t off = 0;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (i < nbits) {
      const uint32_t p = pos[i];
      if (((v >> i) & 1u) && p != TS_NOPOS) off |= 1u << p;
    }
  }
  return off;
}

// Steps with at most 64 out entries run on one wave: lane = out index, offsets scattered
// directly from the record's bit positions (no tables, no workgroup barrier).  The record
// (in LDS) is read as a few aligned dwords issued together, and the scatter is
// branch-free on the position bytes held in registers.
__device__ __forceinline__ bool small_step(const ts_step& st) { return st.out_rank <= 6; }

// 8 bytes of the record starting at byte `off` (compile-time), from aligned dword reads.
__device__ __forceinline__ uint64_t rec_bytes8(const uint32_t* w, int off) {
  const int q = off >> 2, r = off & 3;
  const uint64_t lo = (uint64_t)w[q] | ((uint64_t)w[q + 1] << 32);
  if (r == 0) return lo;
  const uint64_t hi = w[q + 2];
  return (lo >> (8 * r)) | (hi << (64 - 8 * r));
}

// Offset of an index whose bits 0..5 are lb[0..5] (0/1) placed at the positions packed
// in pw (one byte per bit, TS_NOPOS = none), only the first n of them taken.  Evaluated
// per lane in vector registers, one bit-field extract and one shift-or per bit: bytes
// from the n-th on are forced to TS_NOPOS first, every TS_NOPOS byte shifts its bit to
// position 31 (255 mod 32), and bit 31 is cleared at the end (real positions are below
// TS_MAX_BITS <= 31).
__device__ __forceinline__ uint32_t scatter6(const uint32_t (&lb)[6], int n, uint64_t pw) {
  pw |= ~0ull << (8 * n);  // n <= 6
  uint32_t off = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) off |= lb[i] << ((uint32_t)(pw >> (8 * i)) & 31u);
  return off & 0x7fffffffu;
}

// A zero the compiler cannot see through: record words read through an address offset
// by it are held per lane, so their decoding runs on the vector ALUs (four per CU)
// rather than the one scalar ALU that every wave of the CU shares.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

__device__ inline void step_small(const ts_step& st, uint8_t* arena, int mode, int lane) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(&st);
  // Header: lhs_off, rhs_off, out_off, ranks (bytes 12..15), flags (byte 16).
  const uint32_t lhs_off = w[0], rhs_off = w[1], out_off = w[2], ranks = w[3], fl = w[4];
  const uint32_t* wv = w + opaque_zero();  // position words, decoded per lane
  const uint64_t lpw = rec_bytes8(wv, (int)offsetof(ts_step, o_lpos));
  const uint64_t rpw = rec_bytes8(wv, (int)offsetof(ts_step, o_rpos));
  const int ro = (int)((ranks >> 16) & 0xffu), nc = (int)(ranks >> 24);
  const bool norhs = fl & TS_STEP_NORHS, kinner = fl & TS_STEP_KINNER;
  const uint8_t* lhs = arena + lhs_off;
  const uint8_t* rhs = arena + (norhs ? lhs_off : rhs_off);
  // Every lane stays active (cross-lane reads below); lanes >= 2^ro compute in-bounds
  // aliases (only the low ro bits of the lane index are scattered) and do not store.
  uint32_t lb[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lb[i] = ((uint32_t)lane >> i) & 1u;
  const uint32_t offl = scatter6(lb, ro, lpw);
  const uint32_t offr = scatter6(lb, ro, rpw);
  uint32_t acc;
  if (nc == 0) {
    acc = lhs[offl];
    if (!norhs) acc &= rhs[offr];
  } else if (kinner) {
    acc = popc_and(lhs + offl, rhs + offr, nc, norhs);
  } else {
    // Lane t holds the offsets of the low (<= 6) summed bits of term t; the term loop
    // reads them across lanes.  Summed bits past the sixth (leaves of more than 6 bits)
    // are walked in blocks of 64 terms, their offsets built by a short loop per block.
    const uint64_t cl = rec_bytes8(wv, (int)offsetof(ts_step, c_lpos));
    const uint64_t cr = rec_bytes8(wv, (int)offsetof(ts_step, c_rpos));
    const int nlo = nc < 6 ? nc : 6, nhi = nc - nlo;
    const int tl = (int)scatter6(lb, nlo, cl), tr = (int)scatter6(lb, nlo, cr);
    acc = 0;
    const int nterm = 1 << nlo;
    for (int chi = 0; chi < (1 << nhi); ++chi) {
      uint32_t hl = 0, hr = 0;
#pragma unroll 1
      for (int i = 0; i < nhi; ++i) {
        if ((chi >> i) & 1) {
          const uint32_t pl = st.c_lpos[6 + i], pr = st.c_rpos[6 + i];
          hl |= pl != TS_NOPOS ? 1u << pl : 0u;
          hr |= pr != TS_NOPOS ? 1u << pr : 0u;
        }
      }
      const uint32_t bl = offl | hl, br = offr | hr;
      for (int c = 0; c < nterm; ++c) {
        const uint32_t a = lhs[bl | (uint32_t)__shfl(tl, c)];
        const uint32_t bb = norhs ? 1u : rhs[br | (uint32_t)__shfl(tr, c)];
        acc += a & bb;
      }
    }
  }
  if (lane < (1 << ro))
    arena[out_off + (uint32_t)lane] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
}

// One packet of a packet-form job (tsc_packetize) on one wave: lane l works for the
// packet's record that owns it (the one whose first lane is the highest <= l, found by
// counting the set bits of the packet's lane mask below l) on out index le = l - first
// lane.  Everything is per lane: the record's header and positions, the K-innermost
// popcount of 2^sum_rank bytes, and the summed-term loop of general steps, whose term
// offsets lane first+t holds for term t; the loops that read across lanes run with the
// whole wave (trip counts from ballots), the lanes that have no such term masked.
// Lanes past the packet's last step alias its index range (bits >= out_rank are not
// scattered) and store nothing.
__device__ inline void step_packet(const ts_step* recs, int my, uint8_t* arena, int mode, int lane) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(recs + my);
  const uint32_t lhs_off = w[0], rhs_off = w[1], out_off = w[2], ranks = w[3], fl4 = w[4];
  const int ro = (int)((ranks >> 16) & 0xffu), nc = (int)(ranks >> 24);
  const bool norhs = fl4 & TS_STEP_NORHS, kinner = fl4 & TS_STEP_KINNER;
  const int first = (int)((fl4 >> 8) & 0xffu);  // pad[0]
  const uint32_t le = (uint32_t)(lane - first);
  uint32_t lb[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lb[i] = (le >> i) & 1u;
  const uint32_t offl = scatter6(lb, ro, rec_bytes8(w, (int)offsetof(ts_step, o_lpos)));
  const uint32_t offr = scatter6(lb, ro, rec_bytes8(w, (int)offsetof(ts_step, o_rpos)));
  const uint8_t* lhs = arena + lhs_off;
  const uint8_t* rhs = arena + (norhs ? lhs_off : rhs_off);
  const bool general = !kinner && nc > 0;
  uint32_t acc = 0;
  if (!general) {
    if (nc == 0) {
      acc = lhs[offl];
      if (!norhs) acc &= rhs[offr];
    } else {
      acc = popc_and(lhs + offl, rhs + offr, nc, norhs);
    }
  }
  if (__ballot(general) != 0) {
    const int nlo = general ? (nc < 6 ? nc : 6) : 0, nhi = general ? nc - nlo : 0;
    const int tl = (int)scatter6(lb, nlo, rec_bytes8(w, (int)offsetof(ts_step, c_lpos)));
    const int tr = (int)scatter6(lb, nlo, rec_bytes8(w, (int)offsetof(ts_step, c_rpos)));
    const int nterm = 1 << nlo, nblk = general ? 1 << nhi : 0;
    const uint8_t* cpl = reinterpret_cast<const uint8_t*>(w) + offsetof(ts_step, c_lpos);
    const uint8_t* cpr = reinterpret_cast<const uint8_t*>(w) + offsetof(ts_step, c_rpos);
    uint32_t g = 0;
    for (int chi = 0; __ballot(chi < nblk) != 0; ++chi) {
      uint32_t hl = 0, hr = 0;
      for (int i = 0; i < nhi; ++i) {
        if ((chi >> i) & 1) {
          const uint32_t pl = cpl[6 + i], pr = cpr[6 + i];
          hl |= pl != TS_NOPOS ? 1u << pl : 0u;
          hr |= pr != TS_NOPOS ? 1u << pr : 0u;
        }
      }
      const uint32_t bl = offl | hl, br = offr | hr;
      for (int c = 0; __ballot(chi < nblk && c < nterm) != 0; ++c) {
        const uint32_t tlc = (uint32_t)__shfl(tl, (first + c) & 63);
        const uint32_t trc = (uint32_t)__shfl(tr, (first + c) & 63);
        if (chi < nblk && c < nterm) g += (uint32_t)lhs[bl | tlc] & (norhs ? 1u : (uint32_t)rhs[br | trc]);
      }
    }
    if (general) acc = g;
  }
  if (le < (1u << ro))
    arena[out_off + le] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
}

// One packet of a lane-form job (tsc_lanes): lane l < count runs its own record, which
// holds absolute arena offsets (all index math was done on the host).  Sum-free and
// K-innermost lanes AND and count 2^nc contiguous bytes; general lanes walk the step's
// summed terms, whose offsets the table lanes first .. first + 2^nc - 1 hold, with the
// whole wave taking part in the cross-lane reads (trip counts from ballots).
__device__ inline void lane_packet(uint4 r, bool active, uint8_t* arena, int mode, int lane) {
  const uint32_t lhs = r.x & 0xffffu, rhs = r.x >> 16, out = r.y & 0xffffu;
  const int tl = (int)(r.y >> 16), tr = (int)(r.z & 0xffffu);
  const uint32_t op = (r.z >> 16) & 0xffu;
  const int first = (int)(r.z >> 24), nc = (int)(op & 7u);
  const bool general = active && (op & TS_LANE_GENERAL), norhs = op & TS_LANE_NORHS;
  uint32_t acc = 0;
  if (active && !general) {
    if (nc == 0)
      acc = (uint32_t)arena[lhs] & (norhs ? 1u : (uint32_t)arena[rhs]);
    else
      acc = popc_and(arena + lhs, arena + rhs, nc, norhs);
  }
  if (__ballot(general) != 0) {
    const int nterm = general ? 1 << nc : 0;
    // Both 16-bit term offsets travel in one cross-lane read.
    const uint32_t tlr = (uint32_t)tl | ((uint32_t)tr << 16);
    uint32_t g = 0;
    for (int c = 0; __ballot(c < nterm) != 0; ++c) {
      const uint32_t t2 = (uint32_t)__shfl((int)tlr, (first + c) & 63);
      const uint32_t tlc = t2 & 0xffffu, trc = t2 >> 16;
      if (c < nterm) g += (uint32_t)arena[lhs + tlc] & (norhs ? 1u : (uint32_t)arena[rhs + trc]);
    }
    if (general) acc = g;
  }
  if (active && out != TS_LANE_NONE)
    arena[out] = (uint8_t)(mode == TS_MODE_MOD256 ? ((acc & 255u) != 0) : (acc != 0));
}

// LDSA: the job arena lives in LDS (dynamic shared memory of arena_stride bytes; every
// tensor then has at most 16 bits, so two table chunks suffice) instead of device memory.
template <bool LDSA>
__global__ __launch_bounds__(WG) void k_batch(ts_batch b) {
  constexpr int NQ = LDSA ? 2 : 4;
  constexpr int NREC = LDSA ? 16 : SREC;
  extern __shared__ __attribute__((aligned(16))) uint8_t dyn_arena[];
  __shared__ LutT<NQ> L;
  __shared__ ts_step s_steps[NREC];
  __shared__ int s_job;
  const int tid = threadIdx.x;
  uint8_t* __restrict__ arena = LDSA ? dyn_arena : b.arenas + (int64_t)blockIdx.x * b.arena_stride;
  const int num_jobs = b.num_jobs_dev ? *b.num_jobs_dev : b.num_jobs;
  int taken = 0;  // queue state of thread 0
  for (;;) {
    if (tid == 0) {
      int j = take_job(b, num_jobs, taken);
      if (j >= 0 && b.early_exit && decided(b.verdicts, b.jobs[j].instance)) {
        if (b.job_done) b.job_done[j] = 2;
        j = -2;  // skipped: take the next entry
      }
      s_job = j;
    }
    __syncthreads();
    const int j = s_job;
    __syncthreads();
    if (j == -1) break;
    if (j < 0) continue;
    const ts_job job = b.jobs[j];
    // Leaves -> arena (16-byte aligned prefix).
    {
      const uint8_t* src = b.leaves + job.leaf_begin;
      const int nbytes = job.num_leaf_bytes;
      const int n16 = nbytes >> 4;
      for (int i = tid; i < n16; i += WG)
        reinterpret_cast<uint4*>(arena)[i] = reinterpret_cast<const uint4*>(src)[i];
      for (int i = (n16 << 4) + tid; i < nbytes; i += WG) arena[i] = src[i];
    }
    for (int s0 = 0; s0 < job.num_steps; s0 += NREC) {
      const int ns = min(NREC, job.num_steps - s0);
      // Stage up to NREC step records (128 bytes = 8 x uint4 each); every wave is past
      // the previous chunk (its last step ends with a barrier).
      const uint4* src = reinterpret_cast<const uint4*>(b.steps + job.step_begin + s0);
      for (int i = tid; i < ns * 8; i += WG) reinterpret_cast<uint4*>(s_steps)[i] = src[i];
      __syncthreads();
      for (int s = 0; s < ns; ++s) {
        const ts_step& st = s_steps[s];
        if (LDSA && small_step(st)) {
          // Wave 0 alone; its LDS writes are ordered before its own later reads.  A
          // workgroup barrier is needed only before a large step and at the chunk end.
          if (tid < 64) {
            step_small(st, arena, b.mode, tid);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
          }
          if (s + 1 == ns || !small_step(s_steps[s + 1])) __syncthreads();
          continue;
        }
        const bool need_c = !(st.flags & TS_STEP_KINNER) && st.sum_rank > 0;
        build_lut(st, L, need_c, tid);
        __syncthreads();
        const uint8_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
        step_u8(st, L, arena + st.lhs_off, rhs, arena + st.out_off, (uint64_t)tid, (uint64_t)WG, b.mode);
        __syncthreads();
      }
    }
    __syncthreads();  // leaves of step-less jobs were written by all threads
    if (tid < 64) {
      // Decision jobs end in a scalar; a non-scalar result counts as satisfiable when
      // any entry is set.
      const int64_t nr = int64_t(1) << job.result_rank;
      uint32_t v = 0;
      for (int64_t i = tid; i < nr; i += 64) v |= arena[job.result_off + i];
      const bool any = __ballot(v != 0) != 0;
      if (tid == 0) {
        if (any) atomicOr(&b.verdicts[job.instance], 1u);
        if (b.job_done) b.job_done[j] = 1;
      }
    }
    __syncthreads();
  }
}

// Programs whose tensors all have at most 6 bits (cube width <= 6): every step has at
// most 64 out entries, so a job needs one wave.  One-wave workgroups with the arena in
// LDS: 4x the concurrent jobs of the 4-wave kernel per CU.  The workgroup is one
// wavefront, whose LDS accesses execute in order, so steps are separated by
// wavefront-scope fences (ordering for the compiler, no wait).  Step records are staged
// NR at a time, the next chunk loaded into registers while the current one runs.
template <int FORM>  // 0: step records, 1: packet form, 2: lane form
__global__ __launch_bounds__(64) void k_batch_wave(ts_batch b) {
  extern __shared__ __attribute__((aligned(16))) uint8_t arena[];
  constexpr int NR = 16;
  static_assert(NR * 8 == 128, "two uint4 of records per lane");
  __shared__ ts_step s_steps[NR];
  uint4* const s16 = reinterpret_cast<uint4*>(s_steps);
  const int lane = threadIdx.x;
  const int num_jobs = b.num_jobs_dev ? *b.num_jobs_dev : b.num_jobs;
  int taken = 0;  // queue state of lane 0
  for (;;) {
    int j = -1;
    if (lane == 0) {
      for (;;) {
        j = take_job(b, num_jobs, taken);
        // The first job starts with the dispatch, before any verdict can be set: no check.
        if (j >= 0 && taken > 1 && b.early_exit && decided(b.verdicts, b.jobs[j].instance)) {
          if (b.job_done) b.job_done[j] = 2;
          continue;  // skipped: take the next entry
        }
        break;
      }
    }
    j = __shfl(j, 0);
    if (j < 0) break;
    const ts_job job = b.jobs[j];
    // The first chunk of step records and the leaves are loaded together: one memory
    // round trip before the first step.
    // Records in uint4 units: 8 per step record, 1 per lane record (lane form: step_begin
    // and num_steps count lane records); staged 128 units (2 KiB) at a time.
    const uint4* rsrc = FORM == 2 ? reinterpret_cast<const uint4*>(b.steps) + job.step_begin
                                  : reinterpret_cast<const uint4*>(b.steps + job.step_begin);
    const int units = FORM == 2 ? job.num_steps : job.num_steps * 8;
    int n = min(NR * 8, units);
    uint4 r0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0};
    if (lane < n) r0 = rsrc[lane];
    if (lane + 64 < n) r1 = rsrc[lane + 64];
    {
      const uint8_t* src = b.leaves + job.leaf_begin;
      const int nbytes = job.num_leaf_bytes;
      const int n16 = nbytes >> 4;
#pragma unroll 1
      for (int i = lane; i < n16; i += 64) reinterpret_cast<uint4*>(arena)[i] = reinterpret_cast<const uint4*>(src)[i];
      if (lane < nbytes - (n16 << 4)) arena[(n16 << 4) + lane] = src[(n16 << 4) + lane];  // < 16 tail bytes
    }
    for (int u0 = 0; u0 < units; u0 += NR * 8) {
      // r0/r1 hold units u0.. (n uint4): stage them, then fetch the next chunk.
      if (lane < n) s16[lane] = r0;
      if (lane + 64 < n) s16[lane + 64] = r1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int nu = min(NR * 8, units - u0), ns = nu >> 3, s0 = u0 >> 3;
      n = max(0, min(NR * 8, units - u0 - NR * 8));
      if (lane < n) r0 = rsrc[u0 + NR * 8 + lane];
      if (lane + 64 < n) r1 = rsrc[u0 + NR * 8 + lane + 64];
      (void)s0;
      if constexpr (FORM == 2) {
        // Packets never straddle a chunk; the first record of each holds its lane count
        // (or, on a filler, the records to skip).  Counts are clamped to the chunk.
        for (int q = 0; q < nu;) {
          const bool in = lane < nu - q;
          const uint4 r = in ? s16[q + lane] : make_uint4(0, 0, 0, 0);
          const uint32_t head = __builtin_amdgcn_readfirstlane(r.w), hop = __builtin_amdgcn_readfirstlane(r.z);
          const int cnt = max((int)(head & 0xffu), 1);
          if ((hop >> 16) & TS_LANE_SKIP) {
            q += cnt;
            continue;
          }
          const int k = min(min(cnt, 64), nu - q);
          lane_packet(r, lane < k, arena, b.mode, lane);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          q += k;
        }
      } else if constexpr (FORM == 1) {
        // Packets never straddle a chunk; the first record of each holds its lane mask
        // (o_lpos[16..23], bytes 36..43).  Counts are clamped to the chunk so a malformed
        // mask cannot stall the loop or index past the staged records.
        for (int s = 0; s < ns;) {
          const uint32_t* w0 = reinterpret_cast<const uint32_t*>(&s_steps[s]);
          const uint64_t mask = (uint64_t)w0[9] | ((uint64_t)w0[10] << 32);
          const int k = min(max(__popcll(mask), 1), ns - s);
          const uint64_t m1 = mask >> 1;
          const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
          step_packet(s_steps, s + min(below, k - 1), arena, b.mode, lane);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          s += k;
        }
      } else {
        for (int s = 0; s < ns; ++s) {
          step_small(s_steps[s], arena, b.mode, lane);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
    }
    const int64_t nr = int64_t(1) << job.result_rank;
    uint32_t v = 0;
    for (int64_t i = lane; i < nr; i += 64) v |= arena[job.result_off + i];
    const bool any = __ballot(v != 0) != 0;
    if (lane == 0) {
      if (any) atomicOr(&b.verdicts[job.instance], 1u);
      if (b.job_done) b.job_done[j] = 1;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the next job's leaves overwrite the arena
  }
}

// Counting semiring: the same walk with uint64 arenas (arena_stride is in bytes) and
// checked arithmetic; each job's scalar result times 2^free_vars is added to its
// instance's count; any overflow sets the instance's flag.
__global__ __launch_bounds__(WG) void k_batch_count(ts_batch b, unsigned long long* __restrict__ counts,
                                                    int32_t* __restrict__ overflow) {
  __shared__ Lut L;
  __shared__ ts_step s_steps[SREC];
  __shared__ int s_job;
  __shared__ int s_ovf;
  const int tid = threadIdx.x;
  uint64_t* __restrict__ arena =
      reinterpret_cast<uint64_t*>(b.arenas + (int64_t)blockIdx.x * b.arena_stride);
  const int num_jobs = b.num_jobs_dev ? *b.num_jobs_dev : b.num_jobs;
  int taken = 0;  // queue state of thread 0
  for (;;) {
    if (tid == 0) {
      s_job = take_job(b, num_jobs, taken);
      s_ovf = 0;
    }
    __syncthreads();
    const int j = s_job;
    if (j < 0) break;
    const ts_job job = b.jobs[j];
    {
      const uint8_t* src = b.leaves + job.leaf_begin;
      for (int i = tid; i < job.num_leaf_bytes; i += WG) arena[i] = src[i];
    }
    bool ovf = false;
    for (int s0 = 0; s0 < job.num_steps; s0 += SREC) {
      const int ns = min(SREC, job.num_steps - s0);
      const uint4* src = reinterpret_cast<const uint4*>(b.steps + job.step_begin + s0);
      for (int i = tid; i < ns * 8; i += WG) reinterpret_cast<uint4*>(s_steps)[i] = src[i];
      __syncthreads();
      for (int s = 0; s < ns; ++s) {
        const ts_step& st = s_steps[s];
        const bool need_c = !(st.flags & TS_STEP_KINNER) && st.sum_rank > 0;
        build_lut(st, L, need_c, tid);
        __syncthreads();
        const uint64_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
        ovf |= step_u64(st, L, arena + st.lhs_off, rhs, arena + st.out_off, (uint64_t)tid, (uint64_t)WG);
        __syncthreads();
      }
    }
    if (ovf) atomicOr(&s_ovf, 1);
    __syncthreads();
    if (tid == 0) {
      // Decision and counting jobs end in a scalar; a non-scalar result adds its entries.
      unsigned long long total = 0;
      bool bad = s_ovf != 0;
      const int64_t nr = int64_t(1) << job.result_rank;
      for (int64_t i = 0; i < nr; ++i) {
        const unsigned long long v = arena[job.result_off + i];
        if (total + v < total) bad = true;
        total += v;
      }
      const int sh = job.free_vars;
      if (sh >= 64 ? total != 0 : (sh > 0 && (total >> (64 - sh)) != 0)) bad = true;
      const unsigned long long add = sh >= 64 ? 0ull : (total << sh);
      const unsigned long long old = atomicAdd(&counts[job.instance], add);
      if (old + add < old) bad = true;
      if (bad) atomicOr(&overflow[job.instance], 1);
      if (b.job_done) b.job_done[j] = 1;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(WG) void k_compact(const int32_t* __restrict__ order, int n,
                                                const ts_job* __restrict__ jobs,
                                                const uint32_t* __restrict__ verdicts,
                                                int32_t* __restrict__ out, int32_t* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  for (int base = blockIdx.x * WG; base < n; base += gridDim.x * WG) {
    const int k = base + threadIdx.x;
    int j = -1;
    bool keep = false;
    if (k < n) {
      j = order[k];
      keep = !decided(verdicts, jobs[j].instance);
    }
    const uint64_t mask = __ballot(keep);
    const int cnt = __popcll(mask);
    int first = 0;
    if (lane == 0 && cnt) first = atomicAdd(count, cnt);
    first = __shfl(first, 0);
    if (keep) out[first + __popcll(mask & ((1ull << lane) - 1ull))] = j;
  }
}

}  // namespace

extern "C" {

int ts_batch_max_workgroups(int32_t* out) {
  if (!out) return TS_EINVAL;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TS_EHIP;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return TS_EHIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_batch<false>), WG, 0) !=
      hipSuccess)
    return TS_EHIP;
  *out = ncu * (per > 0 ? per : 1);
  return TS_OK;
}

int ts_batch_max_workgroups_lds(int64_t arena_bytes, int32_t one_wave, int32_t* out) {
  if (!out || arena_bytes <= 0 || arena_bytes > TS_BATCH_LDS_MAX || (arena_bytes & 255) || one_wave < 0 ||
      one_wave > 3)
    return TS_EINVAL;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TS_EHIP;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return TS_EHIP;
  const void* fn = one_wave == 3   ? reinterpret_cast<const void*>(k_batch_wave<2>)
                   : one_wave == 2 ? reinterpret_cast<const void*>(k_batch_wave<1>)
                   : one_wave == 1 ? reinterpret_cast<const void*>(k_batch_wave<0>)
                                   : reinterpret_cast<const void*>(k_batch<true>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, one_wave ? 64 : WG, (size_t)arena_bytes) != hipSuccess)
    return TS_EHIP;
  *out = ncu * (per > 0 ? per : 1);
  return TS_OK;
}

int ts_batch_run(const ts_batch* b, ts_stream stream) {
  if (!b || b->num_jobs < 0 || b->num_workgroups <= 0 || !b->counter || !b->verdicts) return TS_EINVAL;
  if (b->num_jobs > 0 && (!b->jobs || !b->steps || !b->leaves)) return TS_EINVAL;
  if (b->num_jobs_dev && !b->order) return TS_EINVAL;  // compacted lists are explicit orders
  if (b->num_jobs > 0 && !b->lds_arena && !b->arenas) return TS_EINVAL;
  if (b->arena_stride <= 0 || (b->arena_stride & 255)) return TS_EINVAL;
  if (b->lds_arena && b->arena_stride > TS_BATCH_LDS_MAX) return TS_EINVAL;
  if (b->lds_arena < 0 || b->lds_arena > 4) return TS_EINVAL;
  // Segment x of the queue past the first 2 * num_workgroups positions is served by
  // workgroups w with w % QSEG == x: every non-empty segment needs one.
  if (b->num_workgroups < QSEG && b->num_jobs > 2 * (int64_t)b->num_workgroups) return TS_EINVAL;
  if (b->mode != TS_MODE_GT0 && b->mode != TS_MODE_MOD256) return TS_EINVAL;
  if (b->num_jobs == 0) return TS_OK;
  if (b->lds_arena == 4)
    hipLaunchKernelGGL(k_batch_wave<2>, dim3(b->num_workgroups), dim3(64), (size_t)b->arena_stride,
                       reinterpret_cast<hipStream_t>(stream), *b);
  else if (b->lds_arena == 3)
    hipLaunchKernelGGL(k_batch_wave<1>, dim3(b->num_workgroups), dim3(64), (size_t)b->arena_stride,
                       reinterpret_cast<hipStream_t>(stream), *b);
  else if (b->lds_arena == 2)
    hipLaunchKernelGGL(k_batch_wave<0>, dim3(b->num_workgroups), dim3(64), (size_t)b->arena_stride,
                       reinterpret_cast<hipStream_t>(stream), *b);
  else if (b->lds_arena)
    hipLaunchKernelGGL((k_batch<true>), dim3(b->num_workgroups), dim3(WG), (size_t)b->arena_stride,
                       reinterpret_cast<hipStream_t>(stream), *b);
  else
    hipLaunchKernelGGL((k_batch<false>), dim3(b->num_workgroups), dim3(WG), 0, reinterpret_cast<hipStream_t>(stream),
                       *b);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_batch_count(const ts_batch* b, uint64_t* counts, int32_t* overflow, ts_stream stream) {
  if (!b || !counts || !overflow || b->num_jobs < 0 || b->num_workgroups <= 0 || !b->counter) return TS_EINVAL;
  if (b->num_workgroups < QSEG && b->num_jobs > 2 * (int64_t)b->num_workgroups) return TS_EINVAL;
  if (b->num_jobs > 0 && (!b->jobs || !b->steps || !b->leaves || !b->arenas)) return TS_EINVAL;
  if (b->arena_stride <= 0 || (b->arena_stride & 255)) return TS_EINVAL;
  if (b->num_jobs == 0) return TS_OK;
  hipLaunchKernelGGL(k_batch_count, dim3(b->num_workgroups), dim3(WG), 0, reinterpret_cast<hipStream_t>(stream),
                     *b, reinterpret_cast<unsigned long long*>(counts), overflow);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_batch_compact(const int32_t* order, int32_t n, const ts_job* jobs, const uint32_t* verdicts, int32_t* out,
                     int32_t* out_count, ts_stream stream) {
  if (n < 0 || (n > 0 && (!order || !jobs || !verdicts || !out || !out_count))) return TS_EINVAL;
  if (n == 0) return TS_OK;
  const int grid = (int)std::min<int64_t>((n + WG - 1) / WG, 4096);
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(WG), 0, reinterpret_cast<hipStream_t>(stream), order, n, jobs,
                     verdicts, out, out_count);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // extern "C"
