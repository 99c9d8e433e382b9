// Batched persistent executor (gfx950): one launch walks a device-resident queue of
// contraction jobs.  Each workgroup repeatedly takes the next job index from a global
// counter, copies the job's leaves into its private arena, runs the job's steps with the
// whole workgroup (one __syncthreads between steps) and ORs the scalar result into the
// job's instance verdict.  With early exit, a job whose instance is already decided SAT
// is skipped: the verdict of an instance is the OR over its jobs either way.
//
// This replaces, for many instances at once, the reference's per-instance Python loop
// (tensorsat/contractions/cotengra.py:120-126 -> _autoray.einsum -> FinRel._contract2,
// fin_rel.py:255-288) and the final bool() (Box.__bool__, diagrams.py:949-958).
//
// Every loop bound comes from the job/step records; the dispatch loop ends when the
// counter passes num_jobs, which every workgroup reaches.

#include <hip/hip_runtime.h>

#include "device.h"

using namespace tsd;

namespace {

__global__ __launch_bounds__(WG) void k_batch(ts_batch b) {
  __shared__ Lut L;
  __shared__ int s_job;
  __shared__ int s_skip;
  const int tid = threadIdx.x;
  uint8_t* __restrict__ arena = b.arenas + (int64_t)blockIdx.x * b.arena_stride;
  for (;;) {
    if (tid == 0) {
      const int k = atomicAdd(b.counter, 1);
      int j = k < b.num_jobs ? b.order[k] : -1;
      int skip = 0;
      if (j >= 0 && b.early_exit) {
        const uint32_t v = __hip_atomic_load(&b.verdicts[b.jobs[j].instance], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        skip = v != 0;
      }
      s_job = j;
      s_skip = skip;
    }
    __syncthreads();
    const int j = s_job;
    const int skip = s_skip;
    __syncthreads();
    if (j < 0) break;
    const ts_job job = b.jobs[j];
    if (skip) {
      if (tid == 0 && b.job_done) b.job_done[j] = 2;
      continue;
    }
    // Leaves -> arena (16-byte aligned prefix).
    {
      const uint8_t* src = b.leaves + job.leaf_begin;
      const int nbytes = job.num_leaf_bytes;
      const int n16 = nbytes >> 4;
      for (int i = tid; i < n16; i += WG)
        reinterpret_cast<uint4*>(arena)[i] = reinterpret_cast<const uint4*>(src)[i];
      for (int i = (n16 << 4) + tid; i < nbytes; i += WG) arena[i] = src[i];
    }
    __syncthreads();
    for (int s = 0; s < job.num_steps; ++s) {
      const ts_step& st = b.steps[job.step_begin + s];
      const bool need_c = !(st.flags & TS_STEP_KINNER) && st.sum_rank > 0;
      build_lut(st, L, need_c, tid);
      __syncthreads();
      const uint8_t* rhs = arena + ((st.flags & TS_STEP_NORHS) ? st.lhs_off : st.rhs_off);
      step_u8(st, L, arena + st.lhs_off, rhs, arena + st.out_off, (uint64_t)tid, (uint64_t)WG, b.mode);
      __syncthreads();
    }
    if (tid == 0) {
      // Decision jobs end in a scalar; a non-scalar result counts as satisfiable when
      // any entry is set.
      uint32_t v = 0;
      const int64_t n = int64_t(1) << job.result_rank;
      for (int64_t i = 0; i < n && !v; ++i) v = arena[job.result_off + i] != 0;
      if (v) atomicOr(&b.verdicts[job.instance], 1u);
      if (b.job_done) b.job_done[j] = 1;
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int ts_batch_max_workgroups(int32_t* out) {
  if (!out) return TS_EINVAL;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TS_EHIP;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return TS_EHIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_batch), WG, 0) !=
      hipSuccess)
    return TS_EHIP;
  *out = ncu * (per > 0 ? per : 1);
  return TS_OK;
}

int ts_batch_run(const ts_batch* b, ts_stream stream) {
  if (!b || b->num_jobs < 0 || b->num_workgroups <= 0 || !b->counter || !b->verdicts) return TS_EINVAL;
  if (b->num_jobs > 0 && (!b->jobs || !b->order || !b->steps || !b->leaves || !b->arenas)) return TS_EINVAL;
  if (b->arena_stride <= 0 || (b->arena_stride & 255)) return TS_EINVAL;
  if (b->mode != TS_MODE_GT0 && b->mode != TS_MODE_MOD256) return TS_EINVAL;
  if (b->num_jobs == 0) return TS_OK;
  hipLaunchKernelGGL(k_batch, dim3(b->num_workgroups), dim3(WG), 0, reinterpret_cast<hipStream_t>(stream), *b);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // extern "C"
