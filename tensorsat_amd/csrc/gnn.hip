// Kernels of the learned planner's message-passing network (gfx950): bf16 operands,
// fp32 accumulation.
//
//  * k_gemm: C = A x B (+ bias) (ReLU) with arbitrary strides, so one kernel covers the
//    forward product X W and the backward products dX = dZ W^T and dW = X^T dZ.
//    64x64 block tiles, 4 waves each computing 32x32 with 2x2 v_mfma_f32_16x16x32_bf16,
//    K in steps of 32 staged through LDS with 16-byte loads along whichever index is
//    contiguous.  Output fp32 or bf16, optionally accumulated.
//  * k_segsum: out[r, off + f] = sum over incidences e of row r of |w[e]| X[col[e], f],
//    positive and negative literals (the sign of w[e]) into separate column blocks: the
//    normalised signed neighbour sums of the clause-variable incidence graph.  The edge
//    weights are symmetric (1/sqrt(deg(c) deg(v))), so the transpose incidence with the
//    same weights runs the backward pass.
//  * elementwise: bias-gradient column sums, ReLU masks, the squared-error loss and its
//    gradient, Adam updates on fp32 master weights with a bf16 copy.
//
// There is no counterpart in the reference (the north star adds the learned planner).

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "../../include/tensorsat_amd.h"

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;

__device__ __forceinline__ float bf2f(__bf16 x) { return (float)x; }
__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }

// Staging of one 64 x 32 operand tile into LDS, dst[row][k] = src[row * s_row + k * s_k]
// (zero outside rows < nrows, k < ke): 16-byte loads of 8 elements along the contiguous
// index when it is k (vec 1: one 16-byte LDS store) or the row (vec 2: eight 2-byte LDS
// stores, transposing), element loads otherwise and for ragged edges.
__device__ __forceinline__ int vec_mode(const __bf16* p, int64_t s_row, int64_t s_k) {
  if (reinterpret_cast<uintptr_t>(p) & 15) return 0;
  if (s_k == 1 && (s_row & 7) == 0) return 1;
  if (s_row == 1 && (s_k & 7) == 0) return 2;
  return 0;
}

__device__ __forceinline__ void stage_tile(__bf16 (*dst)[TK + 8], const __bf16* __restrict__ src, int64_t s_row,
                                           int64_t s_k, int row0, int nrows, int k0, int ke, int vec, int tid) {
  if (vec == 1) {  // 64 rows x 4 vectors of 8 k
    const int r = tid >> 2, c = (tid & 3) * 8;
    const int m = row0 + r, k = k0 + c;
    if (m < nrows && k + 8 <= ke) {
      *reinterpret_cast<uint4*>(&dst[r][c]) = *reinterpret_cast<const uint4*>(src + (int64_t)m * s_row + k);
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t)
        dst[r][c + t] = (m < nrows && k + t < ke) ? src[(int64_t)m * s_row + k + t] : f2bf(0.f);
    }
    return;
  }
  if (vec == 2) {  // 32 k x 8 vectors of 8 rows
    const int c = tid >> 3, r = (tid & 7) * 8;
    const int m = row0 + r, k = k0 + c;
    if (k < ke && m + 8 <= nrows) {
      const uint4 v = *reinterpret_cast<const uint4*>(src + m + (int64_t)k * s_k);
      const __bf16* e = reinterpret_cast<const __bf16*>(&v);
#pragma unroll
      for (int t = 0; t < 8; ++t) dst[r + t][c] = e[t];
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t)
        dst[r + t][c] = (k < ke && m + t < nrows) ? src[m + t + (int64_t)k * s_k] : f2bf(0.f);
    }
    return;
  }
  const bool kfast = s_k == 1;  // coalesce along the contiguous index
  for (int e = tid; e < 64 * TK; e += 256) {
    const int r = kfast ? e / TK : e % 64, c = kfast ? e % TK : e / 64;
    const int m = row0 + r, k = k0 + c;
    dst[r][c] = (m < nrows && k < ke) ? src[(int64_t)m * s_row + (int64_t)k * s_k] : f2bf(0.f);
  }
}

// The vector modes split into a global load (the thread's 8 elements, zero outside the
// bounds) and an LDS store, so that the next K tile's loads are in flight while the
// current tile's MFMAs run.
__device__ __forceinline__ uint4 load_vec(const __bf16* __restrict__ src, int64_t s_row, int64_t s_k, int row0,
                                          int nrows, int k0, int ke, int vec, int tid) {
  bf8 x;
  if (vec == 1) {
    const int r = tid >> 2, c = (tid & 3) * 8;
    const int m = row0 + r, k = k0 + c;
    if (m < nrows && k + 8 <= ke) return *reinterpret_cast<const uint4*>(src + (int64_t)m * s_row + k);
#pragma unroll
    for (int t = 0; t < 8; ++t) x[t] = (m < nrows && k + t < ke) ? src[(int64_t)m * s_row + k + t] : f2bf(0.f);
  } else {
    const int c = tid >> 3, r = (tid & 7) * 8;
    const int m = row0 + r, k = k0 + c;
    if (k < ke && m + 8 <= nrows) return *reinterpret_cast<const uint4*>(src + m + (int64_t)k * s_k);
#pragma unroll
    for (int t = 0; t < 8; ++t) x[t] = (k < ke && m + t < nrows) ? src[m + t + (int64_t)k * s_k] : f2bf(0.f);
  }
  return __builtin_bit_cast(uint4, x);
}

__device__ __forceinline__ void store_vec(__bf16 (*dst)[TK + 8], uint4 v, int vec, int tid) {
  if (vec == 1) {
    *reinterpret_cast<uint4*>(&dst[tid >> 2][(tid & 3) * 8]) = v;
    return;
  }
  const bf8 x = __builtin_bit_cast(bf8, v);
  const int c = tid >> 3, r = (tid & 7) * 8;
#pragma unroll
  for (int t = 0; t < 8; ++t) dst[r + t][c] = x[t];
}

// C[m, n] (+)= sum_k A(m, k) B(k, n); A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn],
// C at C[m*ldc + n].  Bounds are checked per element.
__global__ __launch_bounds__(256) void k_gemm(const __bf16* __restrict__ A, int64_t sam, int64_t sak,
                                              const __bf16* __restrict__ B, int64_t sbk, int64_t sbn,
                                              void* __restrict__ C, int64_t ldc, int M, int N, int K,
                                              const float* __restrict__ bias, int relu, int out_bf16,
                                              int accumulate, int kchunk) {
  __shared__ __bf16 sA[TM][TK + 8];
  __shared__ __bf16 sB[TN][TK + 8];  // stored transposed: sB[n][k]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // Split-K: block z covers k in [kb, ke); partial sums are added atomically (fp32 C).
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int vec_a = vec_mode(A, sam, sak), vec_b = vec_mode(B, sbn, sbk);
  f4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  // Lane l of a 16x16x32 tile holds row/col (l % 16) and k = 8*(l/16) .. +7 of both.
  const int rr = lane & 15, kk = (lane >> 4) * 8;
  auto mfma_tile = [&]() {
    bf8 a[2], b[2];
    for (int i = 0; i < 2; ++i)
      for (int t = 0; t < 8; ++t) a[i][t] = sA[wm + 16 * i + rr][kk + t];
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 8; ++t) b[j][t] = sB[wn + 16 * j + rr][kk + t];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  };
  if (vec_a && vec_b) {
    uint4 ra = load_vec(A, sam, sak, m0, M, kb, ke, vec_a, tid);
    uint4 rb = load_vec(B, sbn, sbk, n0, N, kb, ke, vec_b, tid);
    for (int k0 = kb; k0 < ke; k0 += TK) {
      store_vec(sA, ra, vec_a, tid);
      store_vec(sB, rb, vec_b, tid);
      __syncthreads();
      if (k0 + TK < ke) {
        ra = load_vec(A, sam, sak, m0, M, k0 + TK, ke, vec_a, tid);
        rb = load_vec(B, sbn, sbk, n0, N, k0 + TK, ke, vec_b, tid);
      }
      mfma_tile();
      __syncthreads();
    }
  } else {
    for (int k0 = kb; k0 < ke; k0 += TK) {
      stage_tile(sA, A, sam, sak, m0, M, k0, ke, vec_a, tid);
      stage_tile(sB, B, sbn, sbk, n0, N, k0, ke, vec_b, tid);
      __syncthreads();
      mfma_tile();
      __syncthreads();
    }
  }
  // Accumulator layout: lane l holds rows 4*(l/16) + t (t = 0..3), column l % 16.
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[i][j][t];
        if (gridDim.z > 1) {  // split-K partial sum (fp32 output, no bias/ReLU)
          atomicAdd(reinterpret_cast<float*>(C) + (int64_t)m * ldc + n, v);
          continue;
        }
        if (bias) v += bias[n];
        if (relu) v = v > 0.f ? v : 0.f;
        if (out_bf16) {
          __bf16* c = reinterpret_cast<__bf16*>(C) + (int64_t)m * ldc + n;
          *c = f2bf(accumulate ? bf2f(*c) + v : v);
        } else {
          float* c = reinterpret_cast<float*>(C) + (int64_t)m * ldc + n;
          *c = accumulate ? *c + v : v;
        }
      }
}

// One wave per row: lanes cover the F features (F <= 128, two per lane).  Positive and
// negative incidences go to column blocks off_pos / off_neg of the row.
__global__ __launch_bounds__(256) void k_segsum(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                const float* __restrict__ weight, int nrows,
                                                const void* __restrict__ X, int64_t ldx, int xoff_pos,
                                                int xoff_neg, int x_f32, void* __restrict__ out, int64_t ldo,
                                                int off_pos, int off_neg, int F, int out_f32_accumulate,
                                                int pairs) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= nrows) return;
  const int r = wave;
  float p0 = 0.f, p1 = 0.f, q0 = 0.f, q1 = 0.f;
  // pairs: lane l holds features 2l and 2l + 1 (one 4- or 8-byte load per incidence, a
  // wave reading one contiguous row); otherwise features l and l + 64.
  const int f0 = pairs ? 2 * lane : lane, f1 = pairs ? 2 * lane + 1 : lane + 64;
  auto load = [&](int e, float& v0, float& v1, float& w) {
    w = weight[e];
    const int64_t base = (int64_t)col[e] * ldx + (w > 0.f ? xoff_pos : xoff_neg);
    v0 = v1 = 0.f;
    if (x_f32) {
      const float* x = reinterpret_cast<const float*>(X) + base;
      if (pairs && f1 < F) {
        const float2 v = *reinterpret_cast<const float2*>(x + f0);
        v0 = v.x;
        v1 = v.y;
      } else {
        if (f0 < F) v0 = x[f0];
        if (f1 < F) v1 = x[f1];
      }
    } else {
      const __bf16* x = reinterpret_cast<const __bf16*>(X) + base;
      if (pairs && f1 < F) {
        const uint32_t u = *reinterpret_cast<const uint32_t*>(x + f0);
        v0 = bf2f(__builtin_bit_cast(__bf16, (uint16_t)(u & 0xFFFFu)));
        v1 = bf2f(__builtin_bit_cast(__bf16, (uint16_t)(u >> 16)));
      } else {
        if (f0 < F) v0 = bf2f(x[f0]);
        if (f1 < F) v1 = bf2f(x[f1]);
      }
    }
  };
  auto add = [&](float v0, float v1, float w) {
    const float aw = fabsf(w);
    if (w > 0.f) {
      p0 += aw * v0;
      p1 += aw * v1;
    } else {
      q0 += aw * v0;
      q1 += aw * v1;
    }
  };
  // Four incidences per trip, their rows loaded before any is added (four row gathers in
  // flight per wave instead of one).
  int e = rowptr[r];
  const int end = rowptr[r + 1];
  for (; e + 4 <= end; e += 4) {
    float v0[4], v1[4], w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) load(e + u, v0[u], v1[u], w[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) add(v0[u], v1[u], w[u]);
  }
  for (; e < end; ++e) {
    float v0, v1, w;
    load(e, v0, v1, w);
    add(v0, v1, w);
  }
  if (out_f32_accumulate) {
    float* o = reinterpret_cast<float*>(out) + (int64_t)r * ldo;
    if (pairs && f1 < F) {
      float2* d = reinterpret_cast<float2*>(o + off_pos + f0);
      const float2 v = *d;
      *d = make_float2(v.x + p0 + q0, v.y + p1 + q1);
    } else {
      if (f0 < F) o[off_pos + f0] += p0 + q0;
      if (f1 < F) o[off_pos + f1] += p1 + q1;
    }
  } else {
    __bf16* o = reinterpret_cast<__bf16*>(out) + (int64_t)r * ldo;
    if (pairs && f1 < F) {
      const auto pack = [](float a, float b) {
        return (uint32_t)__builtin_bit_cast(uint16_t, f2bf(a)) | ((uint32_t)__builtin_bit_cast(uint16_t, f2bf(b)) << 16);
      };
      *reinterpret_cast<uint32_t*>(o + off_pos + f0) = pack(p0, p1);
      *reinterpret_cast<uint32_t*>(o + off_neg + f0) = pack(q0, q1);
    } else {
      if (f0 < F) {
        o[off_pos + f0] = f2bf(p0);
        o[off_neg + f0] = f2bf(q0);
      }
      if (f1 < F) {
        o[off_pos + f1] = f2bf(p1);
        o[off_neg + f1] = f2bf(q1);
      }
    }
  }
}

// dZ = dH * (H > 0) (ReLU backward), written as bf16 for the next products and fp32
// column sums accumulated into db.
__global__ __launch_bounds__(256) void k_relu_back(const float* __restrict__ dH, int64_t ldd,
                                                   const __bf16* __restrict__ H, int64_t ldh, int rows, int F,
                                                   __bf16* __restrict__ dZ, int64_t ldz, float* __restrict__ db) {
  const int f = threadIdx.x % 128;
  const int rgroup = (blockIdx.x * blockDim.x + threadIdx.x) / 128;
  const int ngroups = (gridDim.x * blockDim.x) / 128;
  float s = 0.f;
  if (f < F) {
    for (int r = rgroup; r < rows; r += ngroups) {
      const float g = bf2f(H[(int64_t)r * ldh + f]) > 0.f ? dH[(int64_t)r * ldd + f] : 0.f;
      dZ[(int64_t)r * ldz + f] = f2bf(g);
      s += g;
    }
    if (db) atomicAdd(&db[f], s);
  }
}

// y[r] = sum_f H[r, f] w[f] + b ; loss = sum_r mask[r] (y[r]-t[r])^2 / count ;
// dy[r] = 2 mask[r] (y[r]-t[r]) / count ; dH[r, f] = dy[r] w[f] ; dw[f] += sum_r dy[r] H[r, f].
__global__ __launch_bounds__(128) void k_head(const __bf16* __restrict__ H, int64_t ldh, int rows,
                                              const float* __restrict__ w, const float* __restrict__ b,
                                              const float* __restrict__ target, const float* __restrict__ mask,
                                              float inv_count, float* __restrict__ y, float* __restrict__ loss,
                                              float* __restrict__ dH, float* __restrict__ dw, float* __restrict__ db,
                                              int F, int train) {
  __shared__ float red[128];
  const int f = threadIdx.x;
  float dw_acc = 0.f, db_acc = 0.f, loss_acc = 0.f;  // this block's rows, added once at the end
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const float h = f < F ? bf2f(H[(int64_t)r * ldh + f]) : 0.f;
    red[f] = h * (f < F ? w[f] : 0.f);
    __syncthreads();
    for (int s = 64; s > 0; s >>= 1) {
      if (f < s) red[f] += red[f + s];
      __syncthreads();
    }
    const float yr = red[0] + b[0];
    __syncthreads();
    if (f == 0) y[r] = yr;
    if (train) {
      const float m = mask ? mask[r] : 1.f;
      const float d = yr - target[r];
      const float g = 2.f * m * d * inv_count;
      loss_acc += m * d * d * inv_count;
      db_acc += g;
      if (f < F) {
        dH[(int64_t)r * F + f] = g * w[f];
        dw_acc += g * h;
      }
    }
  }
  if (train) {
    if (f == 0) {
      atomicAdd(loss, loss_acc);
      atomicAdd(db, db_acc);
    }
    if (f < F) atomicAdd(&dw[f], dw_acc);
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              __bf16* __restrict__ pbf, int64_t n, float lr, float b1, float b2,
                                              float eps, float c1, float c2, float wd) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] -= upd;
    if (pbf) pbf[i] = f2bf(p[i]);
  }
}

__global__ __launch_bounds__(256) void k_cast_bf16(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

inline int blocks_for(int64_t n, int per) { return (int)std::min<int64_t>((n + per - 1) / per, 65535); }

}  // namespace

extern "C" {

int ts_gnn_gemm(const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk, int64_t sbn, void* C,
                int64_t ldc, int32_t M, int32_t N, int32_t K, const float* bias, int32_t relu, int32_t out_bf16,
                int32_t accumulate, ts_stream stream) {
  if (!A || !B || !C || M < 0 || N < 0 || K < 0) return TS_EINVAL;
  if (M == 0 || N == 0) return TS_OK;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, 1);
  if (grid.y > 65535) return TS_EUNSUPPORTED;
  // Few output tiles over a long K (weight gradients X^T dZ): split K across blocks and
  // add fp32 partial sums into C, which the caller has zeroed (requires accumulate).
  int kchunk = K > 0 ? K : 1;
  const int tiles = (int)(grid.x * grid.y);
  if (tiles < 256 && K >= 4096 && !out_bf16 && accumulate && !bias && !relu) {
    int splits = std::min(256 / tiles + 1, (K + 2047) / 2048);
    kchunk = ((K + splits - 1) / splits + TK - 1) / TK * TK;
    grid.z = (K + kchunk - 1) / kchunk;
  }
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const __bf16*>(A), sam, sak, reinterpret_cast<const __bf16*>(B), sbk, sbn, C,
                     ldc, M, N, K, bias, relu, out_bf16, accumulate, kchunk);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_gnn_segsum(const int32_t* rowptr, const int32_t* col, const float* weight, int32_t nrows, const void* X,
                  int64_t ldx, int32_t xoff_pos, int32_t xoff_neg, int32_t x_f32, void* out, int64_t ldo,
                  int32_t off_pos, int32_t off_neg, int32_t F, int32_t out_f32_accumulate, ts_stream stream) {
  if (nrows < 0 || F < 0 || F > 128) return TS_EINVAL;
  if (nrows == 0) return TS_OK;
  if (!rowptr || !col || !weight || !X || !out) return TS_EINVAL;
  const int threads = 256;
  const int64_t waves = nrows;
  // Feature pairs need every pair's first element at an even element offset (8- or
  // 4-byte aligned pair loads and stores).
  const int pairs = ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(out)) & 7) == 0 &&
                    ((ldx | xoff_pos | xoff_neg | ldo | off_pos | off_neg) & 1) == 0;
  hipLaunchKernelGGL(k_segsum, dim3(blocks_for(waves * 64, threads)), dim3(threads), 0,
                     reinterpret_cast<hipStream_t>(stream), rowptr, col, weight, nrows, X, ldx, xoff_pos, xoff_neg,
                     x_f32, out, ldo, off_pos, off_neg, F, out_f32_accumulate, pairs);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_gnn_relu_back(const float* dH, int64_t ldd, const void* H, int64_t ldh, int32_t rows, int32_t F, void* dZ,
                     int64_t ldz, float* db, ts_stream stream) {
  if (!dH || !H || !dZ || rows < 0 || F < 0 || F > 128) return TS_EINVAL;
  if (rows == 0) return TS_OK;
  const int grid = std::min(1024, std::max(1, rows / 16));
  hipLaunchKernelGGL(k_relu_back, dim3(grid), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), dH, ldd,
                     reinterpret_cast<const __bf16*>(H), ldh, rows, F, reinterpret_cast<__bf16*>(dZ), ldz, db);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_gnn_head(const void* H, int64_t ldh, int32_t rows, const float* w, const float* b, const float* target,
                const float* mask, float inv_count, float* y, float* loss, float* dH, float* dw, float* db,
                int32_t F, int32_t train, ts_stream stream) {
  if (!H || !w || !b || !y || rows < 0 || F < 0 || F > 128) return TS_EINVAL;
  if (train && (!target || !loss || !dH || !dw || !db)) return TS_EINVAL;
  if (rows == 0) return TS_OK;
  hipLaunchKernelGGL(k_head, dim3(std::min(rows, 512)), dim3(128), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const __bf16*>(H), ldh, rows, w, b, target, mask, inv_count, y, loss, dH, dw,
                     db, F, train);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_gnn_adam(float* p, const float* g, float* m, float* v, void* pbf, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, float wd, ts_stream stream) {
  if (!p || !g || !m || !v || n < 0 || step < 1) return TS_EINVAL;
  if (n == 0) return TS_OK;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p, g, m,
                     v, reinterpret_cast<__bf16*>(pbf), n, lr, b1, b2, eps, c1, c2, wd);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

int ts_gnn_cast_bf16(const float* x, void* y, int64_t n, ts_stream stream) {
  if (!x || !y || n < 0) return TS_EINVAL;
  if (n == 0) return TS_OK;
  hipLaunchKernelGGL(k_cast_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), x,
                     reinterpret_cast<__bf16*>(y), n);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // extern "C"
