// HIP kernels of libstile_hip.so for MI355X (gfx950, CDNA4, wave64).
//
// Four kernel templates:
//   csr_spmm_f32     plain CSR SpMM (the comparator; reference csrmm template,
//                    MY_sparse/my_fuse_formats.py:1106-1152)
//   csr_sddmm_f32    plain per-nonzero SDDMM (reference golden u_dot_v,
//                    my_fuse_formats.py:4331-4341)
//   hyb_spmm         hybrid fused SpMM: MFMA on dense 16-row blocks + FP32
//                    sparse part, one wave per (16-row window, 64 columns),
//                    every C row written once (replaces the fused ELL +
//                    TensorCore templates, my_fuse_formats.py:1162-1291,
//                    1444-1799, 2426-2465)
//   hyb_sddmm        hybrid fused SDDMM: 1D entries as FP32 wave-segment dot
//                    products + dense 16x16 blocks on MFMA with sampled
//                    scatter (my_fuse_formats.py:2263-2317, 2324-2415,
//                    2519-2560)
#include "stile_internal.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cstdlib>
#include <cstring>

namespace stile {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float4 ld4(const float* p) {
  return *reinterpret_cast<const float4*>(p);
}
__device__ __forceinline__ void st4(float* p, float4 v) {
  *reinterpret_cast<float4*>(p) = v;
}
__device__ __forceinline__ float comp(const float4& v, int t) {
  return t == 0 ? v.x : (t == 1 ? v.y : (t == 2 ? v.z : v.w));
}

// Neumaier compensated accumulation: s + c carries the sum of every x added.
// Contraction is off so (s - t) + x is evaluated as written.
__device__ __forceinline__ void neu_add(float& s, float& c, float x) {
#pragma clang fp contract(off)
  float t = s + x;
  if (fabsf(s) >= fabsf(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

// Elements of B: FP32, BF16 or FP16.  load_b4 returns four consecutive
// columns of row k starting at column col as FP32 (exact for the 16-bit
// types); columns at or beyond N read as 0.  vec: B rows are aligned for one
// vector load of four elements (16 bytes FP32, 8 bytes 16-bit) and N % 4 == 0.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

template <typename TB>
__device__ __forceinline__ float4 cvt4(uint2 w) {
  if constexpr (std::is_same<TB, __bf16>::value) {
    return make_float4(bf16_lo(w.x), bf16_hi(w.x), bf16_lo(w.y), bf16_hi(w.y));
  } else {
    const f16x2 a = __builtin_bit_cast(f16x2, w.x), b = __builtin_bit_cast(f16x2, w.y);
    return make_float4((float)a[0], (float)a[1], (float)b[0], (float)b[1]);
  }
}

// B[k][col .. col + 3] as float4, zero at or beyond N.  vec (N % 4 == 0,
// rows aligned): one unconditional load from a column clamped into the row
// and a select after it.  (A load under an exec mask leaves its registers
// partly unwritten, and the compiler then waits for every load in flight
// before it reuses them: in a loop of gathers that serialises the gathers.)
template <typename TB>
__device__ __forceinline__ float4 load_b4(const TB* __restrict__ B, int64_t ldb, int k, int col, int N, bool vec) {
  const TB* p = B + (int64_t)k * ldb + col;
  if (vec) {
    const TB* q = B + (int64_t)k * ldb + min(col, N - 4);
    float4 r;
    if constexpr (std::is_same<TB, float>::value)
      r = ld4(q);
    else
      r = cvt4<TB>(*reinterpret_cast<const uint2*>(q));
    return col < N ? r : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  float4 r;
  r.x = col < N ? (float)p[0] : 0.f;
  r.y = col + 1 < N ? (float)p[1] : 0.f;
  r.z = col + 2 < N ? (float)p[2] : 0.f;
  r.w = col + 3 < N ? (float)p[3] : 0.f;
  return r;
}

// f(std::true_type) when vec, else f(std::false_type): the row gathers of
// the sparse kernels see vec as a constant, so the 16-byte and the
// per-element paths of load_b4 are not merged into one loop, where the
// per-element path's masked loads make every gather wait for the ones
// before it.
template <typename F>
__device__ __forceinline__ void with_vec(bool vec, F&& f) {
  if (vec)
    f(std::true_type{});
  else
    f(std::false_type{});
}

// NT: non-temporal store (streamed past the caches; for rows nothing reads
// back soon).
template <bool NT = false>
__device__ __forceinline__ void store_c4(float* __restrict__ C, int64_t ldc,
                                         int row, int col, int N, bool vec,
                                         float4 v) {
  float* p = C + (int64_t)row * ldc + col;
  if (vec) {
    if (col < N) {
      if constexpr (NT) {
        typedef float f32v4 __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(f32v4{v.x, v.y, v.z, v.w}, reinterpret_cast<f32v4*>(p));
      } else {
        st4(p, v);
      }
    }
    return;
  }
  if (col < N) p[0] = v.x;
  if (col + 1 < N) p[1] = v.y;
  if (col + 2 < N) p[2] = v.z;
  if (col + 3 < N) p[3] = v.w;
}

// ---------------------------------------------------------------------------
// Plain CSR SpMM.  A segment of LPR lanes owns one row and CPL groups of
// 4*LPR consecutive columns (one float4 per lane per group); a wave holds
// 64/LPR rows.  Grid: (row groups) x (column blocks of 4*LPR*CPL).
// ---------------------------------------------------------------------------
template <int LPR, int CPL>
__global__ __launch_bounds__(256) void csr_spmm_f32_kernel(
    int M, int N, const int32_t* __restrict__ indptr,
    const int32_t* __restrict__ indices, const float* __restrict__ vals,
    const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
    int64_t ldc, bool vec) {
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63;
  const int sub = lane % LPR;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int row = wave * RPW + lane / LPR;
  if (row >= M) return;
  const int col0 = blockIdx.y * (4 * LPR * CPL) + 4 * sub;
  const int p0 = indptr[row], p1 = indptr[row + 1];

  float4 acc[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) acc[q] = make_float4(0.f, 0.f, 0.f, 0.f);

  if (p1 - p0 <= kHubNnz) {
    int p = p0;
    for (; p + 4 <= p1; p += 4) {
      const int k0 = indices[p], k1 = indices[p + 1], k2 = indices[p + 2],
                k3 = indices[p + 3];
      const float v0 = vals[p], v1 = vals[p + 1], v2 = vals[p + 2],
                  v3 = vals[p + 3];
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const int col = col0 + q * 4 * LPR;
        const float4 b0 = load_b4(B, ldb, k0, col, N, vec);
        const float4 b1 = load_b4(B, ldb, k1, col, N, vec);
        const float4 b2 = load_b4(B, ldb, k2, col, N, vec);
        const float4 b3 = load_b4(B, ldb, k3, col, N, vec);
        acc[q].x += v0 * b0.x; acc[q].y += v0 * b0.y; acc[q].z += v0 * b0.z; acc[q].w += v0 * b0.w;
        acc[q].x += v1 * b1.x; acc[q].y += v1 * b1.y; acc[q].z += v1 * b1.z; acc[q].w += v1 * b1.w;
        acc[q].x += v2 * b2.x; acc[q].y += v2 * b2.y; acc[q].z += v2 * b2.z; acc[q].w += v2 * b2.w;
        acc[q].x += v3 * b3.x; acc[q].y += v3 * b3.y; acc[q].z += v3 * b3.z; acc[q].w += v3 * b3.w;
      }
    }
    for (; p < p1; ++p) {
      const int k = indices[p];
      const float v = vals[p];
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const float4 b = load_b4(B, ldb, k, col0 + q * 4 * LPR, N, vec);
        acc[q].x += v * b.x; acc[q].y += v * b.y; acc[q].z += v * b.z; acc[q].w += v * b.w;
      }
    }
  } else {
    // Hub row: compensated FP32 summation.
    float4 cc[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) cc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = p0; p < p1; ++p) {
      const int k = indices[p];
      const float v = vals[p];
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const float4 b = load_b4(B, ldb, k, col0 + q * 4 * LPR, N, vec);
        neu_add(acc[q].x, cc[q].x, v * b.x);
        neu_add(acc[q].y, cc[q].y, v * b.y);
        neu_add(acc[q].z, cc[q].z, v * b.z);
        neu_add(acc[q].w, cc[q].w, v * b.w);
      }
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      acc[q].x += cc[q].x; acc[q].y += cc[q].y; acc[q].z += cc[q].z; acc[q].w += cc[q].w;
    }
  }
#pragma unroll
  for (int q = 0; q < CPL; ++q)
    store_c4(C, ldc, row, col0 + q * 4 * LPR, N, vec, acc[q]);
}

template <int LPR, int CPL>
static hipError_t launch_csr_spmm_cfg(int M, int N, const int32_t* indptr,
                                      const int32_t* indices,
                                      const float* vals, const float* B,
                                      int64_t ldb, float* C, int64_t ldc,
                                      bool vec, hipStream_t s) {
  constexpr int RPW = 64 / LPR;
  const int waves = (M + RPW - 1) / RPW;
  const int cb = 4 * LPR * CPL;
  dim3 grid((waves + 3) / 4, (N + cb - 1) / cb);
  hipLaunchKernelGGL((csr_spmm_f32_kernel<LPR, CPL>), grid, dim3(256), 0, s,
                     M, N, indptr, indices, vals, B, ldb, C, ldc, vec);
  return hipGetLastError();
}

hipError_t launch_csr_spmm_f32(int M, int N, const int32_t* indptr,
                               const int32_t* indices, const float* vals,
                               const float* B, int64_t ldb, float* C,
                               int64_t ldc, hipStream_t s) {
  if (M == 0 || N == 0) return hipSuccess;
  const bool vec = (N % 4 == 0) && (ldb % 4 == 0) && (ldc % 4 == 0) &&
                   ((reinterpret_cast<uintptr_t>(B) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(C) & 15) == 0);
  if (N <= 16)
    return launch_csr_spmm_cfg<4, 1>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
  if (N <= 32)
    return launch_csr_spmm_cfg<8, 1>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
  if (N <= 64)
    return launch_csr_spmm_cfg<16, 1>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
  if (N <= 128)
    return launch_csr_spmm_cfg<32, 1>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
  if (N <= 256)
    return launch_csr_spmm_cfg<64, 1>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
  return launch_csr_spmm_cfg<64, 2>(M, N, indptr, indices, vals, B, ldb, C, ldc, vec, s);
}

// ---------------------------------------------------------------------------
// Plain CSR SDDMM: a segment of LPE lanes owns one nonzero; each lane loads
// float4 slices of X[row] and Y[col] and the segment reduces with xor
// shuffles.  The row of each nonzero is found by binary search in indptr.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_row(const int32_t* __restrict__ indptr,
                                        int M, int p) {
  int lo = 0, hi = M;  // indptr[lo] <= p < indptr[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (indptr[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

template <int LPE>
__device__ __forceinline__ float segment_dot(const float* __restrict__ x,
                                             const float* __restrict__ y,
                                             int Kf, int sub, bool vec) {
  float s = 0.f;
  if (vec) {
    for (int k = 4 * sub; k < Kf; k += 4 * LPE) {
      const float4 a = ld4(x + k), b = ld4(y + k);
      s += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    }
  } else {
    for (int k = sub; k < Kf; k += LPE) s += x[k] * y[k];
  }
#pragma unroll
  for (int off = LPE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPE);
  return s;
}

template <int LPE>
__global__ __launch_bounds__(256) void csr_sddmm_f32_kernel(
    int M, int nnz, int Kf, const int32_t* __restrict__ indptr,
    const int32_t* __restrict__ indices, const float* __restrict__ X,
    int64_t ldx, const float* __restrict__ Y, int64_t ldy,
    float* __restrict__ out, bool vec) {
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int sub = lane % LPE;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int p = wave * EPW + lane / LPE;
  if (p >= nnz) return;
  const int row = find_row(indptr, M, p);
  const int col = indices[p];
  const float s = segment_dot<LPE>(X + (int64_t)row * ldx,
                                   Y + (int64_t)col * ldy, Kf, sub, vec);
  if (sub == 0) out[p] = s;
}

hipError_t launch_csr_sddmm_f32(int M, int nnz, int Kf, const int32_t* indptr,
                                const int32_t* indices, const float* X,
                                int64_t ldx, const float* Y, int64_t ldy,
                                float* out, hipStream_t s) {
  if (nnz == 0) return hipSuccess;
  const bool vec = (Kf % 4 == 0) && (ldx % 4 == 0) && (ldy % 4 == 0) &&
                   ((reinterpret_cast<uintptr_t>(X) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(Y) & 15) == 0);
  int lpe = vec ? Kf / 4 : Kf;
  if (lpe > 16) lpe = 16;
  int l = 1;
  while (l * 2 <= lpe) l *= 2;
  lpe = l;
  const int epw = 64 / lpe;
  const int waves = (nnz + epw - 1) / epw;
  dim3 grid((waves + 3) / 4);
#define STILE_SDDMM_CASE(L)                                                  \
  case L:                                                                    \
    hipLaunchKernelGGL(csr_sddmm_f32_kernel<L>, grid, dim3(256), 0, s, M,    \
                       nnz, Kf, indptr, indices, X, ldx, Y, ldy, out, vec); \
    break;
  switch (lpe) {
    STILE_SDDMM_CASE(1)
    STILE_SDDMM_CASE(2)
    STILE_SDDMM_CASE(4)
    STILE_SDDMM_CASE(8)
    STILE_SDDMM_CASE(16)
  }
#undef STILE_SDDMM_CASE
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Hybrid fused SpMM.
//
// One wave = (window w, 64 output columns n0..n0+63).  Lane l = 16g + j owns
// output columns n0 + 4j .. n0 + 4j + 3 of rows 4g .. 4g+3 of the window: the
// four 16x16 MFMA tiles t = 0..3 of the wave hold column n0 + 4j + t in MFMA
// column j, so every B load and every C store is four consecutive columns and
// MFMA output element (row 4g+i, column j) of tile t is exactly the lane's own
// element.
//
// Dense chunk (32 k-slots): k-slot 4e + g of lane l is column tc_cols[c][g][e]
// of A, A operand element e of lane l (the pairing of k-slots with operand
// elements is the same for A and B, which is all the MFMA needs).
//   BF16: one v_mfma_f32_16x16x32_bf16 per tile;
//   FP16: one v_mfma_f32_16x16x32_f16 per tile, times the chunk's scale;
//   FP32: eight v_mfma_f32_16x16x4_f32 per tile (MFMA e reduces k-slots
//         4e + 0..3; lane group g supplies k-slot 4e + g).
// Sparse part: the lane adds v * B[col][n0+4j..+3] for the FP32 nonzeros of
// its rows (compensated when a row has more than kHubNnz).
//
// B is FP32, BF16 or FP16 (TB); every product and sum of the sparse part is
// FP32 on B's exact values.
//
// Kernels of one run, in stream order.  Plans on the B-stationary kernel
// (BF16 blocks, K <= kLdsMaxK): that kernel computes every dense window with
// its sparse part, and as its tail (when they are few) the sparse-only rows;
// otherwise the sparse kernel runs the sparse-only rows.  Other plans: the
// streaming kernel stores the dense windows (a workgroup for windows of more
// than kBigChunks chunks), the sparse kernel adds the first kSplitNnz sparse
// nonzeros of those windows' rows and writes the sparse-only windows.  Last,
// the pieces of rows longer than kSplitNnz: added with atomicAdd, or, in the
// deterministic mode, summed into a workspace and added in piece order.
// ---------------------------------------------------------------------------

// Gathers in flight per lane group in the sparse and mixed slots.
constexpr int kSparseDepth = 8;

// Sum of v * B[col][n0+4j..+3] over nonzeros [r0, r1) of the sparse part.
// DEPTH 8 keeps eight gathers in flight with the next ids prefetched; DEPTH 4
// (four at a time) is for kernels whose register budget is tight.
template <int DEPTH, typename TB>
__device__ __forceinline__ float4 sparse_row_sum(const SpmmPlanDev& p, const TB* __restrict__ B,
                                                 int64_t ldb, int N, int col, bool vec, int r0, int r1,
                                                 bool compensated) {
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!compensated) {
    int q = r0;
    // Eight gathers in flight, the next eight column ids fetched under them.
    if (DEPTH >= 8 && q + 8 <= r1) {
      int k[8];
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        k[i] = p.res_cols[q + i];
        v[i] = p.res_vals[q + i];
      }
      while (true) {
        float4 b[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) b[i] = load_b4(B, ldb, k[i], col, N, vec);
        const bool more = q + 16 <= r1;
        int kn[8];
        float vn[8];
        if (more) {
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            kn[i] = p.res_cols[q + 8 + i];
            vn[i] = p.res_vals[q + 8 + i];
          }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          s.x += v[i] * b[i].x;
          s.y += v[i] * b[i].y;
          s.z += v[i] * b[i].z;
          s.w += v[i] * b[i].w;
        }
        q += 8;
        if (!more) break;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          k[i] = kn[i];
          v[i] = vn[i];
        }
      }
    }
    for (; q + 4 <= r1; q += 4) {
      const int k0 = p.res_cols[q], k1 = p.res_cols[q + 1], k2 = p.res_cols[q + 2], k3 = p.res_cols[q + 3];
      const float v0 = p.res_vals[q], v1 = p.res_vals[q + 1], v2 = p.res_vals[q + 2], v3 = p.res_vals[q + 3];
      const float4 b0 = load_b4(B, ldb, k0, col, N, vec);
      const float4 b1 = load_b4(B, ldb, k1, col, N, vec);
      const float4 b2 = load_b4(B, ldb, k2, col, N, vec);
      const float4 b3 = load_b4(B, ldb, k3, col, N, vec);
      s.x += v0 * b0.x; s.y += v0 * b0.y; s.z += v0 * b0.z; s.w += v0 * b0.w;
      s.x += v1 * b1.x; s.y += v1 * b1.y; s.z += v1 * b1.z; s.w += v1 * b1.w;
      s.x += v2 * b2.x; s.y += v2 * b2.y; s.z += v2 * b2.z; s.w += v2 * b2.w;
      s.x += v3 * b3.x; s.y += v3 * b3.y; s.z += v3 * b3.z; s.w += v3 * b3.w;
    }
    for (; q < r1; ++q) {
      const int k0 = p.res_cols[q];
      const float v0 = p.res_vals[q];
      const float4 b0 = load_b4(B, ldb, k0, col, N, vec);
      s.x += v0 * b0.x; s.y += v0 * b0.y; s.z += v0 * b0.z; s.w += v0 * b0.w;
    }
  } else {
    float4 cc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int q = r0; q < r1; ++q) {
      const int k0 = p.res_cols[q];
      const float v0 = p.res_vals[q];
      const float4 b0 = load_b4(B, ldb, k0, col, N, vec);
      neu_add(s.x, cc.x, v0 * b0.x);
      neu_add(s.y, cc.y, v0 * b0.y);
      neu_add(s.z, cc.z, v0 * b0.z);
      neu_add(s.w, cc.w, v0 * b0.w);
    }
    s.x += cc.x; s.y += cc.y; s.z += cc.z; s.w += cc.w;
  }
  return s;
}

// Output rows of the 4 slots a lane group stores.  Loaded before any of the
// window's C stores: vector-memory counters complete in issue order on
// CDNA, so a load issued after a store would wait for it.
struct SlotRows {
  int row[4];
};

__device__ __forceinline__ SlotRows load_rows(const SpmmPlanDev& p, int slot0) {
  SlotRows sr;
#pragma unroll
  for (int i = 0; i < 4; ++i) sr.row[i] = p.win_rows[slot0 + i];
  return sr;
}

// Store the dense accumulators of the lane group's 4 rows (acc[t][i] is row
// 4g+i, column col+t).
template <bool NT = false>
__device__ __forceinline__ void store_window(float* __restrict__ C, int64_t ldc, int N, int col, bool vec,
                                             const SlotRows& sr, const f32x4 (&acc)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (sr.row[i] >= 0)
      store_c4<NT>(C, ldc, sr.row[i], col, N, vec, make_float4(acc[0][i], acc[1][i], acc[2][i], acc[3][i]));
}

// Slot of a dense window with a sparse part: C[row] += first kSplitNnz
// nonzeros of the part (the streaming dense kernel stored C[row] before).
template <typename TB>
__device__ __forceinline__ void add_slot(const SpmmPlanDev& p, const TB* __restrict__ B, int64_t ldb, int N,
                                         float* __restrict__ C, int64_t ldc, int col, bool vec, int slot) {
  const int row = p.win_rows[slot];
  const int r0 = p.res_ptr[slot], rend = p.res_ptr[slot + 1];
  const int r1 = min(rend, r0 + kSplitNnz);
  const float4 s = sparse_row_sum<kSparseDepth>(p, B, ldb, N, col, vec, r0, r1, rend - r0 > kHubNnz);
  const float4 c = load_b4(C, ldc, row, col, N, vec);
  store_c4(C, ldc, row, col, N, vec, make_float4(c.x + s.x, c.y + s.y, c.z + s.z, c.w + s.w));
}

// One slot (sparse-only windows, one row per lane group).
template <typename TB>
__device__ __forceinline__ void finish_slot(const SpmmPlanDev& p, const TB* __restrict__ B, int64_t ldb,
                                            int N, float* __restrict__ C, int64_t ldc, int col, bool vec,
                                            int slot) {
  const int row = p.win_rows[slot];
  const int r0 = p.res_ptr[slot], rend = p.res_ptr[slot + 1];
  const int r1 = min(rend, r0 + kSplitNnz);
  const float4 s = sparse_row_sum<kSparseDepth>(p, B, ldb, N, col, vec, r0, r1, rend - r0 > kHubNnz);
  if (row >= 0) store_c4(C, ldc, row, col, N, vec, s);
}

template <int MODE>
__device__ __forceinline__ void mfma_chunk(const SpmmPlanDev& p, int64_t c, int lane, const float4 (&b)[8],
                                           f32x4 (&acc)[4]) {
  if constexpr (MODE == 1) {
    const bf16x8 a = reinterpret_cast<const bf16x8*>(p.tc_vals)[c * 64 + lane];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x8 bt;
#pragma unroll
      for (int e = 0; e < 8; ++e) bt[e] = (__bf16)comp(b[e], t);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bt, acc[t], 0, 0, 0);
    }
  } else if constexpr (MODE == 2) {
    const f16x8 a = reinterpret_cast<const f16x8*>(p.tc_vals)[c * 64 + lane];
    const float sc = p.tc_scale[c];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f16x8 bt;
#pragma unroll
      for (int e = 0; e < 8; ++e) bt[e] = (_Float16)comp(b[e], t);
      f32x4 tmp = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bt, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      acc[t] += sc * tmp;
    }
  } else {
    const float* ap = reinterpret_cast<const float*>(p.tc_vals) + (c * 64 + lane) * 8;
    const float4 a0 = ld4(ap), a1 = ld4(ap + 4);
    const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], comp(b[e], t), acc[t], 0, 0, 0);
    }
  }
}

// B rows of one chunk for lane group g: 8 loads of four columns (addresses
// from the chunk's column ids, already in registers).
template <typename TB>
__device__ __forceinline__ void load_chunk_b(const TB* __restrict__ B, int64_t ldb, int col, int N, bool vec,
                                             const int4& ka, const int4& kb, float4 (&b)[8]) {
  b[0] = load_b4(B, ldb, ka.x, col, N, vec);
  b[1] = load_b4(B, ldb, ka.y, col, N, vec);
  b[2] = load_b4(B, ldb, ka.z, col, N, vec);
  b[3] = load_b4(B, ldb, ka.w, col, N, vec);
  b[4] = load_b4(B, ldb, kb.x, col, N, vec);
  b[5] = load_b4(B, ldb, kb.y, col, N, vec);
  b[6] = load_b4(B, ldb, kb.z, col, N, vec);
  b[7] = load_b4(B, ldb, kb.w, col, N, vec);
}

// Dense part of window w for this lane's columns.  Two chunks per step (16
// independent B loads in flight) and the column ids of the next two chunks
// fetched before the current step's B loads, so a step waits for one memory
// latency instead of two.
template <int MODE, typename TB>
__device__ __forceinline__ void window_dense(const SpmmPlanDev& p, const TB* __restrict__ B, int64_t ldb,
                                             int N, int col, bool vec, int c0, int c1, int lane,
                                             f32x4 (&acc)[4]) {
  const int g = lane >> 4;
  if (c0 >= c1) return;
  auto cols_of = [&](int c, int4& ka, int4& kb) {
    const int4* cp = reinterpret_cast<const int4*>(p.tc_cols + (int64_t)c * 32 + g * 8);
    ka = cp[0];
    kb = cp[1];
  };
  int4 ka0, kb0, ka1, kb1;
  cols_of(c0, ka0, kb0);
  if (c0 + 1 < c1) cols_of(c0 + 1, ka1, kb1); else { ka1 = ka0; kb1 = kb0; }
  int c = c0;
  for (; c + 1 < c1; c += 2) {
    int4 na0 = ka0, nb0 = kb0, na1 = ka1, nb1 = kb1;
    if (c + 2 < c1) cols_of(c + 2, na0, nb0);
    if (c + 3 < c1) cols_of(c + 3, na1, nb1);
    float4 b0[8], b1[8];
    load_chunk_b(B, ldb, col, N, vec, ka0, kb0, b0);
    load_chunk_b(B, ldb, col, N, vec, ka1, kb1, b1);
    mfma_chunk<MODE>(p, c, lane, b0, acc);
    mfma_chunk<MODE>(p, c + 1, lane, b1, acc);
    ka0 = na0; kb0 = nb0; ka1 = na1; kb1 = nb1;
  }
  if (c < c1) {
    float4 b0[8];
    load_chunk_b(B, ldb, col, N, vec, ka0, kb0, b0);
    mfma_chunk<MODE>(p, c, lane, b0, acc);
  }
}

template <int MODE, typename TB>
__device__ __forceinline__ void hyb_spmm_window(const SpmmPlanDev& p, const TB* __restrict__ B, int64_t ldb,
                                                int N, float* __restrict__ C, int64_t ldc, int w, int n0,
                                                bool vec) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int j = lane & 15;
  const int col = n0 + 4 * j;

  const SlotRows sr = load_rows(p, w * 16 + g * 4);
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  window_dense<MODE>(p, B, ldb, N, col, vec, p.tc_ptr[w], p.tc_ptr[w + 1], lane, acc);
  store_window(C, ldc, N, col, vec, sr, acc);
}

// Blocks [0, n_parts * nC) each run one (part of a huge window, 64 columns):
// they add their partial rows into C with atomicAdd (C rows zeroed before),
// or, with ws (deterministic mode), store them to ws[part][16][N] for
// hyb_spmm_part_reduce_kernel.  The next big_blocks each run one (big window,
// 64 columns): wave v takes the v-th contiguous share of the window's chunks,
// waves 1.. leave their partial sums in LDS and wave 0 adds them in wave
// order and stores the 16 rows.  The other blocks run kBigWaves waves each
// over the remaining dense windows [W_big, W_tc), one wave per window and 64
// columns.
template <int MODE, typename TB>
__global__ __launch_bounds__(kBigWaves * 64) void hyb_spmm_kernel(SpmmPlanDev p, const TB* __restrict__ B,
                                                                  int64_t ldb, int N, float* __restrict__ C,
                                                                  int64_t ldc, int nC, int big_blocks,
                                                                  bool vec, float* __restrict__ ws) {
  __shared__ f32x4 part[kBigWaves - 1][4][64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int huge_blocks = p.n_parts * nC;
  const bool in_huge = (int)blockIdx.x < huge_blocks;
  if (in_huge || (int)blockIdx.x < huge_blocks + big_blocks) {
    int w, col, c0, c1, pt = 0;
    if (in_huge) {
      pt = blockIdx.x / nC;
      w = p.part_win[pt];
      col = (blockIdx.x - pt * nC) * 64 + 4 * (lane & 15);
      c0 = p.part_c0[pt];
      c1 = p.part_c1[pt];
    } else {
      w = p.W_huge + (int)(blockIdx.x - huge_blocks) / nC;
      col = (int)(blockIdx.x - huge_blocks - (w - p.W_huge) * nC) * 64 + 4 * (lane & 15);
      c0 = p.tc_ptr[w];
      c1 = p.tc_ptr[w + 1];
    }
    const int share = (c1 - c0 + kBigWaves - 1) / kBigWaves;
    const int a = min(c0 + wv * share, c1), b = min(a + share, c1);
    SlotRows sr;
    if (wv == 0) sr = load_rows(p, w * 16 + (lane >> 4) * 4);
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    with_vec(vec, [&](auto V) { window_dense<MODE>(p, B, ldb, N, col, V, a, b, lane, acc); });
    if (wv > 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t) part[wv - 1][t][lane] = acc[t];
    }
    __syncthreads();
    if (wv != 0) return;
    for (int v = 0; v < kBigWaves - 1; ++v) {
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] += part[v][t][lane];
    }
    if (!in_huge) {
      store_window(C, ldc, N, col, vec, sr, acc);
      return;
    }
    if (ws) {
      // partial rows of the part: ws[pt][slot 4g+i][col..col+3]
      const int g = lane >> 4;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        store_c4(ws + (int64_t)pt * 16 * N, N, 4 * g + i, col, N, vec,
                 make_float4(acc[0][i], acc[1][i], acc[2][i], acc[3][i]));
      return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (sr.row[i] < 0) continue;
      float* cp = C + (int64_t)sr.row[i] * ldc + col;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (col + t < N) atomicAdd(cp + t, acc[t][i]);
    }
    return;
  }
  const int64_t wave = (int64_t)(blockIdx.x - huge_blocks - big_blocks) * kBigWaves + wv;
  if (wave >= (int64_t)(p.W_tc - p.W_big) * nC) return;
  const int w = p.W_big + (int)(wave / nC);
  const int n0 = (int)(wave - (int64_t)(w - p.W_big) * nC) * 64;
  with_vec(vec, [&](auto V) { hyb_spmm_window<MODE>(p, B, ldb, N, C, ldc, w, n0, V); });
}

// Deterministic mode: C rows of huge window w = sum of its parts' partial
// rows in part order.  One wave per (huge window, 64 columns).
__global__ __launch_bounds__(256) void hyb_spmm_part_reduce_kernel(SpmmPlanDev p, int N, float* __restrict__ C,
                                                                   int64_t ldc, int nC, bool vec,
                                                                   const float* __restrict__ ws) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= (int64_t)p.W_huge * nC) return;
  const int w = (int)(wave / nC);
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int col = (int)(wave - (int64_t)w * nC) * 64 + 4 * (lane & 15);
  const SlotRows sr = load_rows(p, w * 16 + g * 4);
  const int q0 = p.hw_p0[w], q1 = p.hw_p0[w + 1];
  float4 acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int q = q0; q < q1; ++q) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 v = load_b4(ws + (int64_t)q * 16 * N, N, 4 * g + i, col, N, vec);
      acc[i].x += v.x; acc[i].y += v.y; acc[i].z += v.z; acc[i].w += v.w;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (sr.row[i] >= 0) store_c4(C, ldc, sr.row[i], col, N, vec, acc[i]);
}

// Sparse parts, one row per lane group: first the n_mix slots of dense
// windows (four per wave and 64 columns, added to the stored dense rows;
// n_mix is 0 when the B-stationary kernel adds them itself), then the
// sparse-only windows [W_tc, W) (four waves per window and 64 columns).  A
// kernel of its own so that its small register footprint sets its occupancy
// (gathers in flight per CU), not the dense path's.
// so_waves: waves for the sparse-only windows (0 when hyb_spmm_sparse_split_kernel
// stores those rows instead).  Waves go column chunk by column chunk (all of a
// part's waves for columns 0..63, then 64..127, ...), so that the B rows the
// grid gathers at one time are one 64-column slice of B: on a 1 M-row graph
// with N = 128 that slice is 256 MB instead of B's 512 MB (GCN scale-20 SpMM
// 2.34 -> 2.26 ms with the same order in the fix-up kernel).
// (256, 7): 71 VGPRs, no spills, 7 waves per SIMD instead of 6; the row
// gathers are latency-bound (GCN scale-20 SpMM 2.42 -> 2.34 ms).
template <typename TB>
__global__ __launch_bounds__(256, 7) void hyb_spmm_sparse_kernel(SpmmPlanDev p, const TB* __restrict__ B,
                                                              int64_t ldb, int N, float* __restrict__ C,
                                                              int64_t ldc, int nC, bool vec, int n_mix,
                                                              int64_t so_waves) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t mix_waves = (int64_t)((n_mix + 3) / 4) * nC;
  if (wave < mix_waves) {
    const int64_t per = (n_mix + 3) / 4;
    const int cch = (int)(wave / per);
    const int q = (int)(wave - cch * per);
    const int col = cch * 64 + 4 * (lane & 15);
    const int m = q * 4 + (lane >> 4);
    if (m < n_mix) with_vec(vec, [&](auto V) { add_slot(p, B, ldb, N, C, ldc, col, V, p.mix_slot[m]); });
    return;
  }
  const int64_t idx = wave - mix_waves;
  if (idx >= so_waves) return;
  const int64_t per = so_waves / nC;
  const int cch = (int)(idx / per);
  const int64_t rr = idx - cch * per;
  const int w = p.W_tc + (int)(rr >> 2);
  const int quarter = (int)(rr & 3);
  const int n0 = cch * 64;
  const int col = n0 + 4 * (lane & 15);
  with_vec(vec, [&](auto V) { finish_slot(p, B, ldb, N, C, ldc, col, V, w * 16 + quarter * 4 + (lane >> 4)); });
}

// Sparse-only rows for launches with too few rows to fill the GPU at one row
// per 16-lane group: one wave per (row, 64 columns).  Task `wave` = (row
// wave / nC, column chunk wave % nC).  The row's nonzeros go 64 at a time:
// lane (g, j) loads nonzero base + 16g + j (column and value, one load each,
// so a round costs one index latency), lane group g's sixteen are passed to
// its lanes by shuffles, and the group's 16 B-row gathers are in flight
// together.  Positions past r1 read the row's last entry with weight 0.
// Fixed order: within a group u = 0..15 per round, then the xor-shuffle sum
// over the four groups.  Rows with more than kHubNnz sparse nonzeros sum with
// Neumaier compensation, as in finish_slot.
template <typename TB>
__device__ __forceinline__ void sparse_split_wave(const SpmmPlanDev& p, const TB* __restrict__ B, int64_t ldb,
                                                  int N, float* __restrict__ C, int64_t ldc, int nC, bool vec,
                                                  int64_t wave) {
  const int lane = threadIdx.x & 63;
  const int64_t rows = (int64_t)(p.W - p.W_tc) * 16;
  if (wave >= rows * nC) return;
  const int64_t ri = wave / nC;
  const int n0 = (int)(wave - ri * nC) * 64;
  const int slot = p.W_tc * 16 + (int)ri;
  const int g = lane >> 4;
  const int col = n0 + 4 * (lane & 15);
  const int row = p.win_rows[slot];
  const int r0 = p.res_ptr[slot], rend = p.res_ptr[slot + 1];
  const int r1 = min(rend, r0 + kSplitNnz);
  const bool compensated = rend - r0 > kHubNnz;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f), cc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int base = r0; base < r1; base += 64) {
    const int q = base + lane;  // = base + 16g + j
    const int qc = min(q, r1 - 1);
    const int kq = p.res_cols[qc];
    const float vq = q < r1 ? p.res_vals[qc] : 0.f;
    float4 b[16];
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int src = (lane & 48) + u;
      const int k = __shfl(kq, src);
      v[u] = __shfl(vq, src);
      b[u] = load_b4(B, ldb, k, col, N, vec);
    }
    if (compensated) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        neu_add(s.x, cc.x, v[u] * b[u].x);
        neu_add(s.y, cc.y, v[u] * b[u].y);
        neu_add(s.z, cc.z, v[u] * b[u].z);
        neu_add(s.w, cc.w, v[u] * b[u].w);
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        s.x += v[u] * b[u].x;
        s.y += v[u] * b[u].y;
        s.z += v[u] * b[u].z;
        s.w += v[u] * b[u].w;
      }
    }
  }
  s.x += cc.x; s.y += cc.y; s.z += cc.z; s.w += cc.w;
#pragma unroll
  for (int off = 16; off <= 32; off <<= 1) {
    s.x += __shfl_xor(s.x, off);
    s.y += __shfl_xor(s.y, off);
    s.z += __shfl_xor(s.z, off);
    s.w += __shfl_xor(s.w, off);
  }
  if (g == 0 && row >= 0) store_c4(C, ldc, row, col, N, vec, s);
}

template <typename TB>
__global__ __launch_bounds__(256) void hyb_spmm_sparse_split_kernel(SpmmPlanDev p, const TB* __restrict__ B,
                                                                    int64_t ldb, int N, float* __restrict__ C,
                                                                    int64_t ldc, int nC, bool vec) {
  with_vec(vec, [&](auto V) {
    sparse_split_wave(p, B, ldb, N, C, ldc, nC, V, (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6));
  });
}

// ---------------------------------------------------------------------------
// B-stationary variant for BF16 plans whose matrices all have K <= kLdsMaxK:
// a workgroup takes one work unit (a run of dense windows of one matrix,
// whose columns are B rows [koff, koff + ksz)) and one 64-column chunk,
// stages B[koff:koff+ksz, n0:n0+64] as BF16 in LDS and runs the unit's
// windows against it, a chunk-balanced run of windows per wave.  Each B row
// then leaves L2 once per workgroup instead of once per window.
//
// LDS row r (B row koff + r) starts at byte r * kLdsRowBytes and holds the
// 64 columns in the order position 16t + 4q + s <-> column n0 + 16q + 4s + t.
// The MFMA B operand of tile t is read with ds_read_b64_tr_b16: within each
// 16-lane group, result element e of lane i is element i % 4 of the 8 bytes
// that lane 4e + i / 4 of the group addresses.  Lane (g, 4e' + q) addresses
// row tc_cols[c][g][e'] (first read) and tc_cols[c][g][e' + 4] (second read)
// at position 16t + 4q, so lane j of group g receives B[col(g, e)][n0 + 4j +
// t] for e = 0..7 -- its MFMA B operand, with no repacking in registers, and
// the lane's four tiles hold four consecutive columns (each C store
// instruction writes four rows of 64 consecutive floats).  The 32 padding
// bytes per row put the 32 bytes one row contributes to a read at bank
// offset 40 r + 8 t (mod 64 banks): the 8 rows of a half-wave fall on
// distinct banks when their row ids differ mod 8, which the plan arranges
// (k-slot order within each chunk).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
  const uint16_t a = __builtin_bit_cast(uint16_t, (__bf16)lo);
  const uint16_t b = __builtin_bit_cast(uint16_t, (__bf16)hi);
  return (uint32_t)a | ((uint32_t)b << 16);
}

__device__ __forceinline__ uint32_t half_sel(uint32_t x, uint32_t y, int hi) {
  // (x.lo, y.lo) or (x.hi, y.hi) as one packed pair: one v_perm_b32, whose
  // selector bytes 0-3 pick bytes of its second operand (x), 4-7 of its
  // first (y)
  return __builtin_amdgcn_perm(y, x, hi ? 0x07060302u : 0x05040100u);
}

constexpr int kLdsThreads = 1024;

typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) unsigned char lds_u8;
typedef __attribute__((address_space(3))) i16x4 lds_i16x4;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ i16x4 lds_tr_read(lds_u8* p) {
  return __builtin_amdgcn_ds_read_tr16_b64_v4i16(reinterpret_cast<lds_i16x4*>(p));
}

// Stage B[koff : koff + ksz, n0 : n0 + 64] into the LDS rows.  Consecutive
// threads take consecutive 16-byte pieces of the block, so every load
// instruction reads whole rows; each thread loads all its pieces (at most
// 8 of 16-bit B, 12 of FP32 B per round: one round up to K = 768) before
// writing them.  Columns at or beyond N are staged as 0.
//   16-bit B (vec16): piece (row r, columns 8b .. 8b + 7) -> for t = 0..3 the
//     pair of columns 8b + t, 8b + 4 + t to position 16t + 4(b / 2) + 2(b % 2)
//     (one v_perm_b32 and one 4-byte write each);
//   FP32 B: piece (row r, columns 4m .. 4m + 3) -> column 4m + t to position
//     16t + m (two packed conversions, four 2-byte writes).
template <typename TB>
__device__ __forceinline__ void stage_b(lds_u8* lds, const TB* __restrict__ B, int64_t ldb, int koff, int ksz,
                                        int n0, int N, bool vec, bool vec16) {
  if constexpr (!std::is_same<TB, float>::value) {
    if (vec16) {
      constexpr int U = kLdsMaxK * 8 / kLdsThreads;
      const int items = ksz * 8;
      uint4 d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int it = min((int)threadIdx.x + u * kLdsThreads, items - 1);
        const int c0 = n0 + 8 * (it & 7);
        d[u] = c0 < N ? *reinterpret_cast<const uint4*>(B + (int64_t)(koff + (it >> 3)) * ldb + c0)
                      : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int it = threadIdx.x + u * kLdsThreads;
        if (it < items) {
          uint32_t dd[4] = {d[u].x, d[u].y, d[u].z, d[u].w};
          if constexpr (std::is_same<TB, _Float16>::value) {
            const float4 lo = cvt4<TB>(make_uint2(dd[0], dd[1])), hi = cvt4<TB>(make_uint2(dd[2], dd[3]));
            dd[0] = bf16_pair(lo.x, lo.y);
            dd[1] = bf16_pair(lo.z, lo.w);
            dd[2] = bf16_pair(hi.x, hi.y);
            dd[3] = bf16_pair(hi.z, hi.w);
          }
          const int b = it & 7;
          lds_u8* row = lds + (it >> 3) * kLdsRowBytes + 8 * (b >> 1) + 4 * (b & 1);
#pragma unroll
          for (int t = 0; t < 4; ++t)
            *reinterpret_cast<lds_u32*>(row + 32 * t) = half_sel(dd[t >> 1], dd[2 + (t >> 1)], t & 1);
        }
      }
      return;
    }
  }
  // 12 loads of 16 bytes in flight per thread: one round up to K = 768
  // (48 registers; the persistent variants have no room for 64); 8 for the
  // unaligned 16-bit fallback, which shares the kernel with the 16-byte path
  constexpr int U = std::is_same<TB, float>::value ? 12 : 8;
  const int items = ksz * 16;
  for (int base = threadIdx.x; base < items; base += U * kLdsThreads) {
    float4 f[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int it = min(base + u * kLdsThreads, items - 1);
      f[u] = load_b4(B, ldb, koff + (it >> 4), n0 + 4 * (it & 15), N, vec);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int it = base + u * kLdsThreads;
      if (it < items) {
        lds_u8* pos = lds + (it >> 4) * kLdsRowBytes + 2 * (it & 15);
        const uint32_t p01 = bf16_pair(f[u].x, f[u].y), p23 = bf16_pair(f[u].z, f[u].w);
        *reinterpret_cast<lds_u16*>(pos) = (uint16_t)p01;
        *reinterpret_cast<lds_u16*>(pos + 32) = (uint16_t)(p01 >> 16);
        *reinterpret_cast<lds_u16*>(pos + 64) = (uint16_t)p23;
        *reinterpret_cast<lds_u16*>(pos + 96) = (uint16_t)(p23 >> 16);
      }
    }
  }
}

// Column ids of k-slots (g, e') and (g, e' + 4) (the rows this lane
// addresses in the transposed reads) and the lane's A operand of one chunk
// or, for a scatter chunk (sc, wave-uniform), the 16 bytes of k-slot words
// that hold the lane's one entry.
struct TrChunk {
  int k0, k1;
  bf16x8 a;
  bool sc;
};

__device__ __forceinline__ TrChunk load_tr_chunk(const SpmmPlanDev& p, int c, int cidx, int lane, bool sc) {
  TrChunk r;
  const int32_t* cp = p.tc_cols + (int64_t)c * 32 + cidx;
  r.k0 = cp[0];
  r.k1 = cp[4];
  // scatter chunk: lane (G, row) needs the word of k-slot 8G + row % 8, in
  // 16-byte group 2G + (row % 8) / 4 of the chunk's A block
  const int q = sc ? 2 * (lane >> 4) + ((lane >> 2) & 1) : lane;
  r.a = reinterpret_cast<const bf16x8*>(p.tc_vals)[(int64_t)c * 64 + q];
  r.sc = sc;
  return r;
}

// The MFMA A operand of a scatter chunk for lane (G, row): element row % 8
// holds the value of k-slot 8G + row % 8 when that entry's row is this
// lane's (its word's bit 16 = row >= 8), every other element 0.
__device__ __forceinline__ bf16x8 scatter_operand(const bf16x8& raw, int lane) {
  const uint4 w4 = __builtin_bit_cast(uint4, raw);
  const int s = lane & 3;
  const uint32_t v = s == 0 ? w4.x : s == 1 ? w4.y : s == 2 ? w4.z : w4.w;
  const uint32_t val = ((v >> 16) & 1u) == (uint32_t)((lane >> 3) & 1) ? (v & 0xffffu) : 0u;
  const uint32_t h = (lane & 1) ? val << 16 : val;
  const int word = (lane >> 1) & 3;
  uint4 o;
  o.x = word == 0 ? h : 0u;
  o.y = word == 1 ? h : 0u;
  o.z = word == 2 ? h : 0u;
  o.w = word == 3 ? h : 0u;
  return __builtin_bit_cast(bf16x8, o);
}

// cst = 8q - koff * kLdsRowBytes: the lane's position offset within a row,
// and the unit's first B row
__device__ __forceinline__ void tr_chunk_mfma(lds_u8* lds, int cst, const TrChunk& r, f32x4 (&acc)[4]) {
  const bf16x8 a = r.sc ? scatter_operand(r.a, threadIdx.x & 63) : r.a;
  // 24-bit multiplies (column ids < 2^24): one full-rate v_mad_u32_u24 each
  lds_u8* p0 = lds + (int)(__umul24((unsigned)r.k0, (unsigned)kLdsRowBytes) + (unsigned)cst);
  lds_u8* p1 = lds + (int)(__umul24((unsigned)r.k1, (unsigned)kLdsRowBytes) + (unsigned)cst);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const i16x4 lo = lds_tr_read(p0 + 32 * t);
    const i16x4 hi = lds_tr_read(p1 + 32 * t);
    const i16x8 b = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, b), acc[t], 0, 0, 0);
  }
}

// Rows of the lane group's 4 window slots.
__device__ __forceinline__ SlotRows load_win(const SpmmPlanDev& p, int slot0) { return load_rows(p, slot0); }

// One round of a window's sparse part (kMixed plans): e holds the round's
// entry of slot 4g + (lane & 3) (column, value bits); lane (g, j) adds
// v * B[col][n0 + 4j + t] of the group's four slots to acc[t][i].  BF16 B is
// read from the staged LDS block (position 16t + j of row col - koff), other
// B types from memory (FP32 products on B's exact values either way).
template <typename TB>
__device__ __forceinline__ void sparse_round(const TB* __restrict__ B, int64_t ldb, int N, int col, bool vec,
                                             lds_u8* lds, int koff, int2 e, f32x4 (&acc)[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int src = (lane & 48) + i;
    const int k = __shfl(e.x, src);
    const float v = __int_as_float(__shfl(e.y, src));
    float4 b;
    if constexpr (std::is_same<TB, __bf16>::value) {
      lds_u8* rp = lds + (k - koff) * kLdsRowBytes + 2 * (lane & 15);
      b.x = bf16_lo(*reinterpret_cast<lds_u16*>(rp));
      b.y = bf16_lo(*reinterpret_cast<lds_u16*>(rp + 32));
      b.z = bf16_lo(*reinterpret_cast<lds_u16*>(rp + 64));
      b.w = bf16_lo(*reinterpret_cast<lds_u16*>(rp + 96));
    } else {
      b = load_b4(B, ldb, k, col, N, vec);
    }
    acc[0][i] += v * b.x;
    acc[1][i] += v * b.y;
    acc[2][i] += v * b.z;
    acc[3][i] += v * b.w;
  }
}

// One work item (unit item / nC, column chunk item % nC) of the B-stationary
// kernel.  tc_ptr and sp_off are p.tc_ptr and p.sp_off, passed again as
// restrict parameters so that the compiler may read them with scalar loads
// inside the loop that stores C.  Returns early per wave (a wave with no
// windows) or per workgroup (an empty unit); the caller's barrier before the
// next item's staging keeps the LDS block until every wave is done with it.
// kMixed: windows may have sparse parts, added (in rounds, p.sp_ent) to the
// dense sums before the window's rows are stored; a window's first rounds
// (four, two with FP32 or FP16 B) are requested when it starts, later ones
// in batches of as many when it ends.  Scatter chunks (p.tc_kind) build their A operand from
// one k-slot word per lane.
typedef float f32v4_t __attribute__((ext_vector_type(4)));

// The window's rows as float4 stores of four columns each, one per row.
// kVec (C rows 16-byte aligned, N a multiple of 4, C below 2 GiB): buffer
// stores with L2 write-through (sc1) over a descriptor of [C, C + cbytes),
// measured 4 % faster on the block-pruned layer at N = 4096 than
// non-temporal global stores; stores that fall outside C (rows of -1,
// columns at or beyond N) get an offset past the descriptor, which the
// hardware drops, so every lane issues the same stores on every path.
// Otherwise store_window.
template <bool kVec>
__device__ __forceinline__ void store_window_lds(float* __restrict__ C, int64_t ldc, int N, int col, bool vec,
                                                 int cbytes, const SlotRows& sr, const f32x4 (&acc)[4]) {
  if constexpr (kVec) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(C, 0, cbytes, 0x00020000);
    constexpr uint32_t kOut = 0x80000000u;  // past every descriptor (cbytes < 2^31)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ok = sr.row[i] >= 0 && col < N;
      const uint32_t off = ok ? ((uint32_t)sr.row[i] * (uint32_t)ldc + (uint32_t)col) * 4u : kOut;
      __builtin_amdgcn_raw_buffer_store_b128(f32v4_t{acc[0][i], acc[1][i], acc[2][i], acc[3][i]}, rs, (int)off, 0,
                                            16);
    }
  } else {
    store_window<true>(C, ldc, N, col, vec, sr, acc);
  }
}

template <typename TB, bool kMixed, bool kVec, bool kSc>
__device__ __forceinline__ void lds_item(const SpmmPlanDev& p, const int32_t* __restrict__ tc_ptr,
                                         const int32_t* __restrict__ sp_off, const uint32_t* __restrict__ tc_kind,
                                         const int32_t* __restrict__ win_r0, const int4* __restrict__ units,
                                         const TB* __restrict__ B, int64_t ldb, int N, float* __restrict__ C,
                                         int64_t ldc, int nC, bool vec, bool vec16, int cbytes, int item,
                                         lds_u8* lds) {
  constexpr int kWaves = kLdsThreads / 64;
  const int chunk = item % nC;
  const int4 unit = units[item / nC];
  const int n0 = chunk * 64;
  const int koff = unit.z, ksz = unit.w;
  // The wave index is made visibly uniform so that the window and chunk
  // bookkeeping below lives in scalar registers and reads tc_ptr with scalar
  // loads: a vector load of the next window bound inside the loop would make
  // every step wait for all outstanding vector loads and stores.
  const int lane = threadIdx.x & 63, wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, j = lane & 15;
  const int col = n0 + 4 * j;  // the lane's columns: col .. col + 3
  const int cidx = g * 8 + ((lane >> 2) & 3);
  const int cst = 8 * (lane & 3) - koff * kLdsRowBytes;
  const int eidx = 4 * g + (lane & 3);  // slot of the round entry this lane requests
  // This wave's contiguous run of the unit's windows, balanced by chunk count.
  const int wg_beg = unit.x;
  const int wg_end = unit.y;
  if (wg_beg >= wg_end) return;  // the whole workgroup
  const int cb = tc_ptr[wg_beg], ce = tc_ptr[wg_end];
  auto window_at = [&](int target) {  // first window w in [wg_beg, wg_end] with tc_ptr[w] >= target
    int lo = wg_beg, hi = wg_end;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tc_ptr[mid] < target) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const int w_beg = window_at(cb + (int)(((int64_t)(ce - cb) * wib) / kWaves));
  const int w_end = window_at(cb + (int)(((int64_t)(ce - cb) * (wib + 1)) / kWaves));
  const bool has = w_beg < w_end;  // wave-uniform
  const int c_beg = has ? tc_ptr[w_beg] : 0, c_end = has ? tc_ptr[w_end] : 1;

  // The operands of chunk c + 3 are requested while chunk c is multiplied
  // (indices clamped to the wave's last chunk, so every load is
  // unconditional): a wave's loop is bound by the latency of these requests,
  // so three are kept in flight.  Four register sets rotate through a loop
  // unrolled by four, so no set is copied (a copy would wait for the newest
  // load).  Every window of the run holds at least one chunk, so at most one
  // window ends before each chunk.  The first two chunks' operands are
  // requested before the B block is staged, so their latency overlaps the
  // staging, the third one after it (fewer registers live while the staging
  // holds its loads).
  // A window's rows are not carried in registers: for consecutive rows they
  // are computed when the window is stored, from its first row (r0w, a
  // scalar load); other windows load theirs then.  (Rows loaded ahead into
  // registers that the consecutive path overwrites made every window end
  // wait for all loads in flight, the next chunks' operands included.)
  const int last = c_end - 1;
  // chunk kind (scatter or dense): kSc (the plan has scatter chunks, tc_kind
  // set) -- a scalar load of its word of tc_kind; otherwise dense, with no
  // per-chunk bookkeeping (measured 4 % of the block-pruned layer's launch)
  auto is_sc = [&](int idx) { return kSc && ((tc_kind[idx >> 5] >> (idx & 31)) & 1u) != 0; };
  int w = w_beg;
  int w_next = 0;
  int r0w = 0;
  TrChunk r0{}, r1{}, r2{};
  // kMixed: sp_off of the current window and the next, and the lane's
  // entries of the current window's rounds 0 .. kPfRounds - 1, requested when
  // the window starts and used when it ends.  The loads are unconditional
  // (indices clamped to the last entry; a plan with mixed windows has at
  // least one round), so the wait before their use is an exact count.
  constexpr int kPfRounds = std::is_same<TB, __bf16>::value ? 4 : 2;  // FP32 / FP16 B: fewer, for registers
  constexpr int kBatchRounds = std::is_same<TB, __bf16>::value ? 4 : 2;  // later rounds, per request batch
  int so0 = 0, so1 = 0;
  const int sp_last = kMixed ? sp_off[p.W_tc] - 1 : 0;
  int2 ecur[kPfRounds];
  auto request_rounds = [&](int so) {
#pragma unroll
    for (int u = 0; u < kPfRounds; ++u) ecur[u] = p.sp_ent[min(so + 16 * u + eidx, sp_last)];
  };
  if (has) {
    w_next = tc_ptr[w + 1];
    r0w = win_r0[w];
    r0 = load_tr_chunk(p, c_beg, cidx, lane, is_sc(c_beg));
    r1 = load_tr_chunk(p, min(c_beg + 1, last), cidx, lane, is_sc(min(c_beg + 1, last)));
    if constexpr (kMixed) {
      so0 = sp_off[w];
      so1 = sp_off[w + 1];
      request_rounds(so0);
    }
  }
  stage_b(lds, B, ldb, koff, ksz, n0, N, vec, vec16);
  if (has) r2 = load_tr_chunk(p, min(c_beg + 2, last), cidx, lane, is_sc(min(c_beg + 2, last)));
  __syncthreads();
  if (!has) return;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  auto finish = [&]() {
    if constexpr (kMixed) {
      const int rounds = (so1 - so0) >> 4;  // wave-uniform
      if (rounds > 0) {
#pragma unroll
        for (int u = 0; u < kPfRounds; ++u)
          if (u < rounds) sparse_round(B, ldb, N, col, vec, lds, koff, ecur[u], acc);
        // rounds beyond those: entries requested a batch at a time, one
        // memory latency per batch
        for (int r = kPfRounds; r < rounds; r += kBatchRounds) {
          int2 e[kBatchRounds];
#pragma unroll
          for (int u = 0; u < kBatchRounds; ++u)
            e[u] = r + u < rounds ? p.sp_ent[so0 + 16 * (r + u) + eidx] : make_int2(koff, 0);
#pragma unroll
          for (int u = 0; u < kBatchRounds; ++u)
            if (r + u < rounds) sparse_round(B, ldb, N, col, vec, lds, koff, e[u], acc);
        }
      }
    }
    SlotRows cur;
    if (r0w >= 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) cur.row[i] = r0w + 4 * g + i;
    } else {
      // The loaded rows are consumed inside this branch (the empty asm reads
      // them), so the wait for the load (and for the loads in flight before
      // it) stays on this path: where the branches join no load is pending
      // on the registers the consecutive path writes.
      cur = load_win(p, w * 16 + g * 4);
      asm volatile("" : "+v"(cur.row[0]), "+v"(cur.row[1]), "+v"(cur.row[2]), "+v"(cur.row[3]));
    }
    store_window_lds<kVec>(C, ldc, N, col, vec, cbytes, cur, acc);
  };
  TrChunk r3;
  int c = c_beg;
  auto step = [&](const TrChunk& use, TrChunk& fill) {
    fill = load_tr_chunk(p, min(c + 3, last), cidx, lane, is_sc(min(c + 3, last)));
    if (c >= w_next) {  // window w is complete: store it
      const int w_next2 = tc_ptr[w + 2];
      const int r0n = win_r0[w + 1];
      int so2 = 0;
      if constexpr (kMixed) so2 = sp_off[w + 2];
      finish();
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
      ++w;
      w_next = w_next2;
      r0w = r0n;
      if constexpr (kMixed) {
        so0 = so1;
        so1 = so2;
        request_rounds(so0);
      }
    }
    tr_chunk_mfma(lds, cst, use, acc);
    ++c;
  };
  while (c + 4 <= c_end) {
    step(r0, r3);
    step(r1, r0);
    step(r2, r1);
    step(r3, r2);
  }
  if (c < c_end) step(r0, r3);
  if (c < c_end) step(r1, r0);
  if (c < c_end) step(r2, r1);
  finish();
}

// units: the work-unit list of this launch.  kLoop = false: one workgroup
// per item (grid = items <= compute units); true: persistent workgroups that
// restage the LDS block for each item they take: items blockIdx.x,
// blockIdx.x + gridDim.x, ... or, with xcd (grid a multiple of kXcds), items
// xr.start[x] + blockIdx.x / kXcds, + gridDim.x / kXcds, ... below
// xr.start[x + 1] for x = blockIdx.x % kXcds.  Two instantiations because
// the loop costs the one-item launches about 0.6 us (more registers live
// across the item body).
//
// Tail: the last tail_blocks workgroups of the grid then run the sparse-only
// rows' tasks (one wave per (row, 64 columns), as hyb_spmm_sparse_split_kernel)
// in place of a second launch: block gridDim.x - 1 - i takes tasks 16i ..
// 16i + 15 of tail_tasks.  Under the XCD grouping those are the last slot of
// each XCD, which holds the fewest items when a range does not divide evenly;
// one-item launches add tail blocks past the items when CUs are free.
template <bool kLoop, typename TB, bool kMixed, bool kVec>
__global__ __launch_bounds__(kLdsThreads) void hyb_spmm_lds_bf16_kernel(SpmmPlanDev p,
                                                                        const int32_t* __restrict__ tc_ptr,
                                                                        const int32_t* __restrict__ sp_off,
                                                                        const uint32_t* __restrict__ tc_kind,
                                                                        const int32_t* __restrict__ win_r0,
                                                                        const int4* __restrict__ units,
                                                                        const TB* __restrict__ B,
                                                                        int64_t ldb, int N, float* __restrict__ C,
                                                                        int64_t ldc, int nC, int n_items, bool vec,
                                                                        bool vec16, int cbytes, LdsXcdRanges xr,
                                                                        bool xcd, int tail_blocks, int tail_tasks) {
  extern __shared__ uint4 lds_dyn[];  // [ksz][kLdsRowBytes]
  lds_u8* lds = (lds_u8*)lds_dyn;
  const int tb = (int)gridDim.x - 1 - (int)blockIdx.x;
  auto run = [&](int item) {  // tc_kind: kernel-uniform
    if (tc_kind)
      lds_item<TB, kMixed, kVec, true>(p, tc_ptr, sp_off, tc_kind, win_r0, units, B, ldb, N, C, ldc, nC, vec, vec16,
                                       cbytes, item, lds);
    else
      lds_item<TB, kMixed, kVec, false>(p, tc_ptr, sp_off, tc_kind, win_r0, units, B, ldb, N, C, ldc, nC, vec,
                                        vec16, cbytes, item, lds);
  };
  if (!kLoop) {
    if ((int)blockIdx.x < n_items) run(blockIdx.x);
    if (tb < tail_blocks) {
      const int task = tb * (kLdsThreads / 64) + (int)(threadIdx.x >> 6);
      if (task < tail_tasks) with_vec(vec, [&](auto V) { sparse_split_wave(p, B, ldb, N, C, ldc, nC, V, task); });
    }
    return;
  }
  int beg = blockIdx.x, end = n_items, step = gridDim.x;
  if (xcd) {
    const int x = blockIdx.x % kXcds;
    int r0 = 0, r1 = 0;
#pragma unroll
    for (int i = 0; i < kXcds; ++i)  // selects, not a dynamically indexed copy
      if (i == x) {
        r0 = xr.start[i];
        r1 = xr.start[i + 1];
      }
    beg = r0 + (int)(blockIdx.x / kXcds);
    end = r1;
    step = gridDim.x / kXcds;
  }
  for (int item = beg; item < end; item += step) {
    __syncthreads();  // every wave is done reading the previous item's block
    run(item);
  }
  if (tb < tail_blocks) {
    const int task = tb * (kLdsThreads / 64) + (int)(threadIdx.x >> 6);
    if (task < tail_tasks) with_vec(vec, [&](auto V) { sparse_split_wave(p, B, ldb, N, C, ldc, nC, V, task); });
  }
}

// Remaining pieces of long sparse rows, execution slots [pc0, pc1) (piece
// p.piece_exec[slot], ordered by first column so that pieces of different
// rows gathering the same B rows run together), column chunk by column chunk
// as in hyb_spmm_sparse_kernel: lane group g of
// the wave takes nonzeros q = beg + g, beg + g + 4, ...; the four partial
// float4 are reduced across groups and added to C with atomicAdd (C already
// holds the window pass's value of the row), or, with ws (deterministic
// mode), stored to ws[piece][N] for hyb_spmm_piece_reduce_kernel.
template <typename TB>
__global__ __launch_bounds__(256) void hyb_spmm_fixup_kernel(SpmmPlanDev p, const TB* __restrict__ B,
                                                             int64_t ldb, int N, float* __restrict__ C,
                                                             int64_t ldc, int nC, bool vec, int pc0, int pc1,
                                                             float* __restrict__ ws) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int npc = pc1 - pc0;
  const int cch = wave / npc;
  if (cch >= nC) return;
  const int pe = pc0 + (wave - cch * npc);
  const int pc = p.piece_exec[pe];
  const int n0 = cch * 64;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int col = n0 + 4 * (lane & 15);
  const int beg = p.piece_beg[pc], end = p.piece_end[pc];
  with_vec(vec, [&](auto V) {
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f), cc = make_float4(0.f, 0.f, 0.f, 0.f);
  // eight gathers in flight per lane group (positions past the end read the
  // piece's last entry with weight 0)
  for (int q0 = beg + g; q0 < end; q0 += 32) {
    int k[8];
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int q = q0 + 4 * i;
      const int qc = min(q, end - 1);
      k[i] = p.res_cols[qc];
      v[i] = q < end ? p.res_vals[qc] : 0.f;
    }
    float4 b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = load_b4(B, ldb, k[i], col, N, V);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      neu_add(s.x, cc.x, v[i] * b[i].x);
      neu_add(s.y, cc.y, v[i] * b[i].y);
      neu_add(s.z, cc.z, v[i] * b[i].z);
      neu_add(s.w, cc.w, v[i] * b[i].w);
    }
  }
  s.x += cc.x; s.y += cc.y; s.z += cc.z; s.w += cc.w;
#pragma unroll
  for (int off = 16; off <= 32; off <<= 1) {
    s.x += __shfl_xor(s.x, off);
    s.y += __shfl_xor(s.y, off);
    s.z += __shfl_xor(s.z, off);
    s.w += __shfl_xor(s.w, off);
  }
  if (g != 0) return;
  if (ws) {
    store_c4(ws, N, pc, col, N, V, s);
    return;
  }
  float* cp = C + (int64_t)p.piece_row[pc] * ldc + col;
  if (col < N) atomicAdd(cp, s.x);
  if (col + 1 < N) atomicAdd(cp + 1, s.y);
  if (col + 2 < N) atomicAdd(cp + 2, s.z);
  if (col + 3 < N) atomicAdd(cp + 3, s.w);
  });
}

// Deterministic mode: C[row] += the row's pieces in piece order, for split
// rows [sr0, sr1).  One 16-lane group per (split row, 64 columns); C[row] is
// read and written by that group alone.
__global__ __launch_bounds__(256) void hyb_spmm_piece_reduce_kernel(SpmmPlanDev p, int N, float* __restrict__ C,
                                                                    int64_t ldc, int nC, bool vec, int sr0,
                                                                    int sr1, const float* __restrict__ ws) {
  const int64_t task = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int r = sr0 + (int)(task / nC);
  if (r >= sr1) return;
  const int col = (int)(task - (int64_t)(r - sr0) * nC) * 64 + 4 * (threadIdx.x & 15);
  const int row = p.srow_row[r];
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int q = p.srow_p0[r]; q < p.srow_p1[r]; ++q) {
    const float4 v = load_b4(ws, N, q, col, N, vec);
    s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
  }
  const float4 c = load_b4(C, ldc, row, col, N, vec);
  store_c4(C, ldc, row, col, N, vec, make_float4(c.x + s.x, c.y + s.y, c.z + s.z, c.w + s.w));
}

template <bool kLoop, typename TB>
static void set_lds_bytes(int bytes) {
  for (const void* f : {reinterpret_cast<const void*>(hyb_spmm_lds_bf16_kernel<kLoop, TB, false, false>),
                        reinterpret_cast<const void*>(hyb_spmm_lds_bf16_kernel<kLoop, TB, true, false>),
                        reinterpret_cast<const void*>(hyb_spmm_lds_bf16_kernel<kLoop, TB, false, true>),
                        reinterpret_cast<const void*>(hyb_spmm_lds_bf16_kernel<kLoop, TB, true, true>)})
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

void prepare_spmm_kernels(const SpmmPlanDev& p) {
  const int bytes = p.lds_max_k * kLdsRowBytes;
  if (!p.lds_units || bytes <= 65536) return;
  set_lds_bytes<false, float>(bytes);
  set_lds_bytes<true, float>(bytes);
  set_lds_bytes<false, __bf16>(bytes);
  set_lds_bytes<true, __bf16>(bytes);
  set_lds_bytes<false, _Float16>(bytes);
  set_lds_bytes<true, _Float16>(bytes);
}

int device_cus() {
  // per-device cache; threads that race on the first query store the same value
  static std::atomic<int> cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

int64_t spmm_workspace_floats(const SpmmPlanDev& p, int N) {
  return ((int64_t)p.npieces + (int64_t)p.n_parts * 16) * N;
}

template <typename TB>
static hipError_t launch_hyb_spmm_t(const SpmmPlanDev& p, const TB* B, int64_t ldb, int N, float* C, int64_t ldc,
                                    bool det, float* ws, hipStream_t s) {
  const int eb = (int)sizeof(TB);
  const bool vec = (N % 4 == 0) && (ldb % 4 == 0) && (ldc % 4 == 0) &&
                   ((reinterpret_cast<uintptr_t>(B) & (4 * eb - 1)) == 0) &&
                   ((reinterpret_cast<uintptr_t>(C) & 15) == 0);
  const bool vec16 = eb == 2 && vec && N % 8 == 0 && ldb % 8 == 0 && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
  const int nC = (N + 63) / 64;
  // workspace regions of the deterministic mode: pieces, then huge-window parts
  float* ws_pieces = det ? ws : nullptr;
  float* ws_parts = det ? ws + (int64_t)p.npieces * N : nullptr;
  // B-stationary kernel for every BF16 plan whose matrices all have K <=
  // kLdsMaxK (measured on the block-pruned layer: 9.0 us vs 15.0 us
  // streaming at N = 128, 20.8 vs 82 us at N = 4096); otherwise the
  // streaming kernel.  Unit list: the plan's pick for nC column chunks (up to
  // kLdsPickMax), else the smallest list with at least one item per CU.
  // (launch options of the plan: opt_kernel 1 = streaming kernel, refused
  // at plan level for plans with scatter chunks; opt_lds_list >= 0 = that
  // unit list; opt_lds_xcd 0 = items round-robin instead of per-XCD ranges)
  bool lds = p.lds_units != nullptr && p.W_tc > 0;
  if (p.opt_kernel == 1 && !p.tc_kind) lds = false;
  // Sparse-only rows: with the LDS kernel and at most kLdsTailMaxTasks
  // (row, 64 columns) tasks they are the LDS launch's tail; otherwise a
  // launch of their own follows the dense kernels.
  const int64_t so_rows = (int64_t)(p.W - p.W_tc) * 16;
  const int cus = device_cus();
  const bool fold = lds && so_rows > 0 && so_rows * nC <= kLdsTailMaxTasks &&
                    (so_rows * nC + kLdsThreads / 64 - 1) / (kLdsThreads / 64) <= cus / kXcds * kXcds;
  // The LDS kernel runs every dense window with its sparse part; otherwise
  // the streaming kernel does the dense parts (its first blocks the big
  // windows) and the sparse kernel the sparse parts.
  if (lds) {
    int t = p.opt_lds_list >= 0 && p.opt_lds_list < kLdsLists ? p.opt_lds_list : -1;
    if (t < 0 && nC <= kLdsPickMax) t = p.lds_pick[nC];
    if (t < 0) {
      t = 0;
      while (t + 1 < kLdsLists && (int64_t)p.lds_cnt[t] * nC < cus) ++t;
    }
    const int64_t items = (int64_t)p.lds_cnt[t] * nC;
    if (items > INT32_MAX) return hipErrorInvalidValue;
    // one workgroup per CU at most (each holds a staged block of LDS)
    int grid = (int)std::min<int64_t>(items, cus);
    LdsXcdRanges xr{};
    const bool xcd = items > cus && nC <= kLdsXcdMaxNC && cus >= kXcds && p.lds_xcd_host && p.opt_lds_xcd != 0;
    if (xcd) {
      const int32_t* st = p.lds_xcd_host + ((size_t)t * (kLdsXcdMaxNC + 1) + nC) * (kXcds + 1);
      if (st[0] != 0 || st[kXcds] != items) return hipErrorInvalidValue;  // ranges cover every item once
      for (int x = 0; x < kXcds; ++x)
        if (st[x + 1] < st[x]) return hipErrorInvalidValue;
      std::copy(st, st + kXcds + 1, xr.start);
      grid = cus / kXcds * kXcds;
    }
    const int tail_tasks = fold ? (int)(so_rows * nC) : 0;
    const int tail_blocks = (tail_tasks + kLdsThreads / 64 - 1) / (kLdsThreads / 64);
    // one-item launches: tail blocks of their own past the items while CUs
    // are free, else the last item blocks take the tail
    if (items <= cus && items + tail_blocks <= cus) grid = (int)items + tail_blocks;
    if (items <= cus) grid = std::max(grid, tail_blocks);
    const size_t lds_bytes = (size_t)p.lds_max_k * kLdsRowBytes;
    const int4* units = p.lds_units + p.lds_off[t];
    const bool mixed = p.n_mix > 0 && p.sp_off != nullptr;
    // C addressed by a buffer descriptor (the kVec variants) when its extent
    // is below 2 GiB
    const int64_t c_extent = ((int64_t)(p.M - 1) * ldc + N) * 4;
    const int cbytes = p.M > 0 && c_extent < ((int64_t)1 << 31) ? (int)c_extent : 0;
#define STILE_LDS_LAUNCH1(LOOP, MIX, VEC)                                                                    \
  hipLaunchKernelGGL((hyb_spmm_lds_bf16_kernel<LOOP, TB, MIX, VEC>), dim3(grid), dim3(kLdsThreads), lds_bytes, s, \
                     p,                                                                                        \
                     p.tc_ptr, p.sp_off, p.tc_kind, p.win_r0, units, B, ldb, N, C, ldc, nC, (int)items, vec, vec16, \
                     cbytes, xr, LOOP && xcd, tail_blocks, tail_tasks)
#define STILE_LDS_LAUNCH(LOOP, MIX)                \
  do {                                             \
    if (vec && cbytes > 0)                         \
      STILE_LDS_LAUNCH1(LOOP, MIX, true);          \
    else                                           \
      STILE_LDS_LAUNCH1(LOOP, MIX, false);         \
  } while (0)
    if (items <= cus) {
      if (mixed) STILE_LDS_LAUNCH(false, true); else STILE_LDS_LAUNCH(false, false);
    } else {
      if (mixed) STILE_LDS_LAUNCH(true, true); else STILE_LDS_LAUNCH(true, false);
    }
#undef STILE_LDS_LAUNCH
#undef STILE_LDS_LAUNCH1
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else if (p.W_tc > 0) {
    if (p.n_parts > 0 && !det) {
      hipError_t e = launch_zero_window_rows(p, p.W_huge, C, ldc, N, s);
      if (e != hipSuccess) return e;
    }
    const int64_t big_blocks = (int64_t)(p.W_big - p.W_huge) * nC;
    const int64_t blocks = (int64_t)p.n_parts * nC + big_blocks +
                           ((int64_t)(p.W_tc - p.W_big) * nC + kBigWaves - 1) / kBigWaves;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    dim3 grid((unsigned)blocks), block(kBigWaves * 64);
    const int bb = (int)big_blocks;
    switch (p.blk_dtype) {
      case 1:
        hipLaunchKernelGGL((hyb_spmm_kernel<1, TB>), grid, block, 0, s, p, B, ldb, N, C, ldc, nC, bb, vec,
                           ws_parts);
        break;
      case 2:
        hipLaunchKernelGGL((hyb_spmm_kernel<2, TB>), grid, block, 0, s, p, B, ldb, N, C, ldc, nC, bb, vec,
                           ws_parts);
        break;
      default:
        hipLaunchKernelGGL((hyb_spmm_kernel<0, TB>), grid, block, 0, s, p, B, ldb, N, C, ldc, nC, bb, vec,
                           ws_parts);
        break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (det && p.W_huge > 0) {
      const int64_t waves = (int64_t)p.W_huge * nC;
      hipLaunchKernelGGL(hyb_spmm_part_reduce_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, p, N, C,
                         ldc, nC, vec, ws_parts);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  // Sparse-only rows (unless folded) one per lane group, or, when that
  // leaves fewer than 16 waves per CU for rows of 8 or more nonzeros on
  // average, one per wave; before them, unless the LDS kernel added them,
  // the sparse parts of the dense windows' rows.
  const int n_mix = lds ? 0 : p.n_mix;
  const int64_t so_left = fold ? 0 : so_rows;
  const bool split = so_left > 0 && so_left * nC / 4 < 4096 && p.nnz_sparse_only >= 8 * so_left;
  const int64_t sparse_waves = (int64_t)((n_mix + 3) / 4) * nC + (split ? 0 : so_left * nC / 4);
  if (split) {
    const int64_t blocks = (so_rows * nC + 3) / 4;
    hipLaunchKernelGGL(hyb_spmm_sparse_split_kernel<TB>, dim3((unsigned)blocks), dim3(256), 0, s, p, B, ldb, N, C,
                       ldc, nC, vec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (sparse_waves > 0) {
    const int64_t blocks = (sparse_waves + 3) / 4;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hyb_spmm_sparse_kernel<TB>, dim3((unsigned)blocks), dim3(256), 0, s, p, B, ldb, N, C, ldc,
                       nC, vec, n_mix, split ? (int64_t)0 : so_left * nC / 4);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // Pieces of long rows.  The LDS kernel summed its windows' rows whole, so
  // only the sparse-only rows' pieces remain after it.
  const int pc0 = lds ? p.npieces_dense : 0;
  if (p.npieces > pc0) {
    const int64_t fw = (int64_t)(p.npieces - pc0) * nC;
    hipLaunchKernelGGL(hyb_spmm_fixup_kernel<TB>, dim3((unsigned)((fw + 3) / 4)), dim3(256), 0, s, p, B, ldb, N,
                       C, ldc, nC, vec, pc0, p.npieces, ws_pieces);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (det) {
      const int sr0 = lds ? p.nsrows_dense : 0;
      const int64_t tasks = (int64_t)(p.nsrows - sr0) * nC;
      if (tasks > 0) {
        hipLaunchKernelGGL(hyb_spmm_piece_reduce_kernel, dim3((unsigned)((tasks + 15) / 16)), dim3(256), 0, s, p,
                           N, C, ldc, nC, vec, sr0, p.nsrows, ws_pieces);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
    }
  }
  return hipSuccess;
}

hipError_t launch_hyb_spmm(const SpmmPlanDev& p, int b_dtype, const void* B, int64_t ldb, int N, float* C,
                           int64_t ldc, bool deterministic, float* ws, hipStream_t s) {
  if (p.W == 0 || N == 0) return hipSuccess;
  switch (b_dtype) {
    case 1:
      return launch_hyb_spmm_t(p, static_cast<const __bf16*>(B), ldb, N, C, ldc, deterministic, ws, s);
    case 2:
      return launch_hyb_spmm_t(p, static_cast<const _Float16*>(B), ldb, N, C, ldc, deterministic, ws, s);
    default:
      return launch_hyb_spmm_t(p, static_cast<const float*>(B), ldb, N, C, ldc, deterministic, ws, s);
  }
}

// ---------------------------------------------------------------------------
// Hybrid fused SDDMM.  Waves [0, waves1d) run the 1D entries (segments of LPE
// lanes, kSddmmU consecutive entries each); the remaining waves each run one
// dense 16x16 block:
// lane l = 16g + j loads float4 slices of X[rows[j]] and Y[cols[j]] at
// features kb + 4g .. kb + 4g + 3 and feeds component e to MFMA e
// (v_mfma_f32_16x16x4_f32, so k-slot g of MFMA e is feature kb + 4g + e).
// The 16x16 result goes through LDS to the sampled scatter.
// ---------------------------------------------------------------------------
constexpr int kSddmmU = 4;

template <int LPE>
// (256, 8): registers for 8 waves per SIMD (59 VGPRs, no spills); the 1D
// entries are latency-bound gathers and more waves in flight measured faster
// (R-MAT scale 16, K = 32: 16.3 -> 15.2 us with every entry in the 1D part).
__global__ __launch_bounds__(256, 8) void hyb_sddmm_kernel(SddmmPlanDev p,
                                                        const float* __restrict__ X,
                                                        int64_t ldx,
                                                        const float* __restrict__ Y,
                                                        int64_t ldy, int Kf,
                                                        float* __restrict__ out,
                                                        int waves1d, bool vec) {
  __shared__ float blk[4][256];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wave = blockIdx.x * 4 + wib;
  if (wave < waves1d) {
    // a segment of LPE lanes computes kSddmmU consecutive entries, their
    // index loads and row gathers all in flight together
    constexpr int EPW = 64 / LPE;
    const int sub = lane % LPE;
    const int e0 = (wave * EPW + lane / LPE) * kSddmmU;
    if (e0 >= p.n1d) return;
    int row[kSddmmU], col[kSddmmU], dst[kSddmmU];
#pragma unroll
    for (int u = 0; u < kSddmmU; ++u) {
      const int e = min(e0 + u, p.n1d - 1);
      row[u] = p.e_row[e];
      col[u] = p.e_col[e];
      dst[u] = p.e_out[e];
    }
    float s[kSddmmU];
#pragma unroll
    for (int u = 0; u < kSddmmU; ++u) s[u] = 0.f;
    if (vec) {
      // consecutive entries of one row (CSR order) share the X load; an
      // entry whose row repeats skips its load and takes the previous X
      // value in a select after every load is issued (a copy inside the
      // branch would wait for the previous load before the next one)
      bool same[kSddmmU];
#pragma unroll
      for (int u = 0; u < kSddmmU; ++u) same[u] = u > 0 && row[u] == row[u - 1];
      for (int k = 4 * sub; k < Kf; k += 4 * LPE) {
        float4 a[kSddmmU], b[kSddmmU];
#pragma unroll
        for (int u = 0; u < kSddmmU; ++u) {
          if (!same[u]) a[u] = ld4(X + (int64_t)row[u] * ldx + k);
          b[u] = ld4(Y + (int64_t)col[u] * ldy + k);
        }
#pragma unroll
        for (int u = 1; u < kSddmmU; ++u)
          if (same[u]) a[u] = a[u - 1];
#pragma unroll
        for (int u = 0; u < kSddmmU; ++u) s[u] += a[u].x * b[u].x + a[u].y * b[u].y + a[u].z * b[u].z + a[u].w * b[u].w;
      }
    } else {
      for (int k = sub; k < Kf; k += LPE) {
#pragma unroll
        for (int u = 0; u < kSddmmU; ++u) s[u] += X[(int64_t)row[u] * ldx + k] * Y[(int64_t)col[u] * ldy + k];
      }
    }
#pragma unroll
    for (int u = 0; u < kSddmmU; ++u) {
#pragma unroll
      for (int off = LPE / 2; off > 0; off >>= 1) s[u] += __shfl_xor(s[u], off, LPE);
    }
    if (sub == 0) {
#pragma unroll
      for (int u = 0; u < kSddmmU; ++u)
        if (e0 + u < p.n1d) out[dst[u]] = s[u];
    }
    return;
  }
  const int b = wave - waves1d;
  if (b >= p.nblk) return;
  const int g = lane >> 4, j = lane & 15;
  const float* xr = X + (int64_t)p.blk_rows[b * 16 + j] * ldx;
  const float* yr = Y + (int64_t)p.blk_cols[b * 16 + j] * ldy;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < Kf; kb += 16) {
    const int k = kb + 4 * g;
    float4 xa, ya;
    if (vec && k + 3 < Kf) {
      xa = ld4(xr + k);
      ya = ld4(yr + k);
    } else {
      xa.x = k < Kf ? xr[k] : 0.f;         ya.x = k < Kf ? yr[k] : 0.f;
      xa.y = k + 1 < Kf ? xr[k + 1] : 0.f; ya.y = k + 1 < Kf ? yr[k + 1] : 0.f;
      xa.z = k + 2 < Kf ? xr[k + 2] : 0.f; ya.z = k + 2 < Kf ? yr[k + 2] : 0.f;
      xa.w = k + 3 < Kf ? xr[k + 3] : 0.f; ya.w = k + 3 < Kf ? yr[k + 3] : 0.f;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, ya.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, ya.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, ya.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, ya.w, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) blk[wib][(4 * g + i) * 16 + j] = acc[i];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int s0 = p.blk_ptr[b], s1 = p.blk_ptr[b + 1];
  for (int s = s0 + lane; s < s1; s += 64) out[p.blk_out[s]] = blk[wib][p.blk_pos[s]];
}

hipError_t launch_hyb_sddmm(const SddmmPlanDev& p, const float* X,
                            int64_t ldx, const float* Y, int64_t ldy, int Kf,
                            float* out, hipStream_t s) {
  const bool vec = (Kf % 4 == 0) && (ldx % 4 == 0) && (ldy % 4 == 0) &&
                   ((reinterpret_cast<uintptr_t>(X) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(Y) & 15) == 0);
  int lpe = vec ? Kf / 4 : Kf;
  if (lpe > 16) lpe = 16;
  if (lpe < 1) lpe = 1;
  int l = 1;
  while (l * 2 <= lpe) l *= 2;
  lpe = l;
  const int epw = 64 / lpe * kSddmmU;
  const int waves1d = (p.n1d + epw - 1) / epw;
  const int64_t waves = (int64_t)waves1d + p.nblk;
  if (waves == 0) return hipSuccess;
  dim3 grid((unsigned)((waves + 3) / 4));
#define STILE_HSDDMM_CASE(L)                                                  \
  case L:                                                                     \
    hipLaunchKernelGGL(hyb_sddmm_kernel<L>, grid, dim3(256), 0, s, p, X, ldx, \
                       Y, ldy, Kf, out, waves1d, vec);                        \
    break;
  switch (lpe) {
    STILE_HSDDMM_CASE(1)
    STILE_HSDDMM_CASE(2)
    STILE_HSDDMM_CASE(4)
    STILE_HSDDMM_CASE(8)
    STILE_HSDDMM_CASE(16)
  }
#undef STILE_HSDDMM_CASE
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row softmax over CSR rows: one wave per row; max and sum reduced with xor
// shuffles across the 64 lanes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_softmax_kernel(int M, const int32_t* __restrict__ indptr,
                                                          const float* in, float* out, float scale) {
  const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (row >= M) return;
  const int lane = threadIdx.x & 63;
  const int p0 = indptr[row], p1 = indptr[row + 1];
  if (p0 >= p1) return;
  float m = -INFINITY;
  for (int p = p0 + lane; p < p1; p += 64) m = fmaxf(m, scale * in[p]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  float s = 0.f;
  for (int p = p0 + lane; p < p1; p += 64) s += __expf(scale * in[p] - m);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  const float inv = 1.f / s;
  for (int p = p0 + lane; p < p1; p += 64) out[p] = __expf(scale * in[p] - m) * inv;
}

hipError_t launch_row_softmax(int M, const int32_t* indptr, const float* in, float* out, float scale,
                              hipStream_t s) {
  if (M == 0) return hipSuccess;
  const int64_t waves = M;
  hipLaunchKernelGGL(row_softmax_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, M, indptr, in, out,
                     scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused sparse attention: out = softmax(scale * Q K^T on the mask) V for BF16
// Q, K, V (rows x 64, leading dimension ld), FP32 out, over a plan built by
// stile_plan_create_attention (block-diagonal masks; tc_mask marks the mask
// entries among each chunk's k-slots).  One wave per window of 16 query rows;
// lane l holds query q = l & 15 and lane group g = l >> 4.
//
// Per chunk of 32 key columns (k-slots):
//   S^T = K Q^T as two 16x16 tiles (keys x queries), two v_mfma_f32_16x16x32_
//     bf16 each over the 64 features (lane group g carries features
//     32s + 8g .. +7 of step s for both operands, one 16-byte load each);
//     lane (g, q) then holds S[q][slot 16T + 4g + i], i = 0..3, T = 0, 1;
//   online softmax per query (running max m and sum l; the chunk max is
//     reduced over the four lanes of a query with xor shuffles);
//   O^T += V^T P^T as four d-tiles of one v_mfma_f32_16x16x32_bf16: the
//     lane's eight probabilities are its B operand as they are (k-slot e of
//     lane group g = slot 4g + e for e < 4, 16 + 4g + e - 4 otherwise); the
//     V rows of the chunk go through LDS transposed (Vt[d][slot]) so the A
//     operand is two ds_read_b64 per d-tile.
// Sparse entries of a window row (the plan's sparse part) follow one at a
// time: the four lanes of the query share the dot product and the update.
// Out row r = O[r] / l[r] (0 for a row without entries).
// ---------------------------------------------------------------------------
constexpr int kAttD = 64;
// Vt row stride in uint2 (4 BF16): 9, not 8, so that the eight d rows one
// ds_write_b64 touches start on different LDS banks
constexpr int kVtStride = 9;

__global__ __launch_bounds__(256) void attention_fused_kernel(SpmmPlanDev p, const __bf16* __restrict__ Qm,
                                                              const __bf16* __restrict__ Km,
                                                              const __bf16* __restrict__ Vm, int64_t ld,
                                                              float scale, float* __restrict__ out, int64_t ldo) {
  __shared__ uint2 vt_all[4][kAttD * kVtStride];  // per wave: Vt[d][quad of slots], 4 BF16 each
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wv;
  if (w >= p.W) return;
  uint2* vt = vt_all[wv];
  const int q = lane & 15, g = lane >> 4;
  const int row = p.win_rows[w * 16 + q];
  const int64_t qrow = row >= 0 ? row : 0;  // padding query rows compute on row 0, store nothing
  const bf16x8 qb0 = *reinterpret_cast<const bf16x8*>(Qm + qrow * ld + 8 * g);
  const bf16x8 qb1 = *reinterpret_cast<const bf16x8*>(Qm + qrow * ld + 32 + 8 * g);
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float m = -INFINITY, lsum = 0.f;
  const int c0 = p.tc_ptr[w], c1 = p.tc_ptr[w + 1];
  const int a = lane >> 3, o = lane & 7;
  // the column ids this lane needs from a chunk: K rows of slots q and
  // 16 + q (its A-operand rows), V rows of slots 4a + u (a = lane >> 3)
  int ids[6];
  auto load_ids = [&](int c) {
    const int32_t* cc = p.tc_cols + (int64_t)c * 32;
    ids[0] = cc[(q & 3) * 8 + (q >> 2)];
    ids[1] = cc[(q & 3) * 8 + 4 + (q >> 2)];
#pragma unroll
    for (int u = 0; u < 4; ++u) ids[2 + u] = cc[u * 8 + a];
  };
  // K and V of chunk c + 1 are requested before chunk c is processed (two
  // register sets, loop unrolled by two so neither is copied); column ids run
  // one chunk further ahead.
  struct AttRegs {
    bf16x8 k00, k01, k10, k11;
    uint4 vv[4];
    uint32_t mw;
  };
  auto load_regs = [&](int c, AttRegs& r) {
    r.k00 = *reinterpret_cast<const bf16x8*>(Km + (int64_t)ids[0] * ld + 8 * g);
    r.k01 = *reinterpret_cast<const bf16x8*>(Km + (int64_t)ids[0] * ld + 32 + 8 * g);
    r.k10 = *reinterpret_cast<const bf16x8*>(Km + (int64_t)ids[1] * ld + 8 * g);
    r.k11 = *reinterpret_cast<const bf16x8*>(Km + (int64_t)ids[1] * ld + 32 + 8 * g);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      r.vv[u] = *reinterpret_cast<const uint4*>(Vm + (int64_t)ids[2 + u] * ld + 8 * o);
    r.mw = p.tc_mask[(int64_t)c * 16 + q];
  };
  auto process = [&](const AttRegs& r) {
    const bf16x8 &k00 = r.k00, &k01 = r.k01, &k10 = r.k10, &k11 = r.k11;
    const uint4* vv = r.vv;
    const uint32_t mw = r.mw;
    f32x4 st0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(k00, qb0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    st0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(k01, qb1, st0, 0, 0, 0);
    f32x4 st1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(k10, qb0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    st1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(k11, qb1, st1, 0, 0, 0);
    // Vt[d][slot] for d = 8o + e, slots 4a .. 4a + 3
    {
      const uint32_t* v0 = reinterpret_cast<const uint32_t*>(&vv[0]);
      const uint32_t* v1 = reinterpret_cast<const uint32_t*>(&vv[1]);
      const uint32_t* v2 = reinterpret_cast<const uint32_t*>(&vv[2]);
      const uint32_t* v3 = reinterpret_cast<const uint32_t*>(&vv[3]);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int h = e & 1, dw = e >> 1;
        vt[(8 * o + e) * kVtStride + a] = make_uint2(half_sel(v0[dw], v1[dw], h), half_sel(v2[dw], v3[dw], h));
      }
    }
    // online softmax
    float sv[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      sv[i] = (mw >> (4 * g + i)) & 1u ? st0[i] : -INFINITY;
      sv[4 + i] = (mw >> (16 + 4 * g + i)) & 1u ? st1[i] : -INFINITY;
    }
    float cm = sv[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) cm = fmaxf(cm, sv[i]);
    cm = fmaxf(cm, __shfl_xor(cm, 16));
    cm = fmaxf(cm, __shfl_xor(cm, 32));
    const float m_new = fmaxf(m, cm);
    const float alpha = m_new == -INFINITY ? 1.f : __expf((m - m_new) * scale);
    uint32_t pk[4];
    float psum = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float pa = sv[2 * i] == -INFINITY ? 0.f : __expf((sv[2 * i] - m_new) * scale);
      const float pb = sv[2 * i + 1] == -INFINITY ? 0.f : __expf((sv[2 * i + 1] - m_new) * scale);
      pk[i] = bf16_pair(pa, pb);
      // the sum uses the BF16-rounded probabilities the MFMA multiplies
      psum += __builtin_bit_cast(float, pk[i] << 16) + __builtin_bit_cast(float, pk[i] & 0xffff0000u);
    }
    lsum = lsum * alpha + psum;
    m = m_new;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] *= alpha;
    // LDS writes of this wave complete before its reads below
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    const bf16x8 pb = __builtin_bit_cast(bf16x8, make_uint4(pk[0], pk[1], pk[2], pk[3]));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int d = 16 * t + q;
      const uint2 lo = vt[d * kVtStride + g], hi = vt[d * kVtStride + 4 + g];
      const bf16x8 va = __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, pb, acc[t], 0, 0, 0);
    }
    // reads of Vt done before the next chunk overwrites it
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    };
  if (c0 < c1) {
    const int last = c1 - 1;
    AttRegs R0, R1;
    load_ids(c0);
    load_regs(c0, R0);
    load_ids(min(c0 + 1, last));
    int c = c0;
    while (c + 2 <= c1) {
      load_regs(min(c + 1, last), R1);
      load_ids(min(c + 2, last));
      process(R0);
      ++c;
      load_regs(min(c + 1, last), R0);
      load_ids(min(c + 2, last));
      process(R1);
      ++c;
    }
    if (c < c1) process(R0);
  }
  // sparse entries of the window rows, one per step for all 16 queries
  {
    const int r0 = p.res_ptr[w * 16 + q], r1 = p.res_ptr[w * 16 + q + 1];
    int n = r1 - r0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) n = max(n, __shfl_xor(n, off));
    for (int j = 0; j < n; ++j) {
      const bool on = r0 + j < r1;
      const int col = on ? p.res_cols[r0 + j] : 0;
      const bf16x8 ka = *reinterpret_cast<const bf16x8*>(Km + (int64_t)col * ld + 8 * g);
      const bf16x8 kb = *reinterpret_cast<const bf16x8*>(Km + (int64_t)col * ld + 32 + 8 * g);
      float dot = 0.f;
#pragma unroll
      for (int e = 0; e < 8; ++e) dot += (float)qb0[e] * (float)ka[e] + (float)qb1[e] * (float)kb[e];
      dot += __shfl_xor(dot, 16);
      dot += __shfl_xor(dot, 32);
      const float sval = on ? dot : -INFINITY;
      const float m_new = fmaxf(m, sval);
      const float alpha = m_new == -INFINITY ? 1.f : __expf((m - m_new) * scale);
      const float pr = sval == -INFINITY ? 0.f : __expf((sval - m_new) * scale);
      const float prb = __builtin_bit_cast(float, bf16_pair(pr, 0.f) << 16);
      lsum = lsum * alpha + (g == 0 ? prb : 0.f);
      m = m_new;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint2 vr = *reinterpret_cast<const uint2*>(Vm + (int64_t)col * ld + 16 * t + 4 * g);
        acc[t][0] = acc[t][0] * alpha + prb * __builtin_bit_cast(float, vr.x << 16);
        acc[t][1] = acc[t][1] * alpha + prb * __builtin_bit_cast(float, vr.x & 0xffff0000u);
        acc[t][2] = acc[t][2] * alpha + prb * __builtin_bit_cast(float, vr.y << 16);
        acc[t][3] = acc[t][3] * alpha + prb * __builtin_bit_cast(float, vr.y & 0xffff0000u);
      }
    }
  }
  float ltot = lsum + __shfl_xor(lsum, 16);
  ltot += __shfl_xor(ltot, 32);
  if (row < 0) return;
  const float inv = ltot > 0.f ? 1.f / ltot : 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    st4(out + (int64_t)row * ldo + 16 * t + 4 * g,
        make_float4(acc[t][0] * inv, acc[t][1] * inv, acc[t][2] * inv, acc[t][3] * inv));
}

hipError_t launch_attention(const SpmmPlanDev& p, const void* Q, const void* K, const void* V, int64_t ld,
                            float scale, float* out, int64_t ldo, hipStream_t s) {
  if (p.W == 0) return hipSuccess;
  hipLaunchKernelGGL(attention_fused_kernel, dim3((unsigned)((p.W + 3) / 4)), dim3(256), 0, s, p,
                     static_cast<const __bf16*>(Q), static_cast<const __bf16*>(K), static_cast<const __bf16*>(V),
                     ld, scale, out, ldo);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ELL-bucket SpMM (the SparseTIR hyb comparator): every ELL row of a bucket
// holds `width` (column, value) slots; C[row] += sum over slots v * B[col].
// One 16-lane group per ELL row and 64 columns; rows of C appear in several
// buckets (column parts, split long rows), so results are added with
// atomicAdd into a zero-filled C.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ell_bucket_spmm_kernel(int rows, int width, const int32_t* __restrict__ orow,
                                                              const int32_t* __restrict__ cols,
                                                              const float* __restrict__ vals,
                                                              const float* __restrict__ B, int64_t ldb, int N,
                                                              float* __restrict__ C, int64_t ldc, int nC,
                                                              bool vec) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t item = wave * 4 + (lane >> 4);
  if (item >= (int64_t)rows * nC) return;
  const int r = (int)(item / nC);
  const int col = (int)(item - (int64_t)r * nC) * 64 + 4 * (lane & 15);
  if (col >= N) return;
  const int32_t* cr = cols + (int64_t)r * width;
  const float* vr = vals + (int64_t)r * width;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  int k = 0;
  for (; k + 4 <= width; k += 4) {
    const int k0 = cr[k], k1 = cr[k + 1], k2 = cr[k + 2], k3 = cr[k + 3];
    const float v0 = vr[k], v1 = vr[k + 1], v2 = vr[k + 2], v3 = vr[k + 3];
    const float4 b0 = load_b4(B, ldb, k0, col, N, vec), b1 = load_b4(B, ldb, k1, col, N, vec);
    const float4 b2 = load_b4(B, ldb, k2, col, N, vec), b3 = load_b4(B, ldb, k3, col, N, vec);
    s.x += v0 * b0.x + v1 * b1.x + v2 * b2.x + v3 * b3.x;
    s.y += v0 * b0.y + v1 * b1.y + v2 * b2.y + v3 * b3.y;
    s.z += v0 * b0.z + v1 * b1.z + v2 * b2.z + v3 * b3.z;
    s.w += v0 * b0.w + v1 * b1.w + v2 * b2.w + v3 * b3.w;
  }
  for (; k < width; ++k) {
    const float v = vr[k];
    const float4 b = load_b4(B, ldb, cr[k], col, N, vec);
    s.x += v * b.x; s.y += v * b.y; s.z += v * b.z; s.w += v * b.w;
  }
  float* c = C + (int64_t)orow[r] * ldc + col;
  atomicAdd(c, s.x);
  if (col + 1 < N) atomicAdd(c + 1, s.y);
  if (col + 2 < N) atomicAdd(c + 2, s.z);
  if (col + 3 < N) atomicAdd(c + 3, s.w);
}

hipError_t launch_ell_bucket_spmm(int rows, int width, const int32_t* orow, const int32_t* cols, const float* vals,
                                  const float* B, int64_t ldb, int N, float* C, int64_t ldc, hipStream_t s) {
  if (rows == 0 || N == 0) return hipSuccess;
  const bool vec = (N % 4 == 0) && (ldb % 4 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
  const int nC = (N + 63) / 64;
  const int64_t items = (int64_t)rows * nC;
  const int64_t blocks = (items + 15) / 16;
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ell_bucket_spmm_kernel, dim3((unsigned)blocks), dim3(256), 0, s, rows, width, orow, cols, vals,
                     B, ldb, N, C, ldc, nC, vec);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Value refresh.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void set_values_kernel(const float* __restrict__ vals,
                                                         const int32_t* __restrict__ tc_ids, int64_t n_tc,
                                                         void* tc_vals, int mode,
                                                         const int32_t* __restrict__ res_ids, int64_t n_res,
                                                         float* res_vals, const int32_t* __restrict__ sp_ids,
                                                         int64_t n_sp, int2* sp_ent) {
  const int64_t n = n_tc + n_res + n_sp;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n_tc) {
      const int id = tc_ids[i];
      if (id == kKeepId) continue;
      const float v = id >= 0 ? vals[id] : 0.f;
      if (mode == 1)
        reinterpret_cast<__bf16*>(tc_vals)[i] = (__bf16)v;
      else
        reinterpret_cast<float*>(tc_vals)[i] = v;
    } else if (i < n_tc + n_res) {
      const int64_t j = i - n_tc;
      res_vals[j] = vals[res_ids[j]];
    } else {
      // the value half of a round entry; padding entries (id -1) stay 0
      const int64_t j = i - n_tc - n_res;
      const int id = sp_ids[j];
      reinterpret_cast<float*>(sp_ent)[2 * j + 1] = id >= 0 ? vals[id] : 0.f;
    }
  }
}

hipError_t launch_set_values(const SpmmPlanDev& p, const float* vals, const int32_t* tc_ids, int64_t n_tc,
                             const int32_t* res_ids, int64_t n_res, const int32_t* sp_ids, int64_t n_sp,
                             hipStream_t s) {
  const int64_t n = n_tc + n_res + n_sp;
  if (n == 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(set_values_kernel, dim3((unsigned)blocks), dim3(256), 0, s, vals, tc_ids, n_tc,
                     const_cast<void*>(p.tc_vals), p.blk_dtype, res_ids, n_res, const_cast<float*>(p.res_vals),
                     sp_ids, n_sp, const_cast<int2*>(p.sp_ent));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Zero fill.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_zero_2d_kernel(float* __restrict__ p, int rows, int cols,
                                                           int64_t ld) {
  const int64_t n = (int64_t)rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    p[r * ld + c] = 0.f;
  }
}

// C rows of windows [0, nwin) set to zero (the huge windows' rows, before
// their parts add into them).  One block per row slot.
__global__ __launch_bounds__(256) void zero_window_rows_kernel(const int32_t* __restrict__ win_rows, float* C,
                                                               int64_t ldc, int N) {
  const int row = win_rows[blockIdx.x];
  if (row < 0) return;
  float* c = C + (int64_t)row * ldc;
  for (int i = threadIdx.x; i < N; i += blockDim.x) c[i] = 0.f;
}

hipError_t launch_zero_window_rows(const SpmmPlanDev& p, int nwin, float* C, int64_t ldc, int N, hipStream_t s) {
  if (nwin == 0 || N == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_window_rows_kernel, dim3((unsigned)(nwin * 16)), dim3(256), 0, s, p.win_rows, C, ldc, N);
  return hipGetLastError();
}

hipError_t launch_fill_zero_2d(float* p, int rows, int cols, int64_t ld, hipStream_t s) {
  const int64_t n = (int64_t)rows * cols;
  if (n == 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_zero_2d_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, rows, cols, ld);
  return hipGetLastError();
}

hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n <= INT32_MAX) return launch_fill_zero_2d(p, 1, (int)n, n, s);
  return hipErrorInvalidValue;
}

}  // namespace stile
