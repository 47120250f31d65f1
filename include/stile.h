/*
 * stile.h -- C ABI of libstile_hip.so, the MI355X (gfx950) kernel library for
 * STile's hybrid sparse-tile SpMM and SDDMM.
 *
 * Everything here is plain C: pointers, sizes, status codes.  No torch, no HIP
 * types (a hipStream_t is passed as void*; NULL means the default stream).
 *
 * Ownership: the caller owns every device pointer passed to a *_run call
 * (B, C, X, Y, out).  A StilePlan owns device copies of the packed tile
 * buffers; it is created once per (matrix, tile set) and reused for any N.
 * *_run calls never allocate, never synchronise and are safe inside HIP graph
 * capture.  Status: 0 = ok, otherwise a STILE_E* code; stile_last_error()
 * returns a thread-local message for the last failure on the calling thread.
 *
 * What each entry point replaces in the reference (STile, MY_sparse/):
 *   - the TVM PackedFunc `f` built by tvm.build in my_fuse_formats.py:3237
 *     and called as f(*args_nd) at my_fuse_formats.py:4173 (SpMM) and in
 *     measure_latency_BSDDMM (my_fuse_formats.py:4292-4380, SDDMM), whose
 *     kernel body is the horizontally fused templates of
 *     my_fuse_formats.py:1162-1291 (ELL), :1444-1799 (TensorCore),
 *     :2263-2317 (1D SDDMM), :2324-2415 (TC SDDMM), fused by :2426-2560;
 *   - the argument lists built by prepare_inputs_BSPMM
 *     (my_fuse_formats.py:3321-3537) and prepare_inputs_BSDDMM
 *     (my_fuse_formats.py:3544-3678), which are the StileSpmmPack /
 *     StileSddmmPack layouts below, group by group.
 */
#ifndef STILE_H_
#define STILE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STILE_OK 0
#define STILE_E_INVALID 1   /* bad argument: null pointer, bad size, index out of range */
#define STILE_E_HIP 2       /* a HIP runtime call failed */
#define STILE_E_NOMEM 3     /* device or host allocation failed at plan creation */
#define STILE_E_UNSUPPORTED 4

/* Arithmetic type of the dense (MFMA) blocks of a plan. */
#define STILE_BLK_F32 0   /* v_mfma_f32_16x16x4_f32: FP32 products, FP32 accumulation */
#define STILE_BLK_BF16 1  /* v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation */
#define STILE_BLK_F16 2   /* v_mfma_f32_16x16x32_f16 with a per-block FP32 scale */

/* ---------------------------------------------------------------------------
 * SpMM packed groups (layouts of prepare_inputs_BSPMM).
 * ------------------------------------------------------------------------- */

/* One "sparse_template_ell" group (my_fuse_formats.py:3372-3419).
 * T tiles of m_blk (I) rows x nnz_cols (K) slots.
 *   a                   [T][I*K]  values; padding slots hold 0
 *   indices_k           [T*I][K]  original column of each slot
 *   indices_reordered_i [T*I]     original output row of each tile row;
 *                                 == M (the sentinel row) for rows whose
 *                                 products must be discarded
 *                                 (my_branch_and_bound.py:1682-1700). */
typedef struct {
  int32_t m_blk;
  int32_t nnz_cols;
  int32_t num_tiles;
  int32_t real_atomic; /* the group's atomic flag (my_fuse_formats.py:57-78) */
  const float* a;
  const int32_t* indices_k;
  const int32_t* indices_reordered_i;
  const int32_t* a_ids; /* optional [T][I*K]: index of each slot's value in A's
                           CSR value array, -1 for padding; lets
                           stile_plan_set_values refresh the values (NULL: static) */
} StileEllGroup;

/* One "TensorCore_template" group (my_fuse_formats.py:3424-3504).
 * nnzb dense tiles of 16 rows x group_size (K) columns in mb row windows.
 *   a         [nnzb][16][K] values (zero where the tile covers no nonzero)
 *   indptr    [mb+1]        tiles of window slot s are indptr[s]..indptr[s+1]
 *   indices   [nnzb][K]     original column of each tile column
 *                           (k_notsorted: [nnzb] 16-wide block column ids)
 *   i_indices rows_2d ? [mb][16] original rows of each window slot (padded
 *                       with the window's last row)
 *                     : [mb] window ids w, rows 16w..16w+15 */
typedef struct {
  int32_t group_size;
  int32_t num_windows;
  int32_t nnzb;
  int32_t rows_2d;
  int32_t k_notsorted;
  int32_t real_atomic;
  const float* a;
  const int32_t* i_indices;
  const int32_t* indptr;
  const int32_t* indices;
  const int32_t* a_ids; /* optional [nnzb][16][K], as StileEllGroup.a_ids */
} StileTcGroup;

typedef struct {
  int32_t M; /* rows of A (and of C) */
  int32_t K; /* columns of A == rows of B */
  int32_t num_ell;
  int32_t num_tc;
  const StileEllGroup* ell;
  const StileTcGroup* tc;
} StileSpmmPack;

/* ---------------------------------------------------------------------------
 * SDDMM packed groups (layouts of prepare_inputs_BSDDMM).
 * ------------------------------------------------------------------------- */

/* One "1D_sddmm" group (my_fuse_formats.py:3565-3598): len entries
 * out[remap[e]] = <X[mid_idx[e]], Y[j_idx[e]]>.  Trailing padding repeats
 * the last entry, which is harmless (same value, same slot). */
typedef struct {
  int32_t len;
  const int32_t* mid_idx;
  const int32_t* j_idx;
  const int32_t* remap;
} StileSddmm1DGroup;

/* One "TC_sddmm" group (my_fuse_formats.py:3602-3656): nnzb dense 16x16
 * output blocks X[i_indices[b]] . Y[j_indices[b]]^T, of which the sampled
 * entries s in indptr[b]..indptr[b+1] are written:
 *   out[indices[s]] = block[indices_dense[s] / blk_k][indices_dense[s] % blk_k] */
typedef struct {
  int32_t nnzb;
  int32_t blk_k;              /* tile_sizes[2][1]; 16 */
  const int32_t* i_indices;   /* [nnzb][16] */
  const int32_t* j_indices;   /* [nnzb][blk_k] */
  const int32_t* indptr;      /* [nnzb+1] */
  const int32_t* indices_dense;
  const int32_t* indices;
} StileSddmmTcGroup;

typedef struct {
  int32_t M;      /* rows of the sampling pattern (rows of X) */
  int32_t N;      /* columns of the pattern (rows of Y) */
  int32_t nnz;    /* length of the output array */
  int32_t num_1d;
  int32_t num_tc;
  const StileSddmm1DGroup* g1d;
  const StileSddmmTcGroup* tc;
} StileSddmmPack;

typedef struct StilePlan StilePlan;

typedef struct {
  int32_t kind;           /* 1 = SpMM, 2 = SDDMM */
  int32_t blk_dtype;
  int32_t M, K;           /* SpMM: A is M x K; SDDMM: pattern is M x K(=N cols) */
  int32_t num_windows;    /* SpMM: 16-row output windows */
  int32_t num_chunks;     /* SpMM: 16x32 MFMA chunks; SDDMM: 16x16 blocks */
  int64_t block_nnz;      /* nonzeros carried by MFMA blocks */
  int64_t sparse_nnz;     /* nonzeros carried by the FP32 1D/ELL path */
  int64_t device_bytes;   /* bytes of plan buffers resident on the device */
  int32_t hub_rows;       /* rows with > 4096 sparse nonzeros: each 256-term
                             piece of such a row is summed with Neumaier
                             compensation, the pieces then plainly added */
  int32_t num_split;      /* SpMM: pieces of rows with > 256 sparse nonzeros,
                             added to C by the fix-up kernel */
} StilePlanInfo;

/* ---------------------------------------------------------------------------
 * Entry points.
 * ------------------------------------------------------------------------- */

/* Plain CSR SpMM, FP32 throughout: C[M][N] = A * B.  The comparator kernel
 * (the reference's csrmm template, my_fuse_formats.py:1106-1152, which its
 * search excludes, my_search_formats.py:332-333).  All pointers are device
 * pointers; B is K x N with row stride ldb, C is M x N with row stride ldc.
 * Rows with more than 4096 nonzeros are summed with compensated (Neumaier)
 * FP32 arithmetic. */
int stile_csr_spmm_f32(int32_t M, int32_t K, int32_t N, const int32_t* indptr,
                       const int32_t* indices, const float* vals,
                       const float* B, int64_t ldb, float* C, int64_t ldc,
                       void* stream);

/* Plain CSR SDDMM, FP32: out[p] = <X[row(p)], Y[indices[p]]> for every
 * nonzero p < nnz (== indptr[M]) of the M x N pattern (indptr, indices),
 * feature length Kf.  Replaces the reference's golden dgl.ops.u_dot_v
 * (my_fuse_formats.py:4331). */
int stile_csr_sddmm_f32(int32_t M, int32_t N, int32_t nnz, int32_t Kf,
                        const int32_t* indptr, const int32_t* indices,
                        const float* X, int64_t ldx, const float* Y,
                        int64_t ldy, float* out, void* stream);

/* One bucket of the ELL-bucket ("hyb") format of SparseTIR's comparator
 * (SPMM/SPMMbench_spmm_hyb.py:156-316, column_part_hyb): `rows` ELL rows of
 * `width` slots; device arrays row[rows] (output row of each ELL row),
 * cols[rows][width], vals[rows][width] (0 in padding slots). */
typedef struct StileEllBucket {
  int32_t rows;
  int32_t width;
  const int32_t* row;
  const int32_t* cols;
  const float* vals;
} StileEllBucket;

/* C (M x N) = sum over buckets of their ELL products (FP32; C is zero-filled,
 * then each bucket adds with atomicAdd, one launch per bucket).  The
 * comparator standing in for the reference's SparseTIR hyb baseline. */
int stile_ell_buckets_spmm_f32(const StileEllBucket* buckets, int32_t count, int32_t M, int32_t N, const float* B,
                               int64_t ldb, float* C, int64_t ldc, void* stream);

/* Row softmax over a CSR pattern (FP32): for every row i,
 * out[p] = exp(scale*in[p] - m_i) / sum_q exp(scale*in[q] - m_i) over the
 * row's nonzeros p, q (m_i = max of scale*in over the row); empty rows write
 * nothing.  The middle step of sparse attention: SDDMM(Q, K) -> softmax ->
 * SpMM with V (in the reference only as a mask pattern,
 * gen_formats_v2.py:2926-3003).  in and out may alias. */
int stile_csr_row_softmax_f32(int32_t M, const int32_t* indptr, const float* in, float* out, float scale,
                              void* stream);

/* Build a hybrid SpMM plan from the packed groups (host pointers; copied).
 * The plan regroups every group's products by 16-row output window, so C
 * needs no zero-fill and most rows are stored once with a fixed summation
 * order.  Two kinds of rows are added with atomicAdd, in completion order,
 * unless the run passes STILE_RUN_DETERMINISTIC: the pieces of rows with more
 * than 256 sparse nonzeros, and the parts of windows with more than 256
 * dense chunks (such windows are zeroed first). */
int stile_plan_create_spmm(const StileSpmmPack* pack, int32_t blk_dtype,
                           StilePlan** out);

/* Build a hybrid SDDMM plan from the packed groups (host pointers; copied). */
int stile_plan_create_sddmm(const StileSddmmPack* pack, int32_t blk_dtype,
                            StilePlan** out);

/* C[M][N] = A * B with the plan's tile set.  B: K x N FP32 (row stride ldb),
 * C: M x N FP32 (row stride ldc).  Every row of C is written.  Same as
 * stile_spmm_run_ex(plan, STILE_BLK_F32, B, ldb, N, C, ldc, 0, NULL, 0,
 * stream). */
int stile_spmm_run(const StilePlan* plan, const float* B, int64_t ldb,
                   int32_t N, float* C, int64_t ldc, void* stream);

/* Run flag: sum the pieces of rows with more than 256 sparse nonzeros, and
 * the parts of windows with more than 256 dense chunks, in a fixed order
 * through the caller's workspace instead of with atomicAdd in completion
 * order: results are then bit-identical run to run.  (Without it only those
 * rows can differ in their last bits between runs.) */
#define STILE_RUN_DETERMINISTIC 1

/* Bytes of workspace stile_spmm_run_ex needs for width N and these flags
 * (0 without STILE_RUN_DETERMINISTIC, or when the plan has no split rows or
 * huge windows). */
int stile_spmm_workspace_bytes(const StilePlan* plan, int32_t N, int32_t flags, int64_t* bytes);

/* C[M][N] = A * B with B's elements of type b_dtype (STILE_BLK_F32, _BF16 or
 * _F16: the reference keeps B in the run dtype, FP16 by default,
 * MY_sparse/my_fuse_formats.py:3529).  B: K x N with row stride ldb
 * (elements); C: M x N FP32.  The sparse (1D/ELL) products are FP32 on B's
 * exact values; the dense blocks run on the plan's MFMA type (a 16-bit B
 * staged for BF16 blocks is rounded to BF16 only if it is FP16 or FP32).
 * flags: 0 or STILE_RUN_DETERMINISTIC; workspace: a device buffer of at
 * least stile_spmm_workspace_bytes bytes, 16-byte aligned (may be NULL when
 * that is 0).  Launches kernels only (no allocation, no synchronisation). */
int stile_spmm_run_ex(const StilePlan* plan, int32_t b_dtype, const void* B, int64_t ldb, int32_t N, float* C,
                      int64_t ldc, int32_t flags, void* workspace, int64_t workspace_bytes, void* stream);

/* Size of the static device scratch of stile_spmm_run_noalloc (per GPU). */
#define STILE_NOALLOC_SCRATCH_BYTES (16 << 20)

/* stile_spmm_run_ex that allocates nothing and takes no workspace: with
 * STILE_RUN_DETERMINISTIC the workspace is a static device array of
 * STILE_NOALLOC_SCRATCH_BYTES in the library (one per GPU, held until the
 * process ends), so a call whose stile_spmm_workspace_bytes exceeds that is
 * rejected with STILE_E_INVALID before any launch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (several streams or threads) race. */
int stile_spmm_run_noalloc(const StilePlan* plan, int32_t b_dtype, const void* B, int64_t ldb, int32_t N, float* C,
                           int64_t ldc, int32_t flags, void* stream);

/* out[p] = <X[row(p)], Y[col(p)]> for every nonzero p of the plan's pattern;
 * X: M x Kf, Y: N x Kf FP32 with row strides ldx, ldy; out has pack->nnz
 * entries in the pattern's CSR order (the reference's remap ids). */
int stile_sddmm_run(const StilePlan* plan, const float* X, int64_t ldx,
                    const float* Y, int64_t ldy, int32_t Kf, float* out,
                    void* stream);

/* Many uneven matrices in one launch: build one plan over the block-diagonal
 * composition of count packs.  Matrix i's rows are offset by the sum of the
 * previous packs' M and its columns by the sum of their K, so the caller
 * stacks the B_i (row-wise, same N) into one buffer and receives the C_i
 * stacked the same way from one stile_spmm_run.  Replaces the reference's
 * batched templates (my_fuse_formats.py:1296-1428, 1806-2261), which only
 * repeat one pattern over a batch of B. */
int stile_plan_create_spmm_batched(const StileSpmmPack* packs, int32_t count,
                                   int32_t blk_dtype, StilePlan** out);

/* Build options of an SpMM plan (a variant parameter of the kernels, like
 * the launch options below: the analogue of the reference's per-format
 * kernel parameters its benchmark scripts sweep,
 * SPMM/SPMMbench_our_method.py:288-320). */
typedef struct StileSpmmOptions {
  /* B-stationary plans (BF16 blocks, every matrix K <= 1024): a window's TC
   * columns holding at most this many nonzeros go to scatter chunks, whose
   * k-slots carry one entry each (4 bytes of A per entry instead of 32 per
   * column), when that shrinks the window's A bytes.  0 = none, at most 16. */
  int32_t scatter_max_nnz;
} StileSpmmOptions;

/* stile_plan_create_spmm_batched with build options (opts may be NULL: all
 * zero).  count = 1 builds a plan of one matrix. */
int stile_plan_create_spmm_ex(const StileSpmmPack* packs, int32_t count, int32_t blk_dtype,
                              const StileSpmmOptions* opts, StilePlan** out);

/* Launch options of an SpMM plan (kernel variant parameters, read by every
 * later run of the plan; results do not depend on them):
 *   STILE_OPT_KERNEL    0 (default) = the B-stationary kernel when the plan
 *                       allows it, 1 = the streaming kernel (refused with
 *                       STILE_E_UNSUPPORTED for plans with scatter chunks);
 *   STILE_OPT_LDS_LIST  -1 (default) = the plan's work-unit list for the
 *                       run's column-chunk count, 0..11 = that list;
 *   STILE_OPT_LDS_XCD   1 (default) = persistent B-stationary launches give
 *                       each XCD a contiguous range of work items, 0 = items
 *                       round-robin over the workgroups. */
#define STILE_OPT_KERNEL 1
#define STILE_OPT_LDS_LIST 2
#define STILE_OPT_LDS_XCD 3
int stile_plan_set_option(StilePlan* plan, int32_t option, int32_t value);

/* Same for SDDMM: pattern i's rows are offset by the previous M, its columns
 * by the previous N and its output ids by the previous nnz. */
int stile_plan_create_sddmm_batched(const StileSddmmPack* packs, int32_t count,
                                    int32_t blk_dtype, StilePlan** out);

/* Replace the plan's matrix values by vals[0..nnz) (device pointer, A's CSR
 * value order) without rebuilding the plan: every packed slot is refreshed
 * from the value id it was built with.  Requires a plan whose groups all
 * carried a_ids; FP32 and BF16 plans.  Not in the reference (its values are
 * the 0/1 position space); used for training, where A's values change every
 * step (the SDDMM backward multiplies by the incoming gradient). */
int stile_plan_set_values(StilePlan* plan, const float* vals, int64_t nnz, void* stream);

/* Fused sparse attention (no counterpart in the reference, which has these
 * masks only as SpMM patterns, gen_formats_v2.py:2926-3003, and batches one
 * pattern over many B, my_fuse_formats.py:1296-1428).  The plan is built like
 * stile_plan_create_spmm_batched from `count` square mask packs (block
 * diagonal; every packed slot is a mask entry) with BF16 blocks, plus per-chunk
 * entry bitmaps; it also runs as an SpMM plan. */
int stile_plan_create_attention(const StileSpmmPack* packs, int32_t count, StilePlan** out);

/* out[r] = sum_c softmax_c(scale * <Q[r], K[c]>) V[c] over the mask entries
 * (r, c) of the plan, for every row r of the stacked instances; rows without
 * entries get 0.  Q, K, V: BF16 device arrays (rows x d, leading dimension ld,
 * a multiple of 8, 16-byte aligned); out: FP32 (rows x d, leading dimension
 * ldo, a multiple of 4); d must be 64; scale > 0.  Launches one kernel on
 * `stream`, allocates nothing (capturable). */
int stile_attention_run(const StilePlan* plan, const void* Q, const void* K, const void* V, int64_t ld, int32_t d,
                        float scale, float* out, int64_t ldo, void* stream);

int stile_plan_info(const StilePlan* plan, StilePlanInfo* info);
int stile_plan_destroy(StilePlan* plan);

/* Message of the last failure on this thread ("" if none). */
const char* stile_last_error(void);
/* ABI version, bumped on any layout change of the structs above. */
int32_t stile_abi_version(void);  /* 3 */

#ifdef __cplusplus
}
#endif

#endif /* STILE_H_ */
