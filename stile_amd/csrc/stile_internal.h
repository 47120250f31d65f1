// Internal declarations shared by the kernel and plan translation units of
// libstile_hip.so.  Not part of the C ABI (include/stile.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stile {

// Hub rows: a row whose sparse (non-MFMA) part has more than this many
// nonzeros is summed with compensated (Neumaier) FP32 arithmetic, per
// kSplitNnz-term piece (the pieces are then added plainly).
constexpr int kHubNnz = 4096;

// Rows whose sparse part is longer than this are cut: the window pass sums
// the first kSplitNnz nonzeros, the rest go to pieces of kSplitNnz that a
// fix-up kernel adds with atomicAdd after the window pass (in deterministic
// mode: stores to a workspace, summed in piece order by a reduce kernel).
constexpr int kSplitNnz = 256;

// Value id of an A element that a value refresh leaves as it is (the row
// half of a scatter chunk's k-slot word).
constexpr int32_t kKeepId = -2;
// Windows with more than kBigChunks dense chunks are "big": the streaming
// kernel gives each (big window, 64 columns) a workgroup of kBigWaves waves,
// each wave a contiguous share of the chunks, and sums the partial results
// through LDS in wave order.
constexpr int kBigChunks = 8;
constexpr int kBigWaves = 8;
// Windows with more than kHugeChunks chunks (hub rows of power-law graphs)
// are cut into parts of at most kHugeChunks chunks, one workgroup each, that
// add their partial rows into C with atomicAdd (those rows are zeroed first):
// one workgroup per window would leave its last wave running long after the
// rest of the grid.  Windows up to kHugeChunks keep the fixed-order sums.
constexpr int kHugeChunks = 256;
// BF16 plans whose matrices all have K <= kLdsMaxK run their dense windows on
// the B-stationary kernel (B rows of one matrix staged in LDS, K rows of
// kLdsRowBytes: 64 BF16 columns and 32 bytes of padding, so K <= 1024 fits
// the 160 KiB of LDS of a gfx950 CU).
constexpr int kLdsMaxK = 1024;
constexpr int kLdsRowBytes = 160;
// Work units of the B-stationary kernel: kLdsLists lists of about
// kLdsTargets[t] units each; the plan picks one per column-chunk count
// nC <= kLdsPickMax by a makespan estimate (see build_lds_units).
constexpr int kLdsLists = 12;
constexpr int kLdsTargets[kLdsLists] = {1, 2, 4, 8, 16, 32, 64, 128, 192, 256, 384, 512};
constexpr int kLdsPickMax = 64;
// Persistent launches with nC <= kLdsXcdMaxNC column chunks group items by
// XCD: the GPU dispatches workgroup b to XCD b % kXcds, and XCD x takes the
// contiguous item range [start[x], start[x+1]) (unit-major, chunk-minor, the
// ranges balanced by estimated cost), so that a unit's column chunks and a
// matrix's units run on one XCD together and share its L2.
constexpr int kXcds = 8;
// Sparse-only rows of an LDS-kernel plan with at most this many (row, 64
// columns) tasks run as the LDS launch's tail instead of a launch of their own.
constexpr int kLdsTailMaxTasks = 1024;
constexpr int kLdsXcdMaxNC = 4;
struct LdsXcdRanges {
  int32_t start[kXcds + 1];
};

// Device view of a hybrid SpMM plan (window-major execution layout).
//
// The rows of C are partitioned into 16-row windows (the reference's TC row
// windows, my_branch_and_bound.py:238-389, extended to every row).  Window w
// owns output rows win_rows[16w .. 16w+15] (-1 = empty slot).  Its dense
// blocks are re-cut into chunks of 32 columns ("k-slots"):
//   chunk c in tc_ptr[w] .. tc_ptr[w+1]
//   tc_cols [c][g][e]       column of A (row of B) for k-slot k = 4e + g
//   tc_vals [c][lane][e]    A[row lane%16][k-slot 4e + lane/16]
// so that lane l of a wave fetches its MFMA A operand with one 16-byte
// (bf16) or 32-byte (fp32) load and its 8 B-row ids with two int4 loads.
// The FP32 part (ELL/1D tiles) is a CSR over the window slots:
//   res_ptr[16w + r] .. res_ptr[16w + r + 1] -> (res_cols, res_vals).
struct SpmmPlanDev {
  int32_t M, K, W;
  int32_t W_tc;              // windows [0, W_tc) hold dense chunks; the rest are sparse-only
  int32_t W_big;             // windows [0, W_big) hold more than kBigChunks chunks
  int32_t W_huge;            // windows [0, W_huge) hold more than kHugeChunks chunks
  int32_t n_parts;           // parts of the huge windows
  const int32_t* part_win;   // [n_parts] window of each part
  const int32_t* part_c0;    // [n_parts] chunk range of each part
  const int32_t* part_c1;
  int32_t nchunks;
  int32_t blk_dtype;
  int32_t npieces;           // split pieces of long sparse rows (in slot order)
  int32_t npieces_dense;     // pieces [0, npieces_dense) belong to slots of dense windows
  const int32_t* piece_row;  // [npieces] output row
  const int32_t* piece_beg;  // [npieces] range in res_cols/res_vals
  const int32_t* piece_end;
  // execution order of the pieces for the fix-up kernel: within [0,
  // npieces_dense) and [npieces_dense, npieces), pieces by their first
  // column, so the pieces running together (of different long rows) gather
  // the same B rows ([npieces])
  const int32_t* piece_exec;
  // split rows (deterministic mode): row r's pieces are [srow_p0[r], srow_p1[r])
  int32_t nsrows;
  int32_t nsrows_dense;      // split rows [0, nsrows_dense) are slots of dense windows
  const int32_t* srow_row;
  const int32_t* srow_p0;
  const int32_t* srow_p1;
  // parts of huge window w (deterministic mode): [hw_p0[w], hw_p0[w + 1])
  const int32_t* hw_p0;      // [W_huge + 1]
  // Sparse parts of the dense windows in rounds, for the B-stationary
  // kernel: window w in [0, W_tc) holds (sp_off[w + 1] - sp_off[w]) / 16
  // rounds; entry sp_ent[sp_off[w] + 16 r + s] = (column, value bits) of the
  // r-th sparse nonzero of window slot s (rows in column order, as in
  // res_cols), or (the matrix's first column, 0) past the slot's end.
  // nullptr when the plan does not qualify for that kernel.
  const int32_t* sp_off;     // [W_tc + 1]
  const int2* sp_ent;
  // B-stationary plans: bit c % 32 of tc_kind[c / 32] set = chunk c is a
  // scatter chunk (k-slot k holds one entry: the 32-bit word (BF16 value) |
  // (row >= 8) << 16 at byte 4k of the chunk's A block, the row being
  // k % 8 or k % 8 + 8); nullptr when the plan has none.  A window's scatter
  // chunks follow its dense chunks.
  const uint32_t* tc_kind;
  // B-stationary plans: first row of window w when its 16 rows are
  // consecutive, else -1 ([W])
  const int32_t* win_r0;
  // launch options (stile_plan_set_option): kernel 0 = B-stationary when the
  // plan allows it, 1 = streaming; unit list -1 = the plan's pick, else that
  // list; XCD ranges 1 = on, 0 = off
  int32_t opt_kernel, opt_lds_list, opt_lds_xcd;
  const int32_t* win_rows;   // [W*16]
  const int32_t* tc_ptr;     // [W+1]
  const int32_t* tc_cols;    // [nchunks*32]
  const void* tc_vals;       // [nchunks*64*8] bf16/f16/f32
  const float* tc_scale;     // [nchunks] (f16 only, else nullptr)
  const int32_t* res_ptr;    // [W*16+1]
  const int32_t* res_cols;
  const float* res_vals;
  int32_t nnz_sparse_only;   // sparse nonzeros in windows [W_tc, W)
  int32_t n_mix;             // slots of dense windows with a sparse part
  const int32_t* mix_slot;   // [n_mix]
  const uint32_t* tc_mask;   // [nchunks][16] entry bits (attention plans), else nullptr
  // B-stationary work units (BF16 plans with every matrix's K <= kLdsMaxK):
  // unit = (first window, end window, first B row, B rows) of one matrix,
  // windows in [0, W_tc).  List t (about kLdsTargets[t] units, chunk-balanced,
  // a matrix's units adjacent) starts at
  // lds_units[lds_off[t]] and holds lds_cnt[t] units; lds_units is nullptr
  // when the plan does not qualify.
  const int4* lds_units;
  int32_t lds_off[kLdsLists], lds_cnt[kLdsLists];
  int32_t lds_max_k;         // largest B-row count of a unit
  int8_t lds_pick[kLdsPickMax + 1];  // list to run for nC column chunks
  // Host memory (read by the launcher only; owned by the plan):
  // [kLdsLists][kLdsXcdMaxNC + 1][kXcds + 1] item ranges per XCD for list t
  // and nC column chunks, or nullptr.
  const int32_t* lds_xcd_host;
};

// Device view of a hybrid SDDMM plan.
//   1D part: n1d entries (row, col, out) computed as FP32 dot products.
//   Block part: nblk 16x16 output blocks: rows blk_rows[b][16],
//   cols blk_cols[b][16]; sampled entries s in blk_ptr[b]..blk_ptr[b+1]
//   write out[blk_out[s]] = block[blk_pos[s] >> 4][blk_pos[s] & 15].
struct SddmmPlanDev {
  int32_t M, N, nnz;
  int32_t n1d, nblk;
  int32_t blk_dtype;
  const int32_t* e_row;
  const int32_t* e_col;
  const int32_t* e_out;
  const int32_t* blk_rows;
  const int32_t* blk_cols;
  const int32_t* blk_ptr;
  const int32_t* blk_pos;
  const int32_t* blk_out;
};

// Kernel launchers (stile_kernels.hip).  All return hipError_t of the launch.
// Compute units of the current device (queried once per device).
int device_cus();

hipError_t launch_csr_spmm_f32(int M, int N, const int32_t* indptr,
                               const int32_t* indices, const float* vals,
                               const float* B, int64_t ldb, float* C,
                               int64_t ldc, hipStream_t s);
hipError_t launch_csr_sddmm_f32(int M, int nnz, int Kf, const int32_t* indptr,
                                const int32_t* indices, const float* X,
                                int64_t ldx, const float* Y, int64_t ldy,
                                float* out, hipStream_t s);
// One-time kernel attributes (dynamic LDS above 64 KiB) for a plan.
void prepare_spmm_kernels(const SpmmPlanDev& p);
// b_dtype: STILE_BLK_F32 / BF16 / F16 (the type of B's elements).
// deterministic: split rows and huge windows sum their pieces in a fixed
// order through ws (workspace_floats(p, N) floats) instead of atomically.
hipError_t launch_hyb_spmm(const SpmmPlanDev& p, int b_dtype, const void* B, int64_t ldb,
                           int N, float* C, int64_t ldc, bool deterministic, float* ws, hipStream_t s);
// Floats of workspace the deterministic mode needs at width N.
int64_t spmm_workspace_floats(const SpmmPlanDev& p, int N);
hipError_t launch_hyb_sddmm(const SddmmPlanDev& p, const float* X,
                            int64_t ldx, const float* Y, int64_t ldy, int Kf,
                            float* out, hipStream_t s);

// Refresh a plan's packed values from A's value array: tc slot i takes
// vals[tc_ids[i]] (0 when -1), converted to the plan's block type; sparse
// entry j takes vals[res_ids[j]].
hipError_t launch_set_values(const SpmmPlanDev& p, const float* vals, const int32_t* tc_ids, int64_t n_tc,
                             const int32_t* res_ids, int64_t n_res, const int32_t* sp_ids, int64_t n_sp,
                             hipStream_t s);

hipError_t launch_attention(const SpmmPlanDev& p, const void* Q, const void* K, const void* V, int64_t ld,
                            float scale, float* out, int64_t ldo, hipStream_t s);
hipError_t launch_ell_bucket_spmm(int rows, int width, const int32_t* orow, const int32_t* cols, const float* vals,
                                  const float* B, int64_t ldb, int N, float* C, int64_t ldc, hipStream_t s);
hipError_t launch_zero_window_rows(const SpmmPlanDev& p, int nwin, float* C, int64_t ldc, int N, hipStream_t s);
hipError_t launch_row_softmax(int M, const int32_t* indptr, const float* in, float* out, float scale,
                              hipStream_t s);

// Zero-fill kernels (used instead of hipMemsetAsync so that every *_run call
// records only kernel nodes under HIP graph capture).
hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t s);
hipError_t launch_fill_zero_2d(float* p, int rows, int cols, int64_t ld, hipStream_t s);

}  // namespace stile
