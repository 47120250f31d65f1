// Host side of libstile_hip.so: the C ABI of include/stile.h, validation of
// the packed tile groups, and their translation into the window-major
// execution layout the hybrid kernels consume (stile_internal.h).
//
// The packed groups follow the reference's argument lists
// (MY_sparse/my_fuse_formats.py:3321-3537 for SpMM, :3544-3678 for SDDMM).
// The reference hands them to one horizontally fused kernel whose ELL and
// TC blocks race on shared output rows and resolve it with atomicAdd
// (nvcc.py:192-441, C zeroed first, my_fuse_formats.py:4175).  Here every
// product that lands on an output row is gathered, at plan time, into the
// 16-row window that owns the row, so the kernel writes each row once.
#include "stile_internal.h"
#include "../../include/stile.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

struct StilePlan {
  int32_t kind = 0;  // 1 SpMM, 2 SDDMM
  int32_t blk_dtype = 0;
  void* dev = nullptr;  // one device allocation holding every plan array
  size_t dev_bytes = 0;
  int device = 0;
  stile::SpmmPlanDev spmm{};
  stile::SddmmPlanDev sddmm{};
  StilePlanInfo info{};
  // value refresh (stile_plan_set_values)
  bool has_ids = false;
  int64_t max_id = -1;
  int64_t n_tc_slots = 0, n_res = 0;
  const int32_t* tc_ids = nullptr;
  const int32_t* res_ids = nullptr;
  int64_t n_sp = 0;
  const int32_t* sp_ids = nullptr;
  // per-XCD item ranges of the B-stationary kernel (spmm.lds_xcd_host)
  std::vector<int32_t> lds_xcd;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return STILE_E_HIP;
}

// Bump allocator over one device buffer; every array starts 256-byte aligned.
struct Layout {
  size_t bytes = 0;
  size_t add(size_t n) {
    size_t off = (bytes + 255) & ~size_t(255);
    bytes = off + n;
    return off;
  }
};

struct HostBlob {
  std::vector<unsigned char> data;
  void put(size_t off, const void* src, size_t n) {
    if (n) std::memcpy(data.data() + off, src, n);
  }
};

int upload(StilePlan* plan, const Layout& lay, const HostBlob& blob) {
  hipError_t e = hipGetDevice(&plan->device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  plan->dev_bytes = std::max<size_t>(lay.bytes, 256);
  e = hipMalloc(&plan->dev, plan->dev_bytes);
  if (e != hipSuccess) {
    plan->dev = nullptr;
    return fail(STILE_E_NOMEM, std::string("hipMalloc of plan: ") + hipGetErrorString(e));
  }
  if (lay.bytes) {
    e = hipMemcpy(plan->dev, blob.data.data(), lay.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy of plan");
  }
  return STILE_OK;
}

template <class T>
const T* dptr(const StilePlan* plan, size_t off) {
  return reinterpret_cast<const T*>(static_cast<const unsigned char*>(plan->dev) + off);
}

uint16_t to_bf16_bits(float f) {
  __hip_bfloat16 h = __float2bfloat16(f);  // round to nearest even
  uint16_t bits;
  std::memcpy(&bits, &h, 2);
  return bits;
}

uint16_t to_f16_bits(float f) {
  _Float16 h = (_Float16)f;
  uint16_t bits;
  std::memcpy(&bits, &h, 2);
  return bits;
}

// ---------------------------------------------------------------------------
// SpMM plan construction.
// ---------------------------------------------------------------------------

struct DenseCol {
  int32_t col;
  float v[16];    // values for window slots 0..15
  int32_t id[16]; // value ids (index into A's value array), -1 = none
};

// One nonzero of a TC column that goes to a scatter chunk (B-stationary
// plans): its column, window slot (row 0..15), value and value id.
struct ScatEnt {
  int32_t col;
  int32_t row;
  float v;
  int32_t id;
};

struct SpmmBuild {
  int32_t M = 0, K = 0;
  // windows
  std::vector<int32_t> win_rows;               // [W*16]
  std::vector<std::vector<DenseCol>> win_cols;  // dense columns per window
  std::vector<std::vector<ScatEnt>> win_scat;   // scatter-chunk entries per window
  std::vector<int32_t> row_win, row_slot;      // per output row
  // sparse part, as (row, col, val) triples
  std::vector<int32_t> s_row, s_col;
  std::vector<float> s_val;
  std::vector<int32_t> s_id;
  int64_t block_nnz = 0;
  bool has_ids = true;     // every group carried a_ids
  bool want_mask = false;  // also build tc_mask (fused attention)
  int64_t max_id = -1;     // largest value id (values array must be longer)
  // block-diagonal batches: first row / column of each matrix, then M / K
  std::vector<int64_t> inst_rows, inst_cols;
};

// Validate one pack and append its windows / products to the build with row
// offset roff and column offset coff (block-diagonal batching).
int add_spmm_pack(SpmmBuild& bd, const StileSpmmPack* pk, int32_t roff, int32_t coff, int64_t ioff) {
  if (!pk) return fail(STILE_E_INVALID, "null pack");
  if (pk->M < 0 || pk->K < 0) return fail(STILE_E_INVALID, "negative M or K");
  if ((pk->num_ell > 0 && !pk->ell) || (pk->num_tc > 0 && !pk->tc) || pk->num_ell < 0 || pk->num_tc < 0)
    return fail(STILE_E_INVALID, "group array missing");
  const int32_t M = pk->M, K = pk->K;
  const size_t base_w = bd.win_rows.size() / 16;
  for (int gi = 0; gi < pk->num_tc; ++gi)
    if (!pk->tc[gi].a_ids) bd.has_ids = false;
  for (int gi = 0; gi < pk->num_ell; ++gi)
    if (!pk->ell[gi].a_ids) bd.has_ids = false;

  // 1. windows from the TC groups, keyed by their first real row.
  std::unordered_map<int32_t, int32_t> key2win;  // first row -> window index
  std::vector<int32_t> local_row_win(M, -1), local_row_slot(M, -1);
  std::vector<std::vector<int32_t>> group_slot_win(pk->num_tc);
  for (int gi = 0; gi < pk->num_tc; ++gi) {
    const StileTcGroup& g = pk->tc[gi];
    if (g.group_size <= 0 || g.group_size % 16 || g.num_windows < 0 || g.nnzb < 0)
      return fail(STILE_E_INVALID, "TC group: bad group_size/num_windows/nnzb");
    if (!g.indptr || !g.indices || !g.i_indices || (g.nnzb > 0 && !g.a))
      return fail(STILE_E_INVALID, "TC group: null buffer");
    if (g.k_notsorted && g.group_size != 16)
      return fail(STILE_E_INVALID, "TC group: k_notsorted requires group_size 16");
    if (g.indptr[0] != 0 || g.indptr[g.num_windows] != g.nnzb)
      return fail(STILE_E_INVALID, "TC group: indptr does not span [0, nnzb]");
    group_slot_win[gi].assign(g.num_windows, -1);
    for (int s = 0; s < g.num_windows; ++s) {
      if (g.indptr[s + 1] < g.indptr[s]) return fail(STILE_E_INVALID, "TC group: indptr not monotone");
      int32_t rows[16];
      if (g.rows_2d) {
        int32_t last = -1;
        for (int r = 0; r < 16; ++r) {
          int32_t v = g.i_indices[(int64_t)s * 16 + r];
          if (v < 0 || v > M) return fail(STILE_E_INVALID, "TC group: i_indices out of range");
          // Padding repeats the window's last row; M is the sentinel row.
          rows[r] = (r > 0 && v == last) || v >= M ? -1 : v;
          if (rows[r] >= 0) last = v;
        }
      } else {
        int32_t w = g.i_indices[s];
        if (w < 0 || (int64_t)w * 16 >= std::max<int64_t>(M, 1))
          return fail(STILE_E_INVALID, "TC group: window id out of range");
        for (int r = 0; r < 16; ++r) rows[r] = (int64_t)w * 16 + r < M ? w * 16 + r : -1;
      }
      int32_t key = -1;
      for (int r = 0; r < 16; ++r)
        if (rows[r] >= 0) { key = rows[r]; break; }
      if (key < 0) continue;  // a window with no real rows carries nothing
      auto it = key2win.find(key);
      int32_t win;
      if (it == key2win.end()) {
        win = (int32_t)(bd.win_rows.size() / 16);
        key2win.emplace(key, win);
        for (int r = 0; r < 16; ++r) {
          if (rows[r] >= 0) {
            if (local_row_win[rows[r]] >= 0)
              return fail(STILE_E_INVALID, "TC groups: a row belongs to two different windows");
            local_row_win[rows[r]] = win;
            local_row_slot[rows[r]] = r;
          }
          bd.win_rows.push_back(rows[r] >= 0 ? rows[r] + roff : -1);
        }
        bd.win_cols.emplace_back();
      } else {
        win = it->second;
        for (int r = 0; r < 16; ++r)
          if (rows[r] >= 0 && (local_row_win[rows[r]] != win || local_row_slot[rows[r]] != r))
            return fail(STILE_E_INVALID, "TC groups disagree on a window's row order");
      }
      group_slot_win[gi][s] = win;
    }
  }
  // 2. rows no TC window owns: windows of 16 consecutive leftover rows.
  {
    std::vector<int32_t> left;
    for (int32_t r = 0; r < M; ++r)
      if (local_row_win[r] < 0) left.push_back(r);
    for (size_t i = 0; i < left.size(); i += 16) {
      int32_t win = (int32_t)(bd.win_rows.size() / 16);
      for (size_t r = 0; r < 16; ++r) {
        if (i + r < left.size()) {
          local_row_win[left[i + r]] = win;
          local_row_slot[left[i + r]] = (int32_t)r;
          bd.win_rows.push_back(left[i + r] + roff);
        } else {
          bd.win_rows.push_back(-1);
        }
      }
      bd.win_cols.emplace_back();
    }
  }
  (void)base_w;
  for (int32_t r = 0; r < M; ++r) {
    bd.row_win.push_back(local_row_win[r]);
    bd.row_slot.push_back(local_row_slot[r]);
  }

  // 3. dense columns per window (merged across tiles and groups).
  for (int gi = 0; gi < pk->num_tc; ++gi) {
    const StileTcGroup& g = pk->tc[gi];
    const int Kg = g.group_size;
    for (int s = 0; s < g.num_windows; ++s) {
      const int32_t win = group_slot_win[gi][s];
      for (int32_t t = g.indptr[s]; t < g.indptr[s + 1]; ++t) {
        for (int kk = 0; kk < Kg; ++kk) {
          int32_t col = g.k_notsorted ? g.indices[t] * 16 + kk : g.indices[(int64_t)t * Kg + kk];
          DenseCol dc;
          dc.col = col + coff;
          bool any = false;
          for (int r = 0; r < 16; ++r) {
            const int64_t at = ((int64_t)t * 16 + r) * Kg + kk;
            const float v = g.a[at];
            const int32_t id = g.a_ids ? g.a_ids[at] : -1;
            // a slot holds a nonzero when it has an id (id-carrying packs) or a
            // nonzero value (value-only packs)
            const bool present = g.a_ids ? id >= 0 : v != 0.f;
            if (g.a_ids && id < -1) return fail(STILE_E_INVALID, "TC group: negative value id");
            if (g.a_ids && !present && v != 0.f) return fail(STILE_E_INVALID, "TC group: value without an id");
            dc.v[r] = present ? v : 0.f;
            dc.id[r] = present && g.a_ids ? id + (int32_t)ioff : -1;
            if (present) {
              if (win < 0) return fail(STILE_E_INVALID, "TC group: nonzero value in a window without rows");
              if (bd.win_rows[(size_t)win * 16 + r] < 0)
                return fail(STILE_E_INVALID, "TC group: nonzero value in a padding row");
              any = true;
              bd.block_nnz++;
              bd.max_id = std::max<int64_t>(bd.max_id, dc.id[r]);
            }
          }
          if (!any) continue;
          if (col < 0 || col >= K) return fail(STILE_E_INVALID, "TC group: column index out of range");
          bd.win_cols[win].push_back(dc);
        }
      }
    }
  }

  // 4. sparse (ELL) products.
  for (int gi = 0; gi < pk->num_ell; ++gi) {
    const StileEllGroup& g = pk->ell[gi];
    if (g.m_blk <= 0 || g.nnz_cols <= 0 || g.num_tiles < 0)
      return fail(STILE_E_INVALID, "ELL group: bad m_blk/nnz_cols/num_tiles");
    if (g.num_tiles > 0 && (!g.a || !g.indices_k || !g.indices_reordered_i))
      return fail(STILE_E_INVALID, "ELL group: null buffer");
    const int I = g.m_blk, Kc = g.nnz_cols;
    for (int64_t t = 0; t < g.num_tiles; ++t) {
      for (int i = 0; i < I; ++i) {
        const int64_t tr = t * I + i;
        const int32_t row = g.indices_reordered_i[tr];
        if (row < 0 || row > M) return fail(STILE_E_INVALID, "ELL group: row index out of range");
        for (int kk = 0; kk < Kc; ++kk) {
          const int64_t at = t * (int64_t)I * Kc + (int64_t)i * Kc + kk;
          const float v = g.a[at];
          const int32_t id = g.a_ids ? g.a_ids[at] : -1;
          if (g.a_ids && id < -1) return fail(STILE_E_INVALID, "ELL group: negative value id");
          const bool present = g.a_ids ? id >= 0 : v != 0.f;
          if (!present) continue;   // padding slots (and explicit zeros of value-only packs)
          if (row == M) continue;   // sentinel row: products are discarded
          const int32_t col = g.indices_k[tr * Kc + kk];
          if (col < 0 || col >= K) return fail(STILE_E_INVALID, "ELL group: column index out of range");
          bd.s_row.push_back(row + roff);
          bd.s_col.push_back(col + coff);
          bd.s_val.push_back(v);
          bd.s_id.push_back(g.a_ids ? id + (int32_t)ioff : -1);
          if (g.a_ids) bd.max_id = std::max<int64_t>(bd.max_id, id + ioff);
        }
      }
    }
  }
  (void)base_w;
  return STILE_OK;
}

// Matrix (batch index) of window w: the one holding its first row.
int32_t window_matrix(const SpmmBuild& bd, int32_t w) {
  if (bd.inst_rows.size() <= 2) return 0;
  for (int r = 0; r < 16; ++r) {
    const int32_t row = bd.win_rows[(size_t)w * 16 + r];
    if (row >= 0)
      return (int32_t)(std::upper_bound(bd.inst_rows.begin(), bd.inst_rows.end(), (int64_t)row) -
                       bd.inst_rows.begin()) - 1;
  }
  return 0;
}

// Estimated time of one workgroup item of the B-stationary kernel in units
// of one 16x32 chunk times 64 columns: the chunks plus staging ksz B rows,
// which costs about one chunk per kLdsStageRows rows.
constexpr int64_t kLdsStageRows = 8;

// Longest workgroup of a launch of `items` items (unit i / nC) over
// min(items, cus) persistent workgroups that take items round-robin.
int64_t lds_makespan(const std::vector<int64_t>& unit_cost, int nC, int cus) {
  const int64_t items = (int64_t)unit_cost.size() * nC;
  const int64_t G = std::min<int64_t>(items, cus);
  int64_t worst = 0;
  for (int64_t b = 0; b < G; ++b) {
    int64_t sum = 0;
    for (int64_t i = b; i < items; i += G) sum += unit_cost[i / nC];
    worst = std::max(worst, sum);
  }
  return worst;
}

// Contiguous item ranges per XCD (items unit-major: item i is unit i / nC),
// cut where the running cost is nearest to x / kXcds of the total.
void lds_xcd_split(const std::vector<int64_t>& unit_cost, int nC, int32_t (&start)[stile::kXcds + 1]) {
  const int64_t items = (int64_t)unit_cost.size() * nC;
  int64_t total = 0;
  for (int64_t c : unit_cost) total += c * nC;
  start[0] = 0;
  int64_t i = 0, run = 0;  // run = cost of items [0, i)
  for (int x = 1; x < stile::kXcds; ++x) {
    const int64_t goal = total * x / stile::kXcds;
    while (i < items && run + unit_cost[i / nC] <= goal) run += unit_cost[i++ / nC];
    // item i straddles the goal: take it when that lands nearer
    if (i < items && run + unit_cost[i / nC] - goal < goal - run) run += unit_cost[i++ / nC];
    start[x] = (int32_t)i;
  }
  start[stile::kXcds] = (int32_t)items;
}

// Longest workgroup under the XCD grouping: XCD x's cus / kXcds workgroups
// take its range round-robin.
int64_t lds_xcd_makespan(const std::vector<int64_t>& unit_cost, int nC, int cus,
                         const int32_t (&start)[stile::kXcds + 1]) {
  const int per = std::max(1, cus / stile::kXcds);
  int64_t worst = 0;
  for (int x = 0; x < stile::kXcds; ++x)
    for (int b = 0; b < per; ++b) {
      int64_t sum = 0;
      for (int64_t i = start[x] + b; i < start[x + 1]; i += per) sum += unit_cost[i / nC];
      worst = std::max(worst, sum);
    }
  return worst;
}

// Work-unit lists of the B-stationary kernel.  The dense windows [0, W_tc)
// are grouped by matrix (contiguous); list t splits each matrix's windows
// into about kLdsTargets[t] * (its share of the chunks) contiguous,
// chunk-balanced units (at least one per matrix).  A matrix's units stay
// adjacent; matrices are ordered by decreasing largest unit (close to
// longest-processing-time-first for workgroups that take units in turn).
// Appends the lists to `units`.  pick[nC] is the list with the smallest
// estimated makespan on `cus` compute units for nC column chunks, under the
// mapping the launcher uses (XCD-grouped for persistent launches with nC <=
// kLdsXcdMaxNC); xcd receives those launches' ranges per list and nC.
void build_lds_units(const SpmmBuild& bd, const std::vector<int32_t>& tc_ptr, int32_t W_tc, int cus,
                     std::vector<int32_t>& units, int32_t (&off)[stile::kLdsLists],
                     int32_t (&cnt)[stile::kLdsLists], int32_t& max_k, int8_t (&pick)[stile::kLdsPickMax + 1],
                     std::vector<int32_t>& xcd) {
  // runs of windows of one matrix
  std::vector<int32_t> run_w0, run_m;
  for (int32_t w = 0; w < W_tc; ++w) {
    const int32_t m = window_matrix(bd, w);
    if (run_m.empty() || run_m.back() != m) {
      run_w0.push_back(w);
      run_m.push_back(m);
    }
  }
  run_w0.push_back(W_tc);
  const int64_t total = tc_ptr[W_tc];
  max_k = 0;
  for (int t = 0; t < stile::kLdsLists; ++t) {
    off[t] = (int32_t)(units.size() / 4);
    const int64_t target = stile::kLdsTargets[t];
    std::vector<std::vector<std::array<int32_t, 4>>> per_run(run_m.size());
    std::vector<int64_t> run_key(run_m.size(), 0);  // largest unit of the run (chunks)
    for (size_t r = 0; r + 1 < run_w0.size(); ++r) {
      const int32_t w0 = run_w0[r], w1 = run_w0[r + 1], m = run_m[r];
      const int64_t c0 = tc_ptr[w0], chunks = tc_ptr[w1] - c0;
      const int64_t n = std::max<int64_t>(1, std::min<int64_t>(w1 - w0, (target * chunks + total - 1) / std::max<int64_t>(total, 1)));
      const int32_t koff = (int32_t)bd.inst_cols[m], ksz = (int32_t)(bd.inst_cols[m + 1] - bd.inst_cols[m]);
      max_k = std::max(max_k, ksz);
      int32_t a = w0;
      for (int64_t u = 1; u <= n && a < w1; ++u) {
        // end: first window whose chunk start reaches u / n of the run
        const int64_t goal = c0 + (chunks * u + n - 1) / n;
        int32_t b = u == n ? w1 : a + 1;
        while (b < w1 && tc_ptr[b] < goal) ++b;
        per_run[r].push_back({a, b, koff, ksz});
        run_key[r] = std::max<int64_t>(run_key[r], tc_ptr[b] - tc_ptr[a]);
        a = b;
      }
    }
    std::vector<size_t> order(run_m.size());
    for (size_t r = 0; r < order.size(); ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return run_key[x] > run_key[y]; });
    for (size_t r : order)
      for (const auto& u : per_run[r]) units.insert(units.end(), u.begin(), u.end());
    cnt[t] = (int32_t)(units.size() / 4) - off[t];
  }
  // list choice per column-chunk count; lists with more than 4 items per CU
  // are not considered once a smaller one fills the GPU (more restaging, and
  // the estimate would cost more than it saves)
  const int S = stile::kXcds + 1;
  xcd.assign((size_t)stile::kLdsLists * (stile::kLdsXcdMaxNC + 1) * S, 0);
  pick[0] = 0;
  for (int nC = 1; nC <= stile::kLdsPickMax; ++nC) {
    int best = 0;
    int64_t best_cost = INT64_MAX;
    for (int t = 0; t < stile::kLdsLists; ++t) {
      const int64_t items = (int64_t)cnt[t] * nC;
      if (t > 0 && items > 4 * (int64_t)cus && (int64_t)cnt[t - 1] * nC >= cus) break;
      std::vector<int64_t> cost(cnt[t]);
      for (int32_t i = 0; i < cnt[t]; ++i) {
        const int32_t* u = &units[(size_t)(off[t] + i) * 4];
        cost[i] = (int64_t)(tc_ptr[u[1]] - tc_ptr[u[0]]) + (u[3] + kLdsStageRows - 1) / kLdsStageRows;
      }
      int64_t c;
      if (nC <= stile::kLdsXcdMaxNC && items <= INT32_MAX) {
        int32_t start[stile::kXcds + 1];
        lds_xcd_split(cost, nC, start);
        std::copy(start, start + S, xcd.begin() + ((size_t)t * (stile::kLdsXcdMaxNC + 1) + nC) * S);
        c = items > cus ? lds_xcd_makespan(cost, nC, cus, start) : lds_makespan(cost, nC, cus);
      } else {
        c = lds_makespan(cost, nC, cus);
      }
      if (c < best_cost) {
        best_cost = c;
        best = t;
      }
    }
    pick[nC] = (int8_t)best;
  }
  // ranges of the lists the pick loop skipped (forced lists)
  for (int t = 0; t < stile::kLdsLists; ++t)
    for (int nC = 1; nC <= stile::kLdsXcdMaxNC; ++nC) {
      int32_t* st = &xcd[((size_t)t * (stile::kLdsXcdMaxNC + 1) + nC) * S];
      if (st[S - 1] != 0 || cnt[t] == 0) continue;
      std::vector<int64_t> cost(cnt[t]);
      for (int32_t i = 0; i < cnt[t]; ++i) {
        const int32_t* u = &units[(size_t)(off[t] + i) * 4];
        cost[i] = (int64_t)(tc_ptr[u[1]] - tc_ptr[u[0]]) + (u[3] + kLdsStageRows - 1) / kLdsStageRows;
      }
      int32_t start[stile::kXcds + 1];
      lds_xcd_split(cost, nC, start);
      std::copy(start, start + S, st);
    }
}

// k-slot order within each chunk of 32 dense columns of a window, for the
// B-stationary kernel: k-slot ks = 4e + g is read by lane group g, and one
// transposed LDS read of a half-wave gathers the B rows of the 8 k-slots
// {16h + 4e' + g : e' = 0..3, g in {2P, 2P + 1}} (h, P = 0, 1), each lane
// group's 4 rows in one 16-lane access.  Rows of kLdsRowBytes bytes fall on
// distinct LDS banks when a lane group's 4 rows differ mod 4 and the
// half-wave's 8 rows differ mod 8.  The columns are dealt to the four k-slot
// sets by residue class (col - koff) mod 8, so every set holds distinct
// residues whenever no class has more than 4 of the chunk's columns (two
// 16-column dense blocks: 2 each); within a set the lower residues go to
// lane group 2P and the higher to 2P + 1 ({0..3} and {4..7} for a full set).
// Only the order changes; the last chunk keeps its real columns in its
// first n k-slots.  (A transposed read whose 8 half-wave rows share a
// residue costs 7 extra LDS cycles; with distinct residues none.)
void order_chunk_slots(std::vector<DenseCol>& cols, int64_t koff) {
  // Which columns share a chunk is free (a window's chunks together cover
  // its columns): deal the columns to the chunks round-robin over the
  // residue classes, so each chunk of 32 holds at most 4 of a class while
  // every class still has columns left.
  {
    std::vector<DenseCol> by[8];
    for (const DenseCol& dc : cols) by[(dc.col - koff) & 7].push_back(dc);
    size_t at[8] = {0, 0, 0, 0, 0, 0, 0, 0}, o = 0;
    while (o < cols.size())
      for (int r = 0; r < 8; ++r)
        if (at[r] < by[r].size()) cols[o++] = by[r][at[r]++];
  }
  for (size_t base = 0; base < cols.size(); base += 32) {
    const int n = (int)std::min<size_t>(32, cols.size() - base);
    // the real k-slots of each set, lane group first (g, then e')
    std::vector<int> slots[4];
    for (int ks = 0; ks < n; ++ks) slots[(ks >> 4) * 2 + ((ks & 3) >> 1)].push_back(ks);
    for (auto& sl : slots)
      std::stable_sort(sl.begin(), sl.end(), [](int a, int b) { return (a & 3) < (b & 3); });
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto res = [&](int i) { return (int)((cols[base + i].col - koff) & 7); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return res(a) < res(b); });
    unsigned used[4] = {0, 0, 0, 0};  // residues present per set
    std::vector<int> members[4];      // chunk column indices dealt to each set
    for (int i = 0; i < n; ++i) {
      const int r = res(order[i]);
      int pick = -1;
      for (int d = 0; d < 4 && pick < 0; ++d) {  // preferred set first, then the others
        const int set = (i + d) & 3;
        if (members[set].size() < slots[set].size() && !(used[set] >> r & 1u)) pick = set;
      }
      for (int set = 0; set < 4 && pick < 0; ++set)
        if (members[set].size() < slots[set].size()) pick = set;
      used[pick] |= 1u << r;
      members[pick].push_back(order[i]);
    }
    std::vector<DenseCol> out(n);
    for (int set = 0; set < 4; ++set) {
      // ascending residue: the lower half to the set's first lane group
      std::stable_sort(members[set].begin(), members[set].end(), [&](int a, int b) { return res(a) < res(b); });
      for (size_t m = 0; m < members[set].size(); ++m) out[slots[set][m]] = cols[base + members[set][m]];
    }
    std::copy(out.begin(), out.end(), cols.begin() + base);
  }
}

// Scatter chunks (B-stationary plans): a chunk whose 32 k-slots hold one
// nonzero each.  The MFMA A operand gives lane (G, r) the elements A[r][8G +
// e], e = 0..7; the entry of k-slot 8G + e is placed so that its row is e or
// e + 8, so lane (G, r) has its one possible value at element r % 8 and the
// chunk stores per k-slot only that value and whether the row is >= 8 (4
// bytes, against 32 for a column of a dense chunk).  Returns the number of
// scatter chunks the entries need.
int32_t scatter_chunks(const std::vector<ScatEnt>& s) {
  int32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (const ScatEnt& e : s) cnt[e.row & 7]++;
  int32_t m = 0;
  for (int r = 0; r < 8; ++r) m = std::max(m, (cnt[r] + 3) / 4);
  return m;
}

// k-slot contents of a window's scatter chunks: per chunk 32 indices into s
// (-1: padding).  Chunk j takes entries 4j .. 4j + 3 of each row-residue
// class e; within a chunk the entries of class e are placed over the four
// lane groups G so that each half-wave's transposed B read (k-slots
// {8G + e: G in a pair, e in a half}) meets distinct B-row residues mod 8
// where it can (the LDS bank groups, as for dense chunks).
std::vector<std::array<int32_t, 32>> place_scatter(const std::vector<ScatEnt>& s, int64_t koff) {
  std::vector<int32_t> cls[8];
  for (size_t i = 0; i < s.size(); ++i) cls[s[i].row & 7].push_back((int32_t)i);
  const int32_t nch = scatter_chunks(s);
  std::vector<std::array<int32_t, 32>> out(nch);
  for (int32_t j = 0; j < nch; ++j) {
    auto& slot = out[j];
    slot.fill(-1);
    for (int h = 0; h < 2; ++h) {
      unsigned used[2] = {0, 0};  // B-row residues read by lane-group pair 0-1 / 2-3
      for (int e = 4 * h; e < 4 * h + 4; ++e) {
        int32_t ents[4];
        int m = 0;
        for (size_t q = (size_t)4 * j; q < cls[e].size() && m < 4; ++q) ents[m++] = cls[e][q];
        // best placement of the m entries over the lane groups 0..3
        int perm[4] = {0, 1, 2, 3}, best[4] = {0, 1, 2, 3}, best_cost = 1 << 30;
        do {
          unsigned u[2] = {used[0], used[1]};
          int cost = 0;
          for (int q = 0; q < m; ++q) {
            const int pair = perm[q] >> 1;
            const unsigned bit = 1u << ((s[ents[q]].col - koff) & 7);
            if (u[pair] & bit) ++cost;
            u[pair] |= bit;
          }
          if (cost < best_cost) {
            best_cost = cost;
            std::copy(perm, perm + 4, best);
          }
        } while (best_cost > 0 && std::next_permutation(perm, perm + 4));
        for (int q = 0; q < m; ++q) {
          slot[8 * best[q] + e] = ents[q];
          used[best[q] >> 1] |= 1u << ((s[ents[q]].col - koff) & 7);
        }
      }
    }
  }
  return out;
}

int finish_spmm(SpmmBuild& bd, int32_t blk_dtype, int32_t scatter_max, StilePlan** out) {
  const int32_t W = (int32_t)(bd.win_rows.size() / 16);
  bd.win_scat.assign(W, {});

  // Dense columns of each window: merge duplicates (a column covered by
  // several tiles of the window), sort by column.
  for (int32_t w = 0; w < W; ++w) {
    auto& cols = bd.win_cols[w];
    std::sort(cols.begin(), cols.end(), [](const DenseCol& a, const DenseCol& b) { return a.col < b.col; });
    size_t o = 0;
    for (size_t i = 0; i < cols.size(); ++i) {
      if (o > 0 && cols[o - 1].col == cols[i].col) {
        for (int r = 0; r < 16; ++r) {
          if (cols[o - 1].id[r] >= 0 && cols[i].id[r] >= 0)
            return fail(STILE_E_INVALID, "two TC tiles cover the same nonzero");
          cols[o - 1].v[r] += cols[i].v[r];
          if (cols[i].id[r] >= 0) cols[o - 1].id[r] = cols[i].id[r];
        }
      } else {
        cols[o++] = cols[i];
      }
    }
    cols.resize(o);
  }
  // BF16 plans whose matrices all have K <= kLdsMaxK run on the B-stationary
  // kernel, which stages one matrix's B rows per workgroup.
  int64_t max_k = 0;
  for (size_t i = 0; i + 1 < bd.inst_cols.size(); ++i) max_k = std::max(max_k, bd.inst_cols[i + 1] - bd.inst_cols[i]);
  const bool lds_ok = blk_dtype == STILE_BLK_BF16 && max_k <= stile::kLdsMaxK && !bd.want_mask;
  // B-stationary plans: a window's TC columns with at most scatter_max
  // nonzeros go to scatter chunks when that shrinks the window's A bytes.
  if (lds_ok && scatter_max > 0) {
    for (int32_t w = 0; w < W; ++w) {
      auto& cols = bd.win_cols[w];
      std::vector<DenseCol> keep;
      std::vector<ScatEnt> sc;
      for (const DenseCol& dc : cols) {
        int n = 0;
        for (int r = 0; r < 16; ++r) n += (dc.id[r] >= 0 || dc.v[r] != 0.f) ? 1 : 0;
        if (n > scatter_max) {
          keep.push_back(dc);
        } else {
          for (int r = 0; r < 16; ++r)
            if (dc.id[r] >= 0 || dc.v[r] != 0.f) sc.push_back(ScatEnt{dc.col, r, dc.v[r], dc.id[r]});
        }
      }
      const int64_t before = (int64_t)(cols.size() + 31) / 32 * 1024;
      const int64_t after = (int64_t)(keep.size() + 31) / 32 * 1024 + (int64_t)scatter_chunks(sc) * 128;
      if (!sc.empty() && after < before) {
        cols.swap(keep);
        bd.win_scat[w].swap(sc);
      }
    }
  }
  auto has_chunks = [&](int32_t w) { return !bd.win_cols[w].empty() || !bd.win_scat[w].empty(); };
  // Windows with more than stile::kBigChunks dense chunks first (most chunks first,
  // W_big of them), then the other windows with dense chunks (W_tc in all),
  // then the sparse-only windows; the kernels map the kinds onto waves
  // differently.
  std::vector<int32_t> order(W);
  for (int32_t w = 0; w < W; ++w) order[w] = w;
  auto nch = [&](int32_t w) {
    return (int64_t)(bd.win_cols[w].size() + 31) / 32 + scatter_chunks(bd.win_scat[w]);
  };
  std::stable_partition(order.begin(), order.end(), [&](int32_t w) { return has_chunks(w); });
  int32_t W_tc = 0;
  while (W_tc < W && has_chunks(order[W_tc])) ++W_tc;
  // B-stationary plans of several matrices keep their dense windows grouped
  // by matrix (and no window counts as big or huge, which only the streaming
  // kernel uses).
  const bool by_matrix = lds_ok && bd.inst_rows.size() > 2;
  int32_t W_big = 0, W_huge = 0;
  if (by_matrix) {
    std::vector<int32_t> mat(W);
    for (int32_t w = 0; w < W; ++w) mat[w] = window_matrix(bd, w);
    std::stable_sort(order.begin(), order.begin() + W_tc, [&](int32_t a, int32_t b) { return mat[a] < mat[b]; });
  } else {
    auto big_end = std::stable_partition(order.begin(), order.begin() + W_tc,
                                         [&](int32_t w) { return nch(w) > stile::kBigChunks; });
    std::stable_sort(order.begin(), big_end, [&](int32_t a, int32_t b) { return nch(a) > nch(b); });
    W_big = (int32_t)(big_end - order.begin());
    while (W_huge < W_big && nch(order[W_huge]) > stile::kHugeChunks) ++W_huge;
  }
  {
    std::vector<int32_t> new_of(W);
    for (int32_t nw = 0; nw < W; ++nw) new_of[order[nw]] = nw;
    std::vector<int32_t> rows2(bd.win_rows.size());
    std::vector<std::vector<DenseCol>> cols2(W);
    std::vector<std::vector<ScatEnt>> scat2(W);
    for (int32_t nw = 0; nw < W; ++nw) {
      for (int r = 0; r < 16; ++r) rows2[(size_t)nw * 16 + r] = bd.win_rows[(size_t)order[nw] * 16 + r];
      cols2[nw].swap(bd.win_cols[order[nw]]);
      scat2[nw].swap(bd.win_scat[order[nw]]);
    }
    bd.win_rows.swap(rows2);
    bd.win_cols.swap(cols2);
    bd.win_scat.swap(scat2);
    for (auto& rw : bd.row_win)
      if (rw >= 0) rw = new_of[rw];
  }

  // Sparse part as CSR over window slots, columns ascending within a slot.
  std::vector<int64_t> cnt((size_t)W * 16 + 1, 0);
  std::vector<int32_t> slot_of(bd.s_row.size());
  for (size_t e = 0; e < bd.s_row.size(); ++e) {
    const int32_t r = bd.s_row[e];
    const int32_t slot = bd.row_win[r] * 16 + bd.row_slot[r];
    slot_of[e] = slot;
    cnt[slot + 1]++;
  }
  for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
  const int64_t nres = cnt.back();
  if (nres > INT32_MAX) return fail(STILE_E_UNSUPPORTED, "more than 2^31 sparse nonzeros");
  std::vector<int32_t> res_ptr(cnt.begin(), cnt.end());
  std::vector<int32_t> res_cols(nres);
  std::vector<float> res_vals(nres);
  std::vector<int32_t> res_ids(bd.has_ids ? nres : 0);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    std::vector<int64_t> src(nres);
    for (size_t e = 0; e < bd.s_row.size(); ++e) src[fill[slot_of[e]]++] = (int64_t)e;
    // columns ascending within a slot (stable: ties keep pack order)
    for (size_t sl = 0; sl + 1 < cnt.size(); ++sl) {
      const int64_t a = cnt[sl], b = cnt[sl + 1];
      if (b - a >= 2)
        std::stable_sort(src.begin() + a, src.begin() + b,
                         [&](int64_t x, int64_t y) { return bd.s_col[x] < bd.s_col[y]; });
    }
    for (int64_t q = 0; q < nres; ++q) {
      res_cols[q] = bd.s_col[src[q]];
      res_vals[q] = bd.s_val[src[q]];
      if (bd.has_ids) res_ids[q] = bd.s_id[src[q]];
    }
  }
  // Slots of dense windows that also have a sparse part: the sparse kernel
  // adds those parts to the rows the dense kernels stored.
  std::vector<int32_t> mix_slot;
  for (int64_t sl = 0; sl < (int64_t)W_tc * 16; ++sl)
    if (cnt[sl + 1] > cnt[sl]) mix_slot.push_back((int32_t)sl);
  // Long rows: the window pass takes the first kSplitNnz nonzeros of a slot,
  // the rest is cut into pieces of kSplitNnz for the fix-up kernel.
  // Pieces are in slot order, so those of dense windows' slots come first
  // (npieces_dense); each split row's pieces are consecutive (srow_*).
  int32_t hub_rows = 0;
  std::vector<int32_t> piece_row, piece_beg, piece_end, srow_row, srow_p0, srow_p1;
  int32_t npieces_dense = 0, nsrows_dense = 0;
  for (size_t sl = 0; sl + 1 < cnt.size(); ++sl) {
    const int64_t a = cnt[sl], b = cnt[sl + 1];
    if (b - a > stile::kHubNnz) hub_rows++;
    if (b - a <= stile::kSplitNnz) continue;
    const int32_t row = bd.win_rows[sl];
    srow_row.push_back(row);
    srow_p0.push_back((int32_t)piece_row.size());
    for (int64_t q = a + stile::kSplitNnz; q < b; q += stile::kSplitNnz) {
      piece_row.push_back(row);
      piece_beg.push_back((int32_t)q);
      piece_end.push_back((int32_t)std::min<int64_t>(b, q + stile::kSplitNnz));
    }
    srow_p1.push_back((int32_t)piece_row.size());
    if ((int64_t)sl < (int64_t)W_tc * 16) {
      npieces_dense = (int32_t)piece_row.size();
      nsrows_dense = (int32_t)srow_row.size();
    }
  }

  if (lds_ok)
    for (int32_t w = 0; w < W_tc; ++w) order_chunk_slots(bd.win_cols[w], bd.inst_cols[window_matrix(bd, w)]);
  // Dense chunks: cut each window's columns into chunks of 32 k-slots (the
  // last one padded with value-0 slots that point at the chunk's first
  // column), then the window's scatter chunks (tc_kind bit set).
  std::vector<int32_t> tc_ptr(W + 1, 0), n_dense(W, 0);
  int64_t nchunks = 0;
  bool any_scatter = false;
  for (int32_t w = 0; w < W; ++w) {
    n_dense[w] = (int32_t)((bd.win_cols[w].size() + 31) / 32);
    const int32_t ns = scatter_chunks(bd.win_scat[w]);
    any_scatter = any_scatter || ns > 0;
    nchunks += n_dense[w] + ns;
    tc_ptr[w + 1] = (int32_t)nchunks;
  }
  if (nchunks > INT32_MAX / 64) return fail(STILE_E_UNSUPPORTED, "too many dense chunks");
  std::vector<int32_t> lds_units;
  int32_t lds_off[stile::kLdsLists] = {}, lds_cnt[stile::kLdsLists] = {}, lds_max_k = 0;
  int8_t lds_pick[stile::kLdsPickMax + 1];
  std::fill(std::begin(lds_pick), std::end(lds_pick), (int8_t)-1);
  std::vector<int32_t> lds_xcd;
  if (lds_ok && W_tc > 0)
    build_lds_units(bd, tc_ptr, W_tc, stile::device_cus(), lds_units, lds_off, lds_cnt, lds_max_k, lds_pick,
                    lds_xcd);
  // Sparse parts of the dense windows in rounds (B-stationary plans).
  std::vector<int32_t> sp_off, sp_ids;
  std::vector<int32_t> sp_ent;  // (column, value bits) pairs
  if (lds_ok && W_tc > 0) {
    sp_off.assign((size_t)W_tc + 1, 0);
    int64_t total = 0;
    for (int32_t w = 0; w < W_tc; ++w) {
      sp_off[w] = (int32_t)total;
      int64_t R = 0;
      for (int sl = 0; sl < 16; ++sl) R = std::max<int64_t>(R, cnt[(size_t)w * 16 + sl + 1] - cnt[(size_t)w * 16 + sl]);
      const int32_t pad_col = (int32_t)bd.inst_cols[window_matrix(bd, w)];
      for (int64_t r = 0; r < R; ++r)
        for (int sl = 0; sl < 16; ++sl) {
          const int64_t q = cnt[(size_t)w * 16 + sl] + r;
          const bool real = q < cnt[(size_t)w * 16 + sl + 1];
          const float v = real ? res_vals[q] : 0.f;
          int32_t vb;
          std::memcpy(&vb, &v, 4);
          sp_ent.push_back(real ? res_cols[q] : pad_col);
          sp_ent.push_back(vb);
          if (bd.has_ids) sp_ids.push_back(real ? res_ids[q] : -1);
        }
      total += R * 16;
      if (total > INT32_MAX) return fail(STILE_E_UNSUPPORTED, "too many sparse entries in dense windows");
    }
    sp_off[W_tc] = (int32_t)total;
  }
  // Parts of the huge windows: consecutive ranges of at most kHugeChunks chunks.
  std::vector<int32_t> part_win, part_c0, part_c1, hw_p0;
  for (int32_t w = 0; w < W_huge; ++w) {
    hw_p0.push_back((int32_t)part_win.size());
    for (int64_t c = tc_ptr[w]; c < tc_ptr[w + 1]; c += stile::kHugeChunks) {
      part_win.push_back(w);
      part_c0.push_back((int32_t)c);
      part_c1.push_back((int32_t)std::min<int64_t>(c + stile::kHugeChunks, tc_ptr[w + 1]));
    }
  }
  hw_p0.push_back((int32_t)part_win.size());
  const size_t val_bytes = blk_dtype == STILE_BLK_F32 ? 4 : 2;
  std::vector<int32_t> tc_cols((size_t)nchunks * 32);
  std::vector<unsigned char> tc_vals((size_t)nchunks * 64 * 8 * val_bytes);
  std::vector<float> tc_scale(blk_dtype == STILE_BLK_F16 ? nchunks : 0);
  std::vector<int32_t> tc_ids(bd.has_ids ? (size_t)nchunks * 64 * 8 : 0);
  // tc_mask[c][r] bit s: k-slot s of chunk c holds an entry of window row r
  std::vector<uint32_t> tc_mask(bd.want_mask ? (size_t)nchunks * 16 : 0);
  std::vector<uint32_t> tc_kind(any_scatter ? (size_t)(nchunks + 31) / 32 : 0);
  for (int32_t w = 0; w < W; ++w) {
    const auto& cols = bd.win_cols[w];
    for (int64_t c = tc_ptr[w]; c < tc_ptr[w] + n_dense[w]; ++c) {
      if (bd.want_mask) {
        const size_t b0 = (size_t)(c - tc_ptr[w]) * 32;
        for (int r = 0; r < 16; ++r) {
          uint32_t bits = 0;
          for (size_t ks = 0; ks < 32 && b0 + ks < cols.size(); ++ks)
            if (cols[b0 + ks].id[r] >= 0 || cols[b0 + ks].v[r] != 0.f) bits |= 1u << ks;
          tc_mask[(size_t)c * 16 + r] = bits;
        }
      }
      const size_t base = (size_t)(c - tc_ptr[w]) * 32;
      float scale = 1.f;
      if (blk_dtype == STILE_BLK_F16) {
        float amax = 0.f;
        for (size_t ks = 0; ks < 32 && base + ks < cols.size(); ++ks)
          for (int r = 0; r < 16; ++r) amax = std::max(amax, std::fabs(cols[base + ks].v[r]));
        // Scale the chunk so its largest magnitude maps to 2^14 (well inside
        // FP16's range, far above its subnormals).
        if (amax > 0.f) scale = std::ldexp(1.f, std::ilogb(amax) - 14);
        tc_scale[c] = scale;
      }
      for (int g = 0; g < 4; ++g)
        for (int e = 0; e < 8; ++e) {
          const size_t ks = 4 * e + g;
          tc_cols[(size_t)c * 32 + g * 8 + e] =
              base + ks < cols.size() ? cols[base + ks].col : cols[base].col;
        }
      for (int lane = 0; lane < 64; ++lane) {
        const int row = lane & 15, g = lane >> 4;
        for (int e = 0; e < 8; ++e) {
          const size_t ks = 4 * e + g;
          const float v = base + ks < cols.size() ? cols[base + ks].v[row] : 0.f;
          const size_t at = ((size_t)c * 64 + lane) * 8 + e;
          if (bd.has_ids) tc_ids[at] = base + ks < cols.size() ? cols[base + ks].id[row] : -1;
          if (blk_dtype == STILE_BLK_F32) {
            std::memcpy(&tc_vals[at * 4], &v, 4);
          } else if (blk_dtype == STILE_BLK_BF16) {
            uint16_t b = to_bf16_bits(v);
            std::memcpy(&tc_vals[at * 2], &b, 2);
          } else {
            uint16_t b = to_f16_bits(v / scale);
            std::memcpy(&tc_vals[at * 2], &b, 2);
          }
        }
      }
    }
    // Scatter chunks (BF16): k-slot k holds one entry, stored in the first
    // 128 bytes of the chunk's A block as the 32-bit word (BF16 value) |
    // (row >= 8) << 16 at byte 4k.  Its value id goes with the value half
    // (A element 2k); the row half keeps its bit across value refreshes.
    // Padding slots: value 0, column = a B row of the matrix whose residue
    // its half-wave read does not meet yet.
    if (!bd.win_scat[w].empty()) {
      const auto& sc = bd.win_scat[w];
      const int32_t m = window_matrix(bd, w);
      const int64_t koff = bd.inst_cols[m], ksz = bd.inst_cols[m + 1] - bd.inst_cols[m];
      const auto slots = place_scatter(sc, koff);
      for (size_t j = 0; j < slots.size(); ++j) {
        const int64_t c = tc_ptr[w] + n_dense[w] + (int64_t)j;
        tc_kind[(size_t)c >> 5] |= 1u << (c & 31);
        if (bd.has_ids) std::fill(tc_ids.begin() + c * 512, tc_ids.begin() + (c + 1) * 512, -1);
        for (int pair = 0; pair < 4; ++pair) {  // k-slots 8G + e, G in {2(pair>>1), +1}, e in 4(pair&1) + 0..3
          unsigned used = 0;
          for (int G = 2 * (pair >> 1); G < 2 * (pair >> 1) + 2; ++G)
            for (int e = 4 * (pair & 1); e < 4 * (pair & 1) + 4; ++e)
              if (slots[j][8 * G + e] >= 0) used |= 1u << ((sc[slots[j][8 * G + e]].col - koff) & 7);
          for (int G = 2 * (pair >> 1); G < 2 * (pair >> 1) + 2; ++G)
            for (int e = 4 * (pair & 1); e < 4 * (pair & 1) + 4; ++e) {
              const int k = 8 * G + e;
              const int32_t q = slots[j][k];
              int32_t col = (int32_t)koff;
              uint32_t word = 0;
              int32_t id = -1;
              if (q >= 0) {
                col = sc[q].col;
                word = (uint32_t)to_bf16_bits(sc[q].v) | (sc[q].row >= 8 ? 1u << 16 : 0u);
                id = sc[q].id;
              } else {
                for (int r = 0; r < 8 && r < ksz; ++r)
                  if (!(used >> r & 1u)) {
                    col = (int32_t)(koff + r);
                    used |= 1u << r;
                    break;
                  }
              }
              tc_cols[(size_t)c * 32 + k] = col;
              std::memcpy(&tc_vals[(size_t)c * 1024 + 4 * k], &word, 4);
              if (bd.has_ids) {
                const size_t at = ((size_t)c * 64 + k / 4) * 8 + 2 * (k % 4);
                tc_ids[at] = id;
                tc_ids[at + 1] = stile::kKeepId;
              }
            }
        }
      }
    }
  }

  Layout lay;
  const size_t o_rows = lay.add(bd.win_rows.size() * 4);
  const size_t o_tcptr = lay.add(tc_ptr.size() * 4);
  const size_t o_tccols = lay.add(tc_cols.size() * 4);
  const size_t o_tcvals = lay.add(tc_vals.size());
  const size_t o_scale = lay.add(tc_scale.size() * 4);
  const size_t o_resptr = lay.add(res_ptr.size() * 4);
  const size_t o_rescols = lay.add(res_cols.size() * 4);
  const size_t o_resvals = lay.add(res_vals.size() * 4);
  // the fix-up kernel's piece order (see SpmmPlanDev::piece_exec)
  std::vector<int32_t> piece_exec(piece_row.size());
  for (size_t q = 0; q < piece_exec.size(); ++q) piece_exec[q] = (int32_t)q;
  auto by_first_col = [&](int32_t a, int32_t b) {
    const int32_t ca = res_cols[piece_beg[a]], cb = res_cols[piece_beg[b]];
    return ca != cb ? ca < cb : a < b;
  };
  std::sort(piece_exec.begin(), piece_exec.begin() + npieces_dense, by_first_col);
  std::sort(piece_exec.begin() + npieces_dense, piece_exec.end(), by_first_col);
  const size_t o_prow = lay.add(piece_row.size() * 4);
  const size_t o_pexec = lay.add(piece_exec.size() * 4);
  const size_t o_pbeg = lay.add(piece_beg.size() * 4);
  const size_t o_pend = lay.add(piece_end.size() * 4);
  const size_t o_tcids = lay.add(tc_ids.size() * 4);
  const size_t o_resids = lay.add(res_ids.size() * 4);
  const size_t o_mix = lay.add(mix_slot.size() * 4);
  const size_t o_mask = lay.add(tc_mask.size() * 4);
  const size_t o_pwin = lay.add(part_win.size() * 4);
  const size_t o_pc0 = lay.add(part_c0.size() * 4);
  const size_t o_pc1 = lay.add(part_c1.size() * 4);
  const size_t o_units = lay.add(lds_units.size() * 4);
  const size_t o_srow = lay.add(srow_row.size() * 4);
  const size_t o_sp0 = lay.add(srow_p0.size() * 4);
  const size_t o_sp1 = lay.add(srow_p1.size() * 4);
  const size_t o_hwp0 = lay.add(hw_p0.size() * 4);
  const size_t o_spoff = lay.add(sp_off.size() * 4);
  const size_t o_spent = lay.add(sp_ent.size() * 4);
  const size_t o_spids = lay.add(sp_ids.size() * 4);
  const size_t o_kind = lay.add(tc_kind.size() * 4);
  // first row of each window whose 16 rows are consecutive (else -1): the
  // B-stationary kernel then computes the rows instead of loading them
  std::vector<int32_t> win_r0(lds_ok ? W : 0, -1);
  for (int32_t w = 0; lds_ok && w < W; ++w) {
    const int32_t r0 = bd.win_rows[(size_t)w * 16];
    bool run = r0 >= 0;
    for (int r = 1; r < 16 && run; ++r) run = bd.win_rows[(size_t)w * 16 + r] == r0 + r;
    if (run) win_r0[w] = r0;
  }
  const size_t o_r0 = lay.add(win_r0.size() * 4);
  HostBlob blob;
  blob.data.assign(lay.bytes, 0);
  blob.put(o_rows, bd.win_rows.data(), bd.win_rows.size() * 4);
  blob.put(o_tcptr, tc_ptr.data(), tc_ptr.size() * 4);
  blob.put(o_tccols, tc_cols.data(), tc_cols.size() * 4);
  blob.put(o_tcvals, tc_vals.data(), tc_vals.size());
  blob.put(o_scale, tc_scale.data(), tc_scale.size() * 4);
  blob.put(o_resptr, res_ptr.data(), res_ptr.size() * 4);
  blob.put(o_rescols, res_cols.data(), res_cols.size() * 4);
  blob.put(o_resvals, res_vals.data(), res_vals.size() * 4);
  blob.put(o_prow, piece_row.data(), piece_row.size() * 4);
  blob.put(o_pexec, piece_exec.data(), piece_exec.size() * 4);
  blob.put(o_pbeg, piece_beg.data(), piece_beg.size() * 4);
  blob.put(o_pend, piece_end.data(), piece_end.size() * 4);
  blob.put(o_tcids, tc_ids.data(), tc_ids.size() * 4);
  blob.put(o_resids, res_ids.data(), res_ids.size() * 4);
  blob.put(o_mix, mix_slot.data(), mix_slot.size() * 4);
  blob.put(o_mask, tc_mask.data(), tc_mask.size() * 4);
  blob.put(o_pwin, part_win.data(), part_win.size() * 4);
  blob.put(o_pc0, part_c0.data(), part_c0.size() * 4);
  blob.put(o_pc1, part_c1.data(), part_c1.size() * 4);
  blob.put(o_units, lds_units.data(), lds_units.size() * 4);
  blob.put(o_srow, srow_row.data(), srow_row.size() * 4);
  blob.put(o_sp0, srow_p0.data(), srow_p0.size() * 4);
  blob.put(o_sp1, srow_p1.data(), srow_p1.size() * 4);
  blob.put(o_hwp0, hw_p0.data(), hw_p0.size() * 4);
  blob.put(o_spoff, sp_off.data(), sp_off.size() * 4);
  blob.put(o_spent, sp_ent.data(), sp_ent.size() * 4);
  blob.put(o_spids, sp_ids.data(), sp_ids.size() * 4);
  blob.put(o_kind, tc_kind.data(), tc_kind.size() * 4);
  blob.put(o_r0, win_r0.data(), win_r0.size() * 4);

  StilePlan* plan = new (std::nothrow) StilePlan();
  if (!plan) return fail(STILE_E_NOMEM, "host allocation of plan");
  plan->kind = 1;
  plan->blk_dtype = blk_dtype;
  int rc = upload(plan, lay, blob);
  if (rc != STILE_OK) {
    if (plan->dev) (void)hipFree(plan->dev);
    delete plan;
    return rc;
  }
  stile::SpmmPlanDev& d = plan->spmm;
  d.M = bd.M;
  d.K = bd.K;
  d.W = W;
  d.W_tc = W_tc;
  d.W_big = W_big;
  d.W_huge = W_huge;
  d.n_parts = (int32_t)part_win.size();
  d.part_win = dptr<int32_t>(plan, o_pwin);
  d.part_c0 = dptr<int32_t>(plan, o_pc0);
  d.part_c1 = dptr<int32_t>(plan, o_pc1);
  d.n_mix = (int32_t)mix_slot.size();
  d.mix_slot = dptr<int32_t>(plan, o_mix);
  d.tc_mask = tc_mask.empty() ? nullptr : dptr<uint32_t>(plan, o_mask);
  d.lds_units = lds_units.empty() ? nullptr : dptr<int4>(plan, o_units);
  for (int t = 0; t < stile::kLdsLists; ++t) {
    d.lds_off[t] = lds_off[t];
    d.lds_cnt[t] = lds_cnt[t];
  }
  d.lds_max_k = lds_max_k;
  std::copy(std::begin(lds_pick), std::end(lds_pick), d.lds_pick);
  plan->lds_xcd = std::move(lds_xcd);
  d.lds_xcd_host = plan->lds_xcd.empty() ? nullptr : plan->lds_xcd.data();
  d.nchunks = (int32_t)nchunks;
  d.npieces = (int32_t)piece_row.size();
  d.npieces_dense = npieces_dense;
  d.nsrows = (int32_t)srow_row.size();
  d.nsrows_dense = nsrows_dense;
  d.srow_row = dptr<int32_t>(plan, o_srow);
  d.srow_p0 = dptr<int32_t>(plan, o_sp0);
  d.srow_p1 = dptr<int32_t>(plan, o_sp1);
  d.hw_p0 = dptr<int32_t>(plan, o_hwp0);
  d.sp_off = sp_off.empty() ? nullptr : dptr<int32_t>(plan, o_spoff);
  d.sp_ent = sp_off.empty() ? nullptr : dptr<int2>(plan, o_spent);
  d.tc_kind = tc_kind.empty() ? nullptr : dptr<uint32_t>(plan, o_kind);
  d.win_r0 = win_r0.empty() ? nullptr : dptr<int32_t>(plan, o_r0);
  d.opt_kernel = 0;
  d.opt_lds_list = -1;
  d.opt_lds_xcd = 1;
  plan->n_sp = (int64_t)sp_ids.size();
  plan->sp_ids = sp_ids.empty() ? nullptr : dptr<int32_t>(plan, o_spids);
  d.piece_row = dptr<int32_t>(plan, o_prow);
  d.piece_exec = dptr<int32_t>(plan, o_pexec);
  d.piece_beg = dptr<int32_t>(plan, o_pbeg);
  d.piece_end = dptr<int32_t>(plan, o_pend);
  d.blk_dtype = blk_dtype;
  d.win_rows = dptr<int32_t>(plan, o_rows);
  d.tc_ptr = dptr<int32_t>(plan, o_tcptr);
  d.tc_cols = dptr<int32_t>(plan, o_tccols);
  d.tc_vals = dptr<unsigned char>(plan, o_tcvals);
  d.tc_scale = tc_scale.empty() ? nullptr : dptr<float>(plan, o_scale);
  d.res_ptr = dptr<int32_t>(plan, o_resptr);
  d.nnz_sparse_only = res_ptr[(size_t)W * 16] - res_ptr[(size_t)W_tc * 16];
  d.res_cols = dptr<int32_t>(plan, o_rescols);
  d.res_vals = dptr<float>(plan, o_resvals);
  plan->has_ids = bd.has_ids;
  plan->max_id = bd.max_id;
  plan->n_tc_slots = (int64_t)tc_ids.size();
  plan->n_res = nres;
  plan->tc_ids = bd.has_ids ? dptr<int32_t>(plan, o_tcids) : nullptr;
  plan->res_ids = bd.has_ids ? dptr<int32_t>(plan, o_resids) : nullptr;
  StilePlanInfo& in = plan->info;
  in.kind = 1;
  in.blk_dtype = blk_dtype;
  in.M = bd.M;
  in.K = bd.K;
  in.num_windows = W;
  in.num_chunks = (int32_t)nchunks;
  in.block_nnz = bd.block_nnz;
  in.sparse_nnz = nres;
  in.device_bytes = (int64_t)plan->dev_bytes;
  in.hub_rows = hub_rows;
  in.num_split = (int32_t)piece_row.size();
  stile::prepare_spmm_kernels(d);
  *out = plan;
  return STILE_OK;
}

int create_spmm(const StileSpmmPack* packs, int32_t count, int32_t blk_dtype, StilePlan** out,
                bool want_mask = false, int32_t scatter_max = 0) {
  if (!out) return fail(STILE_E_INVALID, "null output pointer");
  *out = nullptr;
  if (!packs || count <= 0) return fail(STILE_E_INVALID, "no packs");
  if (blk_dtype < STILE_BLK_F32 || blk_dtype > STILE_BLK_F16)
    return fail(STILE_E_INVALID, "unknown blk_dtype");
  SpmmBuild bd;
  bd.want_mask = want_mask;
  int64_t roff = 0, coff = 0, ioff = 0;
  for (int32_t i = 0; i < count; ++i) {
    const size_t w0 = bd.win_rows.size() / 16;
    const size_t r0 = bd.row_win.size();
    // value ids of matrix i are offset past the largest id of the previous
    // ones: the batched values array is the concatenation of the matrices'
    // value arrays, each as long as its largest id + 1
    ioff = bd.max_id + 1;
    bd.inst_rows.push_back(roff);
    bd.inst_cols.push_back(coff);
    int rc = add_spmm_pack(bd, &packs[i], (int32_t)roff, (int32_t)coff, ioff);
    if (rc != STILE_OK) return rc;
    (void)w0;
    (void)r0;
    roff += packs[i].M;
    coff += packs[i].K;
    if (roff > INT32_MAX || coff > INT32_MAX) return fail(STILE_E_UNSUPPORTED, "batched size exceeds 2^31");
  }
  bd.M = (int32_t)roff;
  bd.K = (int32_t)coff;
  bd.inst_rows.push_back(roff);
  bd.inst_cols.push_back(coff);
  return finish_spmm(bd, blk_dtype, scatter_max, out);
}

// ---------------------------------------------------------------------------
// SDDMM plan construction.
// ---------------------------------------------------------------------------
struct SddmmBuild {
  int32_t M = 0, N = 0, nnz = 0;
  std::vector<int32_t> e_row, e_col, e_out;
  std::vector<int32_t> b_rows, b_cols, b_ptr{0}, b_pos, b_out;
};

int add_sddmm_pack(SddmmBuild& bd, const StileSddmmPack* pk, int32_t roff, int32_t coff, int32_t ooff) {
  if (!pk) return fail(STILE_E_INVALID, "null pack");
  if (pk->M < 0 || pk->N < 0 || pk->nnz < 0) return fail(STILE_E_INVALID, "negative size");
  if ((pk->num_1d > 0 && !pk->g1d) || (pk->num_tc > 0 && !pk->tc) || pk->num_1d < 0 || pk->num_tc < 0)
    return fail(STILE_E_INVALID, "group array missing");
  // Each output entry is computed once: TC blocks first (in group and block
  // order), then the 1D entries not already computed.  An entry covered twice
  // (overlapping tiles, padding repeats) would otherwise be written by two
  // threads with different rounding in no fixed order.
  std::vector<unsigned char> seen(pk->nnz, 0);
  for (int gi = 0; gi < pk->num_tc; ++gi) {
    const StileSddmmTcGroup& g = pk->tc[gi];
    if (g.nnzb < 0) return fail(STILE_E_INVALID, "TC group: negative nnzb");
    if (g.blk_k != 16) return fail(STILE_E_UNSUPPORTED, "TC SDDMM blocks must be 16 columns wide");
    if (g.nnzb > 0 && (!g.i_indices || !g.j_indices || !g.indptr || !g.indices_dense || !g.indices))
      return fail(STILE_E_INVALID, "TC group: null buffer");
    if (g.nnzb > 0 && g.indptr[0] != 0) return fail(STILE_E_INVALID, "TC group: indptr[0] != 0");
    for (int32_t b = 0; b < g.nnzb; ++b) {
      for (int r = 0; r < 16; ++r) {
        const int32_t ri = g.i_indices[(int64_t)b * 16 + r], ci = g.j_indices[(int64_t)b * 16 + r];
        if (ri < 0 || ri >= pk->M || ci < 0 || ci >= pk->N)
          return fail(STILE_E_INVALID, "TC group: block row/col out of range");
        bd.b_rows.push_back(ri + roff);
        bd.b_cols.push_back(ci + coff);
      }
      const int32_t s0 = g.indptr[b], s1 = g.indptr[b + 1];
      if (s1 < s0) return fail(STILE_E_INVALID, "TC group: indptr not monotone");
      for (int32_t s = s0; s < s1; ++s) {
        const int32_t pos = g.indices_dense[s], o = g.indices[s];
        if (pos < 0 || pos >= 256 || o < 0 || o >= pk->nnz)
          return fail(STILE_E_INVALID, "TC group: sampled entry out of range");
        if (seen[o]) continue;
        seen[o] = 1;
        bd.b_pos.push_back(pos);
        bd.b_out.push_back(o + ooff);
      }
      bd.b_ptr.push_back((int32_t)bd.b_pos.size());
    }
  }
  for (int gi = 0; gi < pk->num_1d; ++gi) {
    const StileSddmm1DGroup& g = pk->g1d[gi];
    if (g.len < 0 || (g.len > 0 && (!g.mid_idx || !g.j_idx || !g.remap)))
      return fail(STILE_E_INVALID, "1D group: bad buffers");
    // trailing padding repeats the last entry (my_fuse_formats.py:3586-3588)
    // and is dropped with the other duplicates
    for (int32_t e = 0; e < g.len; ++e) {
      const int32_t r = g.mid_idx[e], c = g.j_idx[e], o = g.remap[e];
      if (r < 0 || r >= pk->M || c < 0 || c >= pk->N || o < 0 || o >= pk->nnz)
        return fail(STILE_E_INVALID, "1D group: index out of range");
      if (seen[o]) continue;
      seen[o] = 1;
      bd.e_row.push_back(r + roff);
      bd.e_col.push_back(c + coff);
      bd.e_out.push_back(o + ooff);
    }
  }
  for (int32_t o = 0; o < pk->nnz; ++o)
    if (!seen[o]) return fail(STILE_E_INVALID, "packed groups leave an output entry uncovered");
  return STILE_OK;
}

int create_sddmm(const StileSddmmPack* packs, int32_t count, int32_t blk_dtype, StilePlan** out) {
  if (!out) return fail(STILE_E_INVALID, "null output pointer");
  *out = nullptr;
  if (!packs || count <= 0) return fail(STILE_E_INVALID, "no packs");
  if (blk_dtype != STILE_BLK_F32)
    return fail(STILE_E_UNSUPPORTED, "SDDMM blocks run FP32 MFMA only");
  SddmmBuild bd;
  int64_t roff = 0, coff = 0, ooff = 0;
  for (int32_t i = 0; i < count; ++i) {
    int rc = add_sddmm_pack(bd, &packs[i], (int32_t)roff, (int32_t)coff, (int32_t)ooff);
    if (rc != STILE_OK) return rc;
    roff += packs[i].M;
    coff += packs[i].N;
    ooff += packs[i].nnz;
    if (roff > INT32_MAX || coff > INT32_MAX || ooff > INT32_MAX)
      return fail(STILE_E_UNSUPPORTED, "batched size exceeds 2^31");
  }
  bd.M = (int32_t)roff;
  bd.N = (int32_t)coff;
  bd.nnz = (int32_t)ooff;
  Layout lay;
  const size_t o_er = lay.add(bd.e_row.size() * 4), o_ec = lay.add(bd.e_col.size() * 4),
               o_eo = lay.add(bd.e_out.size() * 4), o_br = lay.add(bd.b_rows.size() * 4),
               o_bc = lay.add(bd.b_cols.size() * 4), o_bp = lay.add(bd.b_ptr.size() * 4),
               o_bs = lay.add(bd.b_pos.size() * 4), o_bo = lay.add(bd.b_out.size() * 4);
  HostBlob blob;
  blob.data.assign(lay.bytes, 0);
  blob.put(o_er, bd.e_row.data(), bd.e_row.size() * 4);
  blob.put(o_ec, bd.e_col.data(), bd.e_col.size() * 4);
  blob.put(o_eo, bd.e_out.data(), bd.e_out.size() * 4);
  blob.put(o_br, bd.b_rows.data(), bd.b_rows.size() * 4);
  blob.put(o_bc, bd.b_cols.data(), bd.b_cols.size() * 4);
  blob.put(o_bp, bd.b_ptr.data(), bd.b_ptr.size() * 4);
  blob.put(o_bs, bd.b_pos.data(), bd.b_pos.size() * 4);
  blob.put(o_bo, bd.b_out.data(), bd.b_out.size() * 4);
  StilePlan* plan = new (std::nothrow) StilePlan();
  if (!plan) return fail(STILE_E_NOMEM, "host allocation of plan");
  plan->kind = 2;
  plan->blk_dtype = blk_dtype;
  int rc = upload(plan, lay, blob);
  if (rc != STILE_OK) {
    if (plan->dev) (void)hipFree(plan->dev);
    delete plan;
    return rc;
  }
  stile::SddmmPlanDev& d = plan->sddmm;
  d.M = bd.M;
  d.N = bd.N;
  d.nnz = bd.nnz;
  d.n1d = (int32_t)bd.e_row.size();
  d.nblk = (int32_t)(bd.b_ptr.size() - 1);
  d.blk_dtype = blk_dtype;
  d.e_row = dptr<int32_t>(plan, o_er);
  d.e_col = dptr<int32_t>(plan, o_ec);
  d.e_out = dptr<int32_t>(plan, o_eo);
  d.blk_rows = dptr<int32_t>(plan, o_br);
  d.blk_cols = dptr<int32_t>(plan, o_bc);
  d.blk_ptr = dptr<int32_t>(plan, o_bp);
  d.blk_pos = dptr<int32_t>(plan, o_bs);
  d.blk_out = dptr<int32_t>(plan, o_bo);
  StilePlanInfo& in = plan->info;
  in.kind = 2;
  in.blk_dtype = blk_dtype;
  in.M = bd.M;
  in.K = bd.N;
  in.num_windows = 0;
  in.num_chunks = d.nblk;
  in.block_nnz = (int64_t)bd.b_pos.size();
  in.sparse_nnz = (int64_t)bd.e_row.size();
  in.device_bytes = (int64_t)plan->dev_bytes;
  in.hub_rows = 0;
  in.num_split = 0;
  *out = plan;
  return STILE_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// stile_spmm_run_noalloc: the deterministic-mode workspace in this code
// object's own device memory, STILE_NOALLOC_SCRATCH_BYTES per GPU, held
// until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only
// that address is handed to kernels.  The workspace needs no clearing: the
// pieces and parts kernels store every element before the reduce kernels
// read it.  Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[STILE_NOALLOC_SCRATCH_BYTES / 4];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices] = {};

int device_scratch(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(STILE_E_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress of the static scratch");
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return STILE_OK;
}

}  // namespace

extern "C" {

int32_t stile_abi_version(void) { return 3; }

const char* stile_last_error(void) { return g_err.c_str(); }

int stile_csr_spmm_f32(int32_t M, int32_t K, int32_t N, const int32_t* indptr, const int32_t* indices,
                       const float* vals, const float* B, int64_t ldb, float* C, int64_t ldc, void* stream) {
  if (M < 0 || K < 0 || N < 0) return fail(STILE_E_INVALID, "negative size");
  if (M == 0 || N == 0) return STILE_OK;
  if (!indptr || !C || (K > 0 && !B)) return fail(STILE_E_INVALID, "null pointer");
  if (ldb < N || ldc < N) return fail(STILE_E_INVALID, "leading dimension smaller than N");
  if (K == 0) {
    hipError_t e = stile::launch_fill_zero_2d(C, M, N, ldc, (hipStream_t)stream);
    return e == hipSuccess ? STILE_OK : hip_fail(e, "fill_zero launch");
  }
  hipError_t e = stile::launch_csr_spmm_f32(M, N, indptr, indices, vals, B, ldb, C, ldc, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "csr_spmm launch");
}

int stile_csr_sddmm_f32(int32_t M, int32_t N, int32_t nnz, int32_t Kf, const int32_t* indptr,
                        const int32_t* indices, const float* X, int64_t ldx, const float* Y, int64_t ldy,
                        float* out, void* stream) {
  (void)N;
  if (M < 0 || N < 0 || nnz < 0 || Kf < 0) return fail(STILE_E_INVALID, "negative size");
  if (nnz == 0) return STILE_OK;
  if (!indptr || !indices || !out) return fail(STILE_E_INVALID, "null pointer");
  if (Kf == 0) {
    hipError_t e = stile::launch_fill_zero(out, nnz, (hipStream_t)stream);
    return e == hipSuccess ? STILE_OK : hip_fail(e, "fill_zero launch");
  }
  if (!X || !Y) return fail(STILE_E_INVALID, "null pointer");
  if (ldx < Kf || ldy < Kf) return fail(STILE_E_INVALID, "leading dimension smaller than Kf");
  hipError_t e = stile::launch_csr_sddmm_f32(M, nnz, Kf, indptr, indices, X, ldx, Y, ldy, out, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "csr_sddmm launch");
}

int stile_csr_row_softmax_f32(int32_t M, const int32_t* indptr, const float* in, float* out, float scale,
                              void* stream) {
  if (M < 0) return fail(STILE_E_INVALID, "negative M");
  if (M == 0) return STILE_OK;
  if (!indptr || !in || !out) return fail(STILE_E_INVALID, "null pointer");
  hipError_t e = stile::launch_row_softmax(M, indptr, in, out, scale, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "row_softmax launch");
}

int stile_ell_buckets_spmm_f32(const StileEllBucket* buckets, int32_t count, int32_t M, int32_t N, const float* B,
                               int64_t ldb, float* C, int64_t ldc, void* stream) {
  if (M < 0 || N < 0 || count < 0) return fail(STILE_E_INVALID, "negative size");
  if (M == 0 || N == 0) return STILE_OK;
  if (!C || (count > 0 && (!buckets || !B))) return fail(STILE_E_INVALID, "null pointer");
  if (ldb < N || ldc < N) return fail(STILE_E_INVALID, "leading dimension smaller than N");
  for (int32_t i = 0; i < count; ++i) {
    const StileEllBucket& b = buckets[i];
    if (b.rows < 0 || b.width <= 0) return fail(STILE_E_INVALID, "bucket: rows < 0 or width <= 0");
    if (b.rows > 0 && (!b.row || !b.cols || !b.vals)) return fail(STILE_E_INVALID, "bucket: null buffer");
  }
  hipError_t e = stile::launch_fill_zero_2d(C, M, N, ldc, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "fill_zero launch");
  for (int32_t i = 0; i < count; ++i) {
    const StileEllBucket& b = buckets[i];
    e = stile::launch_ell_bucket_spmm(b.rows, b.width, b.row, b.cols, b.vals, B, ldb, N, C, ldc,
                                      (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "ell_bucket_spmm launch");
  }
  return STILE_OK;
}

int stile_plan_create_spmm(const StileSpmmPack* pack, int32_t blk_dtype, StilePlan** out) {
  return create_spmm(pack, 1, blk_dtype, out);
}

int stile_plan_create_spmm_batched(const StileSpmmPack* packs, int32_t count, int32_t blk_dtype,
                                   StilePlan** out) {
  return create_spmm(packs, count, blk_dtype, out);
}

int stile_plan_create_spmm_ex(const StileSpmmPack* packs, int32_t count, int32_t blk_dtype,
                              const StileSpmmOptions* opts, StilePlan** out) {
  const int32_t scatter = opts ? opts->scatter_max_nnz : 0;
  if (scatter < 0 || scatter > 16) return fail(STILE_E_INVALID, "scatter_max_nnz must be in 0..16");
  return create_spmm(packs, count, blk_dtype, out, false, scatter);
}

int stile_plan_set_option(StilePlan* plan, int32_t option, int32_t value) {
  if (!plan || plan->kind != 1) return fail(STILE_E_INVALID, "not an SpMM plan");
  stile::SpmmPlanDev& d = plan->spmm;
  switch (option) {
    case STILE_OPT_KERNEL:
      if (value != 0 && value != 1) return fail(STILE_E_INVALID, "STILE_OPT_KERNEL: 0 or 1");
      if (value == 1 && d.tc_kind)
        return fail(STILE_E_UNSUPPORTED, "a plan with scatter chunks runs on the B-stationary kernel only");
      d.opt_kernel = value;
      return STILE_OK;
    case STILE_OPT_LDS_LIST:
      if (value < -1 || value >= stile::kLdsLists) return fail(STILE_E_INVALID, "STILE_OPT_LDS_LIST: -1 .. 11");
      if (value >= 0 && d.lds_units && d.lds_cnt[value] == 0)
        return fail(STILE_E_INVALID, "STILE_OPT_LDS_LIST: the plan has no units in that list");
      d.opt_lds_list = value;
      return STILE_OK;
    case STILE_OPT_LDS_XCD:
      if (value != 0 && value != 1) return fail(STILE_E_INVALID, "STILE_OPT_LDS_XCD: 0 or 1");
      d.opt_lds_xcd = value;
      return STILE_OK;
    default:
      return fail(STILE_E_INVALID, "unknown option");
  }
}

int stile_plan_create_attention(const StileSpmmPack* packs, int32_t count, StilePlan** out) {
  for (int32_t i = 0; packs && i < count; ++i)
    if (packs[i].M != packs[i].K) return fail(STILE_E_INVALID, "attention masks must be square");
  return create_spmm(packs, count, STILE_BLK_BF16, out, true);
}

int stile_attention_run(const StilePlan* plan, const void* Q, const void* K, const void* V, int64_t ld, int32_t d,
                        float scale, float* out, int64_t ldo, void* stream) {
  if (!plan || plan->kind != 1 || (plan->spmm.nchunks > 0 && !plan->spmm.tc_mask))
    return fail(STILE_E_INVALID, "not an attention plan (stile_plan_create_attention)");
  if (d != 64) return fail(STILE_E_UNSUPPORTED, "fused attention supports head dimension 64");
  if (!(scale > 0.f)) return fail(STILE_E_INVALID, "scale must be positive");
  if (plan->spmm.W == 0) return STILE_OK;
  if (!Q || !K || !V || !out) return fail(STILE_E_INVALID, "null pointer");
  if (ld < d || ld % 8 != 0 || ldo < d || ldo % 4 != 0)
    return fail(STILE_E_INVALID, "leading dimensions: ld >= d and a multiple of 8, ldo >= d and a multiple of 4");
  if ((reinterpret_cast<uintptr_t>(Q) | reinterpret_cast<uintptr_t>(K) | reinterpret_cast<uintptr_t>(V)) & 15 ||
      reinterpret_cast<uintptr_t>(out) & 15)
    return fail(STILE_E_INVALID, "Q, K, V and out must be 16-byte aligned");
  hipError_t e = stile::launch_attention(plan->spmm, Q, K, V, ld, scale, out, ldo, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "attention launch");
}

int stile_plan_create_sddmm(const StileSddmmPack* pack, int32_t blk_dtype, StilePlan** out) {
  return create_sddmm(pack, 1, blk_dtype, out);
}

int stile_plan_create_sddmm_batched(const StileSddmmPack* packs, int32_t count, int32_t blk_dtype,
                                    StilePlan** out) {
  return create_sddmm(packs, count, blk_dtype, out);
}

int stile_spmm_workspace_bytes(const StilePlan* plan, int32_t N, int32_t flags, int64_t* bytes) {
  if (!plan || plan->kind != 1) return fail(STILE_E_INVALID, "not an SpMM plan");
  if (!bytes) return fail(STILE_E_INVALID, "null output pointer");
  if (N < 0) return fail(STILE_E_INVALID, "negative N");
  if (flags & ~STILE_RUN_DETERMINISTIC) return fail(STILE_E_INVALID, "unknown run flags");
  *bytes = (flags & STILE_RUN_DETERMINISTIC) ? stile::spmm_workspace_floats(plan->spmm, N) * 4 : 0;
  return STILE_OK;
}

int stile_spmm_run_ex(const StilePlan* plan, int32_t b_dtype, const void* B, int64_t ldb, int32_t N, float* C,
                      int64_t ldc, int32_t flags, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!plan || plan->kind != 1) return fail(STILE_E_INVALID, "not an SpMM plan");
  if (b_dtype < STILE_BLK_F32 || b_dtype > STILE_BLK_F16) return fail(STILE_E_INVALID, "unknown b_dtype");
  if (flags & ~STILE_RUN_DETERMINISTIC) return fail(STILE_E_INVALID, "unknown run flags");
  if (N < 0) return fail(STILE_E_INVALID, "negative N");
  if (N == 0 || plan->spmm.W == 0) return STILE_OK;
  if (!C || (plan->spmm.K > 0 && !B)) return fail(STILE_E_INVALID, "null pointer");
  if (ldb < N || ldc < N) return fail(STILE_E_INVALID, "leading dimension smaller than N");
  const bool det = (flags & STILE_RUN_DETERMINISTIC) != 0;
  const int64_t need = det ? stile::spmm_workspace_floats(plan->spmm, N) * 4 : 0;
  if (need > 0) {
    if (!workspace || workspace_bytes < need)
      return fail(STILE_E_INVALID, "workspace smaller than stile_spmm_workspace_bytes");
    if (reinterpret_cast<uintptr_t>(workspace) & 15) return fail(STILE_E_INVALID, "workspace not 16-byte aligned");
  }
  hipError_t e = stile::launch_hyb_spmm(plan->spmm, b_dtype, B, ldb, N, C, ldc, det,
                                        static_cast<float*>(workspace), (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "hyb_spmm launch");
}

int stile_spmm_run_noalloc(const StilePlan* plan, int32_t b_dtype, const void* B, int64_t ldb, int32_t N, float* C,
                           int64_t ldc, int32_t flags, void* stream) {
  if (!plan || plan->kind != 1) return fail(STILE_E_INVALID, "not an SpMM plan");
  if (flags & ~STILE_RUN_DETERMINISTIC) return fail(STILE_E_INVALID, "unknown run flags");
  if (N < 0) return fail(STILE_E_INVALID, "negative N");
  const int64_t need =
      (flags & STILE_RUN_DETERMINISTIC) ? stile::spmm_workspace_floats(plan->spmm, N) * 4 : 0;
  if (need > (int64_t)STILE_NOALLOC_SCRATCH_BYTES)
    return fail(STILE_E_INVALID, "deterministic workspace of " + std::to_string(need) +
                                     " bytes exceeds the static scratch of STILE_NOALLOC_SCRATCH_BYTES = " +
                                     std::to_string((int64_t)STILE_NOALLOC_SCRATCH_BYTES) + " bytes");
  float* ws = nullptr;
  if (need > 0) {
    const int rc = device_scratch(&ws);
    if (rc != STILE_OK) return rc;
  }
  return stile_spmm_run_ex(plan, b_dtype, B, ldb, N, C, ldc, flags, ws, need, stream);
}

int stile_spmm_run(const StilePlan* plan, const float* B, int64_t ldb, int32_t N, float* C, int64_t ldc,
                   void* stream) {
  return stile_spmm_run_ex(plan, STILE_BLK_F32, B, ldb, N, C, ldc, 0, nullptr, 0, stream);
}

int stile_sddmm_run(const StilePlan* plan, const float* X, int64_t ldx, const float* Y, int64_t ldy, int32_t Kf,
                    float* out, void* stream) {
  if (!plan || plan->kind != 2) return fail(STILE_E_INVALID, "not an SDDMM plan");
  if (Kf < 0) return fail(STILE_E_INVALID, "negative Kf");
  if (plan->sddmm.nnz == 0) return STILE_OK;
  if (!out || !X || !Y) return fail(STILE_E_INVALID, "null pointer");
  if (ldx < Kf || ldy < Kf) return fail(STILE_E_INVALID, "leading dimension smaller than Kf");
  if (Kf == 0) {
    hipError_t e = stile::launch_fill_zero(out, plan->sddmm.nnz, (hipStream_t)stream);
    return e == hipSuccess ? STILE_OK : hip_fail(e, "fill_zero launch");
  }
  hipError_t e = stile::launch_hyb_sddmm(plan->sddmm, X, ldx, Y, ldy, Kf, out, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "hyb_sddmm launch");
}

int stile_plan_set_values(StilePlan* plan, const float* vals, int64_t nnz, void* stream) {
  if (!plan || plan->kind != 1) return fail(STILE_E_INVALID, "not an SpMM plan");
  if (!plan->has_ids) return fail(STILE_E_INVALID, "plan was built without value ids (a_ids)");
  if (plan->blk_dtype == STILE_BLK_F16) return fail(STILE_E_UNSUPPORTED, "value refresh of FP16 plans");
  if (nnz <= plan->max_id) return fail(STILE_E_INVALID, "values array shorter than the plan's largest value id");
  if (!vals && (plan->n_tc_slots + plan->n_res) > 0) return fail(STILE_E_INVALID, "null values");
  hipError_t e = stile::launch_set_values(plan->spmm, vals, plan->tc_ids, plan->n_tc_slots, plan->res_ids,
                                          plan->n_res, plan->sp_ids, plan->n_sp, (hipStream_t)stream);
  return e == hipSuccess ? STILE_OK : hip_fail(e, "set_values launch");
}

int stile_plan_info(const StilePlan* plan, StilePlanInfo* info) {
  if (!plan || !info) return fail(STILE_E_INVALID, "null pointer");
  *info = plan->info;
  return STILE_OK;
}

int stile_plan_destroy(StilePlan* plan) {
  if (!plan) return STILE_OK;
  int rc = STILE_OK;
  if (plan->dev) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != plan->device) (void)hipSetDevice(plan->device);
    hipError_t e = hipFree(plan->dev);
    if (cur != plan->device) (void)hipSetDevice(cur);
    if (e != hipSuccess) rc = hip_fail(e, "hipFree of plan");
  }
  delete plan;
  return rc;
}

}  // extern "C"
