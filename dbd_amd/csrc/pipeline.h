// Internal declarations shared by the engine's translation units (not part of the C ABI):
// HIP status macros, category tables, and the capturable whole pipeline over raw workspace.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return set_error(DBD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                                      \
  } while (0)

#define CHK(expr)                        \
  do {                                   \
    dbd_status s_ = (expr);              \
    if (s_ != DBD_OK) return s_;         \
  } while (0)

namespace dbd {

// The autotuner's alternative to the default (hub range kernels at every width): hub rows take
// them from this many columns of B only.
constexpr int64_t kHubWideCols = (int64_t)1 << 21;

// A category table: thresholds and kernel shapes of the symbolic and numeric categories, and the
// knobs they were built from (the autotuner's search space).
struct Tables {
  CatTable sym_ct, num_ct;
  KernelShape sym_k[kMaxCats], num_k[kMaxCats];
  int id = 0;
  int g_small = 16, g_big = 64;  // lanes per A entry in tables of <= 1024 / larger slot counts
  double load = 0.5;             // hash-table load limit
  int top = 8192;                // slots of the largest numeric LDS table (symbolic: twice that)
  int64_t hub_min_cols = 0;  // numeric hub category used when ncols >= this
};

// Symbolic key min(P_i, ncols), numeric key nnz_i; hash tables sized so the load stays <= `load`
// (one-wave symbolic tables at half of that); G lanes per A entry for small / large tables; the
// numeric tables stop at `top` slots (4096 or 8192), the symbolic ones at 2 * top. Rows past the
// last table are long rows; the numeric table ends with the hub category, used when B has at
// least hub_min_cols columns.
Tables default_tables(int g_small, int g_big, double load, int top = 8192, int64_t hub_min_cols = 0);

// Whether the numeric binning routes hub rows (long rows of more than kHubEntries A entries) to
// the hub category for a problem of ncols columns.
inline bool hub_enabled(const Tables& t, int64_t ncols) { return ncols >= t.hub_min_cols; }

// Raw device workspace of one pipeline run: per-row arrays (nr + 1), bin counters and offsets
// (kMaxCats per 256-row tile), scan partials, per-A-nonzero B-row descriptors and long-row
// cursors, and the error word (followed by the sink degenerate rows are written to).
struct Work {
  int64_t *P, *nnz, *cnt, *offs_sym, *offs_num, *tmp, *bstart, *cursor;
  int32_t *rows_sym, *rows_num, *blen;
  int* err;
  int64_t nr, nwg;
  // hub rows (all null / 0: none): their slots and per-range column counts for the symbolic pass
  // and the numeric binning, and the enqueue of their range kernels after the numeric categories
  const int32_t* hub_slot = nullptr;
  int32_t* hub_rng = nullptr;
  int nrng = 0;
  void (*hub_fn)(void* arg, const int64_t* Cptr, int32_t* Cidx, void* Cval, hipStream_t s) = nullptr;
  void* hub_arg = nullptr;
};

// The whole pipeline (row analysis, bins, symbolic, scan, bins, numeric) on stream s with
// persistent category kernels that read their row ranges from device memory: no host
// synchronisation and no allocation, so it can be captured in a graph. If nnz(C) exceeds `cap`,
// the numeric kernels drain empty ranges and the capacity bit is set in the error word.
// Without w.hub_slot, hub rows stay in the long category.
void enqueue_pipeline(const Work& w, const Tables& tb, Mat A, const Mat& B, int dtype, int num_cu, int64_t* d_Cptr,
                      int32_t* d_Cidx, void* d_Cval, int64_t cap, hipStream_t s);

}  // namespace dbd
