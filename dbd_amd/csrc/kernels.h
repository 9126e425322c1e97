// Host-side launchers for the gfx950 kernels in kernels.hip.
//
// Every row kernel is grid-stride over a "bin" (list of row ids), and reads the bin's
// [begin, end) from device memory (two int64 slots). The plain pipeline launches exactly as
// many workgroups as the bin needs; the batched/graph pipeline launches a fixed persistent
// grid that drains the same ranges, so nothing depends on host-side counts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dbd {

constexpr int kMaxCats = 16;

// Category thresholds: row key k -> smallest c with k <= lim[c] (c = 0 is "k == 0"),
// else the last category (the long-row path).
struct CatTable {
  int64_t lim[kMaxCats];
  int32_t n;  // number of categories including 0 and the long category
};

struct Mat {
  const int64_t* ptr;
  const int32_t* idx;
  const void* val;
  int64_t nrows, ncols, nnz;
  // For A only: per-nonzero descriptor of the B row it selects (start, length), written by the
  // row analysis in A's order so later passes read it sequentially instead of gathering B.ptr.
  const int64_t* bstart = nullptr;
  const int32_t* blen = nullptr;
  // For B only: the band index of its long rows (launch_heavy_slots / launch_band_index), or null.
  // hslot[k] is row k's slot (-1: not banded); band + slot * nb1 are the slot's band positions.
  const int32_t* hslot = nullptr;
  const int32_t* band = nullptr;
  int64_t nb1 = 0;
};

// One category's kernel shape, chosen by the planner/autotuner.
struct KernelShape {
  int T;   // LDS hash slots (power of two); 0 = long-row window kernel; kHubShape = hub-row kernel
  int NT;  // threads per row group
  int G;   // lanes cooperating on one A entry (B row)
};
constexpr int kHubShape = -1;
// Rows with more A entries than this, among the rows of the long (last by key) category, form the
// hub category: their output is computed in column ranges, one workgroup per (row, range).
#ifndef DBD_HUB_ENTRIES
#define DBD_HUB_ENTRIES 1024
#endif
constexpr int64_t kHubEntries = DBD_HUB_ENTRIES;
// Column ranges of the hub-row kernels (16384 columns); the symbolic long windows hold whole ranges.
constexpr int kHubRangeShift = 14;
constexpr int kHubRange = 1 << kHubRangeShift;

// ---- generic device utilities ----
void launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s);
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
void launch_fill_bytes0(void* p, int64_t nbytes, hipStream_t s);
// Exclusive scan of n int64 (in may equal out); out[n] = total. tmp needs scan_tmp_elems(n).
int64_t scan_tmp_elems(int64_t n);
void launch_exclusive_scan(const int64_t* in, int64_t* out, int64_t n, int64_t* tmp, hipStream_t s);
// out[b] = ptr[min(b*stride, n)] for b in [0, nb]
void launch_gather_strided(const int64_t* ptr, int64_t n, int64_t stride, int64_t nb, int64_t* out, hipStream_t s);

// Inside a captured pipeline: if *total > cap, set the capacity bit (16) in *err and zero
// offs[0 .. n), which empties every bin range that the following row kernels read.
constexpr int kErrCapacityBit = 16;
void launch_capacity_guard(const int64_t* total, int64_t cap, int64_t* offs, int64_t n, int* err, hipStream_t s);

// ---- band index of long B rows ----
// Columns per band; band index rows have band_count(ncols) entries (a multiple of kBandPad
// bands, so every aligned long-row window lies inside, plus one entry holding the row length).
constexpr int kBandShift = 12;
constexpr int kBandPad = 32;
int64_t band_count(int64_t ncols);
// hist[j] (32 int64) = number of B rows with at least 2^j entries.
void launch_len_hist(const int64_t* Bptr, int64_t nrows, int64_t* hist, hipStream_t s);
// slot[k] = index of B row k among the rows of at least minlen entries (-1 for the others),
// slot_row[s] = the row of slot s. flag (nrows + 1 int64) and tmp (scan_tmp_elems(nrows + 1)) are
// scratch; flag[nrows] receives the number of slots.
void launch_heavy_slots(const int64_t* Bptr, int64_t nrows, int64_t minlen, int64_t* flag, int64_t* tmp,
                        int32_t* slot, int32_t* slot_row, hipStream_t s);
// band[s * nb1 + b] = position (from the row start) of slot s's first entry with column >= b * 4096;
// band[s * nb1 + nb1 - 1] = the row's length.
void launch_band_index(const int32_t* slot_row, int64_t nslots, const int64_t* Bptr, const int32_t* Bidx, int64_t nb1,
                       int32_t* band, hipStream_t s);

// ---- planning ----
// P[i] = sum_{k in A_i} len(B_k); nnz[i] = 0 where P[i] == 0; per A nonzero j with column k:
// bstart[j] = B.ptr[k], blen[j] = len(B_k) (both nullable).
void launch_row_products(const Mat& A, const Mat& B, int64_t* P, int64_t* nnz, int64_t* bstart, int32_t* blen,
                         hipStream_t s);
// Stable counting sort of rows [0, n) by category of min(key[i], cap).
// counts/offs have ncat * nwg (+1) int64 where nwg = bin_tiles(n).
int64_t bin_tiles(int64_t n);
// With hub_slot, rows of the category before the last (the long one by key) that have a hub slot
// (>= 0) go to the last category (the hub category).
void launch_bin_count(const int64_t* key, int64_t n, int64_t cap, const CatTable& ct, int64_t* counts,
                      hipStream_t s, const int32_t* hub_slot = nullptr);
void launch_bin_scatter(const int64_t* key, int64_t n, int64_t cap, const CatTable& ct, const int64_t* offs,
                        int32_t* rows, hipStream_t s, const int32_t* hub_slot = nullptr);

// ---- symbolic / numeric / backward per category ----
// grid == 0 -> exact grid from nrows_hint; else a persistent grid of that many workgroups.
int symbolic_groups_per_wg(const KernelShape& k);
// With hub_slot, the long-row kernel gives rows of slot hs >= 0 windows aligned to whole hub ranges
// and writes their columns per range into rng_cnt[hs * nrng + range] (ranges with no column are
// left as they are: zero them first).
void launch_symbolic(const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                     int64_t nrows_hint, int grid, const Mat& A, const Mat& B, int64_t* nnz, int64_t* cursor,
                     int* err, hipStream_t s, const int32_t* hub_slot = nullptr, int32_t* rng_cnt = nullptr,
                     int nrng = 0);
// distinct: most products of the problem are distinct columns of their row (nnzC close to the
// product count); the one-wave shapes then take rows whose products are all distinct through a
// register path instead of the hash table.
void launch_numeric(int dtype, const KernelShape& k, const int32_t* rows, const int64_t* pbeg,
                    const int64_t* pend, int64_t nrows_hint, int grid, const Mat& A, const Mat& B,
                    const int64_t* Cptr, int64_t cbase, int32_t* Cidx, void* Cval, int64_t* cursor, int* err,
                    hipStream_t s, bool distinct = false);
// Hub rows rows[0, nh) in column ranges of 16384 (hub_ranges(ncols) of them; B's band index
// required). prepare: per (row, range) count of the short entries' products into scnt[nh *
// nrng], each row's banded entries (offsets in the row) into hlist at the row's A positions and
// their number into hcount. scatter: those products, (column, value), into buckets at soff (the
// exclusive scan of scnt). range(count): distinct columns per (row, range) into rcnt;
// range(!count): values, emitted at Cptr[row] - cbase + roff[item] - roff[row's first item].
constexpr int kHubMaxRanges = 4096;
int hub_ranges(int64_t ncols);
// Hub slots: flag[j] = 1 for rows[j] with more than kHubEntries A entries; with pos the exclusive
// scan of flag, hub_slot[rows[j]] = pos[j] for the flagged rows (hub_slot filled with -1 first).
void launch_hub_flags(const int32_t* rows, int64_t n, const int64_t* Aptr, int64_t* flag, hipStream_t s);
void launch_hub_slots(const int32_t* rows, int64_t n, const int64_t* flag, const int64_t* pos, int32_t* hub_slot,
                      hipStream_t s);
// rcnt[j * nrng + r] = rng_cnt[hub_slot[rows[j]] * nrng + r] (the symbolic pass's range counts).
void launch_hub_gather_counts(const int32_t* rows, int64_t n, int nrng, const int32_t* hub_slot,
                              const int32_t* rng_cnt, int64_t* rcnt, hipStream_t s);
void launch_hub_prepare(const int32_t* rows, int64_t nh, const Mat& A, const Mat& B, int nrng, int64_t* scnt,
                        int32_t* hlist, int32_t* hcount, hipStream_t s);
void launch_hub_scatter(int dtype, const int32_t* rows, int64_t nh, const Mat& A, const Mat& B, int nrng,
                        const int64_t* soff, int32_t* scol, void* sval, const int* err, hipStream_t s);
// Sets the capacity bit in err when *total (the hub rows' short products) exceeds cap: the hub
// scatter and range kernels then write nothing.
void launch_hub_guard(const int64_t* total, int64_t cap, int* err, hipStream_t s);
void launch_hub_range(int dtype, const int32_t* rows, int64_t nh, int nrng, const Mat& A, const Mat& B,
                      const int32_t* hlist, const int32_t* hcount, const int64_t* soff, const int32_t* scol,
                      const void* sval, const int64_t* roff, const int64_t* Cptr, int64_t cbase, int32_t* Cidx,
                      void* Cval, int* err, hipStream_t s);
void launch_backward(int dtype, const KernelShape& k, const int32_t* rows, const int64_t* pbeg,
                     const int64_t* pend, int64_t nrows_hint, int grid, const Mat& A, const Mat& B,
                     const int64_t* Cptr, const int32_t* Cidx, const void* dC, void* dA, void* dB, int64_t* cursor,
                     int* err, hipStream_t s);
// Is this (T, NT, G) compiled? (the autotuner only proposes compiled shapes)
bool shape_available(int phase, const KernelShape& k);  // phase 0 symbolic, 1 numeric, 2 backward

}  // namespace dbd
