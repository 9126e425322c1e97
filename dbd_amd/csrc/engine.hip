// The DBD engine on one MI355X: planner, category dispatch, chunked output pipeline with
// D2H into pinned host memory and parallel file writers, device two-phase API, backward,
// batched hipGraph form and the autotuner. All exported symbols are declared in
// include/dbd.h.
//
// Reference pipeline being replaced (Z/ = dbd/src of DBD_FULL.zip, S/ = for_small_datasets):
//   divide_Blocks + get_UpperBound   Z/main.cpp:259-310, S/preProc.cu:142-156   -> plan()
//   distribute_Workload dispatcher    Z/distributeWorkload.cpp:25-223           -> run()
//   compute_GPU per block             S/computeGPU.cu:15-98                     -> numeric chunks
//   write_File double buffers         Z/writeFile.cpp:13-154                    -> Writer
// The reference dispatches one 512/4096-row block at a time to a GPU thread through
// usleep-polled flags and synchronous cudaMemcpy. Here blocks are the unit of planning and
// of output streaming, while kernels cover whole chunks of blocks per category, D2H runs on
// its own stream, and writers pwrite chunks at their final file offsets.
#include <fcntl.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "pipeline.h"

using namespace dbd;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  dbd_status ensure(size_t n) {
    if (n <= bytes && p) return DBD_OK;
    release();
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      bytes = 0;
      (void)hipGetLastError();
      return set_error(DBD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return DBD_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  dbd_status ensure(size_t n) {
    if (n <= bytes && p) return DBD_OK;
    release();
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      bytes = 0;
      (void)hipGetLastError();
      return set_error(DBD_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return DBD_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

}  // namespace

// ---- category tables -------------------------------------------------------------------
// Default table. Symbolic key: min(P_i, ncols) (a bound on nnz_i); numeric key: nnz_i.
// Hash tables are sized so the load stays <= `load`; the one-wave symbolic tables (keys only,
// 4 bytes a slot, up to 1024 slots) to half of that, which shortens their probe sequences.
Tables dbd::default_tables(int g_small, int g_big, double load, int top, int64_t hub_min_cols) {
  Tables t;
  t.g_small = g_small;
  t.g_big = g_big;
  t.load = load;
  t.top = top;
  t.hub_min_cols = hub_min_cols;
  const int Ts[] = {64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384};
  const int NTs[] = {64, 64, 64, 64, 128, 256, 512, 1024, 1024};
  int n = 0;
  t.sym_ct.lim[n] = 0;
  t.sym_k[n++] = {0, 0, 0};
  for (int i = 0; i < 9 && Ts[i] <= 2 * top; ++i) {
    t.sym_ct.lim[n] = (int64_t)(Ts[i] * (Ts[i] <= 1024 ? load * 0.5 : load));
    t.sym_k[n++] = {Ts[i], Ts[i] <= 1024 ? 64 : NTs[i], Ts[i] <= 1024 ? g_small : g_big};
  }
  t.sym_ct.lim[n] = INT64_MAX;
  t.sym_k[n++] = {0, 1024, 16};  // long rows
  t.sym_ct.n = n;
  n = 0;
  t.num_ct.lim[n] = 0;
  t.num_k[n++] = {0, 0, 0};
  for (int i = 0; i < 8 && Ts[i] <= top; ++i) {
    t.num_ct.lim[n] = (int64_t)(Ts[i] * load);
    t.num_k[n++] = {Ts[i], NTs[i], Ts[i] <= 1024 ? g_small : g_big};
  }
  t.num_ct.lim[n] = INT64_MAX;
  t.num_k[n++] = {0, 256, 16};  // long rows
  t.num_ct.lim[n] = INT64_MAX;  // hub rows: long rows of more than kHubEntries A entries
  t.num_k[n++] = {kHubShape, 1024, 0};
  t.num_ct.n = n;
  return t;
}

namespace {

// The autotuner's search space: every combination of five knobs, each listed default first, so
// table 0 is the default table and id = mixed-radix digits (g_small, g_big, load, top, hub).
constexpr int kKnobG[2] = {16, 64};
constexpr int kKnobGBig[2] = {64, 16};
constexpr double kKnobLoad[2] = {0.5, 0.25};
constexpr int kKnobTop[2] = {8192, 4096};
constexpr int64_t kKnobHub[3] = {0, kHubWideCols, INT64_MAX};  // always, from 2^21 columns, never
constexpr int kKnobRadix[5] = {2, 2, 2, 2, 3};

std::vector<Tables>& candidate_tables() {
  static std::vector<Tables> c = [] {
    std::vector<Tables> v;
    for (int id = 0; id < 2 * 2 * 2 * 2 * 3; ++id) {
      int d[5], r = id;
      for (int k = 4; k >= 0; --k) {
        d[k] = r % kKnobRadix[k];
        r /= kKnobRadix[k];
      }
      Tables t = default_tables(kKnobG[d[0]], kKnobGBig[d[1]], kKnobLoad[d[2]], kKnobTop[d[3]], kKnobHub[d[4]]);
      t.id = id;
      v.push_back(t);
    }
    return v;
  }();
  return c;
}

// id of the table whose knob k is set to value index x, the other knobs as in table `id`
int table_with_knob(int id, int k, int x) {
  int d[5], r = id;
  for (int j = 4; j >= 0; --j) {
    d[j] = r % kKnobRadix[j];
    r /= kKnobRadix[j];
  }
  d[k] = x;
  int out = 0;
  for (int j = 0; j < 5; ++j) out = out * kKnobRadix[j] + d[j];
  return out;
}

// autotuner cache: problem class -> table id
std::mutex g_tune_mu;
std::map<std::string, int> g_tune_cache;

// ---- writer pool for DBD_MODE_IO ---------------------------------------------------------
struct WriteJob {
  int slot;          // pinned ring slot, or -1 for a host buffer of the hybrid mode's CPU blocks
  hipEvent_t ready;  // the slot's copy is done (nullptr: nothing to wait for)
  const void* idx;
  const void* val;
  int64_t n;
  int64_t off;                // element offset of the chunk in C
  std::shared_ptr<void> keep;  // owner of a host buffer, released when both files have it
};

// File writers of DBD_MODE_IO: one thread per output file (Idx.output, Val.output), each
// writing the chunks in order at their final offsets, so every file is written front to back
// by a single writer (the filesystem serialises writes to one file anyway). A pinned ring
// slot is free again when both files have written it.
struct Writer {
  int fd_idx = -1, fd_val = -1;
  size_t vsz = 8;
  std::mutex mu;
  std::condition_variable cv_job, cv_free;
  std::deque<WriteJob> q[2];
  std::vector<int> slot_busy;  // file writes still pending per slot (0, 1 or 2)
  std::vector<std::thread> th;
  bool stop = false;
  std::atomic<bool> failed{false};
  std::string why;
  double busy_s = 0;
  double bytes = 0;

  void start(int nslots) {
    slot_busy.assign(nslots, 0);
    stop = false;
    failed = false;
    for (int f = 0; f < 2; ++f) th.emplace_back([this, f] { loop(f); });
  }
  void loop(int f) {
    for (;;) {
      WriteJob j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_job.wait(lk, [&] { return stop || !q[f].empty(); });
        if (q[f].empty()) return;
        j = q[f].front();
        q[f].pop_front();
      }
      double t0 = now_s();
      bool ok = j.ready == nullptr || hipEventSynchronize(j.ready) == hipSuccess;
      const size_t es = f == 0 ? 4 : vsz;
      ok = ok && pw(f == 0 ? fd_idx : fd_val, f == 0 ? j.idx : j.val, es * (size_t)j.n, (off_t)(es * (size_t)j.off));
      double t1 = now_s();
      {
        std::lock_guard<std::mutex> lk(mu);
        busy_s += t1 - t0;
        bytes += (double)es * (double)j.n;
        if (!ok && !failed) {
          failed = true;
          why = strerror(errno);
        }
        if (j.slot >= 0) --slot_busy[j.slot];
      }
      j.keep.reset();
      cv_free.notify_all();
    }
  }
  static bool pw(int fd, const void* buf, size_t n, off_t off) {
    const char* p = (const char*)buf;
    while (n > 0) {
      ssize_t w = pwrite(fd, p, n > ((size_t)1 << 30) ? ((size_t)1 << 30) : n, off);
      if (w <= 0) {
        if (w < 0 && errno == EINTR) continue;
        return false;
      }
      p += w;
      n -= (size_t)w;
      off += w;
    }
    return true;
  }
  void acquire(int slot) {
    std::unique_lock<std::mutex> lk(mu);
    cv_free.wait(lk, [&] { return slot_busy[slot] == 0 || failed; });
    slot_busy[slot] = 2;
  }
  void push(const WriteJob& j) {
    {
      std::lock_guard<std::mutex> lk(mu);
      q[0].push_back(j);
      q[1].push_back(j);
    }
    cv_job.notify_all();
  }
  void finish() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_job.notify_all();
    for (auto& t : th) t.join();
    th.clear();
  }
  // Early returns (a failed HIP call mid-run) must not leave joinable threads or open files.
  ~Writer() {
    if (!th.empty()) finish();
    if (fd_idx >= 0) close(fd_idx);
    if (fd_val >= 0) close(fd_val);
  }
};

}  // namespace

struct dbd_ctx {
  dbd_opts o{};
  hipStream_t st = nullptr;
  bool own_st = false;
  hipStream_t cst = nullptr;  // copy stream
  int dev = 0;
  int num_cu = 256;
  size_t vsz = 8;

  // operands
  Mat A{}, B{};
  bool ops = false, same = false;
  DevBuf upA[3], upB[3];
  int64_t a_bytes = 0, b_bytes = 0;
  double h2d_s = 0;
  dbd_csr hostA{}, hostB{};  // the caller's host views (the hybrid mode's CPU blocks read them)
  bool host_ops = false;

  // plan state
  Tables tb;
  bool tuned_set = false;       // a table was chosen for the current operands (tuned or explicit)
  bool table_explicit = false;  // dbd_set_table was called: opts.tune does not override it
  bool planned = false, symb = false;
  int64_t nr = 0, nb = 0, nwg = 0;
  DevBuf P, Pptr, nnz, Cptr, rows_sym, rows_num, cnt, offs_sym, offs_num, tmp, cursor, blk, err, bdesc_s, bdesc_l;
  // band index of B's long rows (for the long-row window kernels), built once per operand set
  DevBuf band_hist, band_flag, band_tmp, band_slot, band_row, band;
  bool band_ready = false;
  // workspace of the hub-row range kernels: banded-entry lists (per A nonzero), per-row counts,
  // per-(row, range) short-product counts and column counts with their scans, short products
  DevBuf hub_list, hub_cnt, hub_scnt, hub_rcnt, hub_tmp, hub_scol, hub_sval;
  // hub slots of the symbolic long rows of more than kHubEntries entries (per row, -1: none) and
  // their column counts per hub range, written by the symbolic long kernel
  DevBuf hub_slot, hub_flag, hub_rng;
  int64_t hub_n = 0;
  int64_t band_minlen = 0, band_rows = 0, band_nb1 = 0;
  std::vector<int64_t> h_blk, h_blockP, h_blockC, h_offs_num, h_sym_tot;
  dbd_plan_info info{};

  // outputs
  DevBuf out_idx[2], out_val[2];
  HostBuf h_cptr, h_idx, h_val;      // whole host result (D2H retain)
  std::vector<HostBuf> ring_idx, ring_val;
  bool dev_result = false, host_result = false;
  hipEvent_t ev_comp[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> ev_slot;
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr, ev_c0 = nullptr, ev_c1 = nullptr;

  // batch graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_key[3] = {nullptr, nullptr, nullptr};
  int64_t batch_cap = 0;  // entries of C the batch caller allocated (exact nnz at prepare)
  // the batch's hub rows: position in the numeric bins, count, bucket capacity (from prepare)
  int64_t bh_pos0 = 0, bh_n = 0, bh_nshort = 0;
  DevBuf g_rng;
};

namespace {

dbd_status ensure_events(dbd_ctx* c) {
  for (int i = 0; i < 2; ++i) {
    if (!c->ev_comp[i]) HIPCHK(hipEventCreateWithFlags(&c->ev_comp[i], hipEventDisableTiming));
    if (!c->ev_copy[i]) HIPCHK(hipEventCreateWithFlags(&c->ev_copy[i], hipEventDisableTiming));
  }
  if (!c->ev_t0) HIPCHK(hipEventCreate(&c->ev_t0));
  if (!c->ev_t1) HIPCHK(hipEventCreate(&c->ev_t1));
  if (!c->ev_c0) HIPCHK(hipEventCreate(&c->ev_c0));
  if (!c->ev_c1) HIPCHK(hipEventCreate(&c->ev_c1));
  return DBD_OK;
}

dbd_status check_err(dbd_ctx* c, const char* phase) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->err.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (e != 0) return set_error(DBD_ERR_INTERNAL, "%s: device consistency check failed (flags 0x%x)", phase, e);
  return DBD_OK;
}

// Whether hub rows can take the range kernels: they read B's band index for long B rows and
// keep one counter per column range in LDS.
bool hub_ranges_ok(const dbd_ctx* c) { return c->B.band != nullptr && hub_ranges(c->B.ncols) <= kHubMaxRanges; }

// Hub slots (after the symbolic binning and the band index): the symbolic long rows of more than
// kHubEntries A entries, when the table sends hub rows to the range kernels for this width. The
// symbolic long kernel then records their column counts per range, and the numeric binning puts
// the ones that stay long into the hub category.
dbd_status assign_hub_slots(dbd_ctx* c) {
  c->hub_n = 0;
  const CatTable& ct = c->tb.sym_ct;
  const int64_t p0 = c->h_sym_tot[(size_t)ct.n - 1], nlong = c->h_sym_tot[(size_t)ct.n] - p0;
  if (nlong <= 0 || !hub_enabled(c->tb, c->B.ncols) || !hub_ranges_ok(c)) return DBD_OK;
  hipStream_t s = c->st;
  const int nrng = hub_ranges(c->B.ncols);
  const int32_t* rows = c->rows_sym.as<int32_t>() + p0;
  CHK(c->hub_slot.ensure(4 * (size_t)c->nr));
  CHK(c->hub_flag.ensure(16 * (size_t)(nlong + 1)));
  CHK(c->hub_tmp.ensure(8 * (size_t)scan_tmp_elems(nlong + 1)));
  int64_t* flag = c->hub_flag.as<int64_t>();
  int64_t* pos = flag + nlong + 1;
  launch_fill_i32(c->hub_slot.as<int32_t>(), c->nr, -1, s);
  launch_hub_flags(rows, nlong, c->A.ptr, flag, s);
  launch_exclusive_scan(flag, pos, nlong, c->hub_tmp.as<int64_t>(), s);
  int64_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, pos + nlong, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (n <= 0 || n * (int64_t)nrng >= INT32_MAX) return DBD_OK;  // none, or too many to count per range
  launch_hub_slots(rows, nlong, flag, pos, c->hub_slot.as<int32_t>(), s);
  CHK(c->hub_rng.ensure(4 * (size_t)(n * nrng)));
  launch_fill_i32(c->hub_rng.as<int32_t>(), n * nrng, 0, s);
  c->hub_n = n;
  return DBD_OK;
}

// Hub rows rows_num[pos0, pos0 + n) of a chunk in column ranges: banded-entry lists and bucketed
// short products, the symbolic pass's (row, range) column counts scanned into output offsets, the
// numeric pass. hub_reserve sizes the workspace for n rows; hub_head lists the banded entries and
// scans the short products' counts per (row, range); hub_tail fills the buckets (at most
// `short_cap` products, else the capacity bit is set and nothing is written) and runs the ranges.
dbd_status hub_reserve(dbd_ctx* c, int64_t n) {
  const int64_t items = n * hub_ranges(c->B.ncols);
  CHK(c->hub_list.ensure(4 * (size_t)std::max<int64_t>(c->A.nnz, 1)));
  CHK(c->hub_cnt.ensure(4 * (size_t)n));
  CHK(c->hub_scnt.ensure(16 * (size_t)(items + 1)));  // counts, then their scan
  CHK(c->hub_rcnt.ensure(16 * (size_t)(items + 1)));
  CHK(c->hub_tmp.ensure(8 * (size_t)scan_tmp_elems(items + 1)));
  return DBD_OK;
}

// The number of short products lands at hub_scnt[2 * items + 1] (the end of their scan).
void hub_head(dbd_ctx* c, int64_t pos0, int64_t n, hipStream_t s) {
  const int nrng = hub_ranges(c->B.ncols);
  const int64_t items = n * nrng;
  int64_t* scnt = c->hub_scnt.as<int64_t>();
  launch_hub_prepare(c->rows_num.as<int32_t>() + pos0, n, c->A, c->B, nrng, scnt, c->hub_list.as<int32_t>(),
                     c->hub_cnt.as<int32_t>(), s);
  launch_exclusive_scan(scnt, scnt + items + 1, items, c->hub_tmp.as<int64_t>(), s);
}

void hub_tail(dbd_ctx* c, int64_t pos0, int64_t n, int64_t short_cap, const int64_t* Cptr, int64_t cbase,
              int32_t* Cidx, void* Cval, hipStream_t s) {
  const int nrng = hub_ranges(c->B.ncols);
  const int64_t items = n * nrng;
  const int32_t* rows = c->rows_num.as<int32_t>() + pos0;
  const int64_t* soff = c->hub_scnt.as<int64_t>() + items + 1;
  int64_t* rcnt = c->hub_rcnt.as<int64_t>();
  int64_t* roff = rcnt + items + 1;
  int* err = c->err.as<int>();
  launch_hub_guard(soff + items, short_cap, err, s);
  launch_hub_scatter(c->o.dtype, rows, n, c->A, c->B, nrng, soff, c->hub_scol.as<int32_t>(), c->hub_sval.p, err, s);
  launch_hub_gather_counts(rows, n, nrng, c->hub_slot.as<int32_t>(), c->hub_rng.as<int32_t>(), rcnt, s);
  launch_exclusive_scan(rcnt, roff, items, c->hub_tmp.as<int64_t>(), s);
  launch_hub_range(c->o.dtype, rows, n, nrng, c->A, c->B, c->hub_list.as<int32_t>(), c->hub_cnt.as<int32_t>(), soff,
                   c->hub_scol.as<int32_t>(), c->hub_sval.p, roff, Cptr, cbase, Cidx, Cval, err, s);
}

// Sizes the buckets: one host synchronisation for the number of short products.
dbd_status hub_buckets(dbd_ctx* c, int64_t n, int64_t* nshort) {
  const int64_t items = n * hub_ranges(c->B.ncols);
  *nshort = 0;
  HIPCHK(hipMemcpyAsync(nshort, c->hub_scnt.as<int64_t>() + 2 * items + 1, 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  CHK(c->hub_scol.ensure(4 * (size_t)std::max<int64_t>(*nshort, 1)));
  CHK(c->hub_sval.ensure(c->vsz * (size_t)std::max<int64_t>(*nshort, 1)));
  return DBD_OK;
}

dbd_status hub_numeric(dbd_ctx* c, int64_t pos0, int64_t n, const int64_t* Cptr, int64_t cbase, int32_t* Cidx,
                       void* Cval) {
  CHK(hub_reserve(c, n));
  hub_head(c, pos0, n, c->st);
  int64_t nshort = 0;
  CHK(hub_buckets(c, n, &nshort));
  hub_tail(c, pos0, n, nshort, Cptr, cbase, Cidx, Cval, c->st);
  return DBD_OK;
}

// Band index of B's long rows: the rows of at least `minlen` entries, minlen the smallest power
// of two from 64 up whose rows' index fits min(8 GiB, device budget / 4). Built once per
// operand set (B does not change between runs).
dbd_status build_band(dbd_ctx* c) {
  hipStream_t s = c->st;
  const int64_t nB = c->B.nrows;
  c->band_ready = true;
  c->band_rows = 0;
  if (nB <= 0) return DBD_OK;
  CHK(c->band_hist.ensure(8 * 32));
  launch_len_hist(c->B.ptr, nB, c->band_hist.as<int64_t>(), s);
  int64_t hist[32];
  HIPCHK(hipMemcpyAsync(hist, c->band_hist.p, sizeof(hist), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int64_t nb1 = band_count(c->B.ncols);
  const int64_t cap = std::min<int64_t>((int64_t)8 << 30, (c->o.device_budget_mib << 20) / 4);
  // from 64 up: shorter B rows are read whole by the span walks (at most 64 elements a window),
  // and leaving them out of the index saves the hub kernels their per-range band reads
  int j = 6;
  while (j < 31 && hist[j] * nb1 * 4 > cap) ++j;
  if (hist[j] == 0 || hist[j] * nb1 * 4 > cap) return DBD_OK;  // nothing to band: cursor walks only
  const int64_t H = hist[j];
  CHK(c->band_flag.ensure(8 * (size_t)(nB + 1)));
  CHK(c->band_tmp.ensure(8 * (size_t)scan_tmp_elems(nB + 1)));
  CHK(c->band_slot.ensure(4 * (size_t)nB));
  CHK(c->band_row.ensure(4 * (size_t)H));
  CHK(c->band.ensure(4 * (size_t)(H * nb1)));
  launch_heavy_slots(c->B.ptr, nB, (int64_t)1 << j, c->band_flag.as<int64_t>(), c->band_tmp.as<int64_t>(),
                     c->band_slot.as<int32_t>(), c->band_row.as<int32_t>(), s);
  launch_band_index(c->band_row.as<int32_t>(), H, c->B.ptr, c->B.idx, nb1, c->band.as<int32_t>(), s);
  HIPCHK(hipGetLastError());
  c->band_minlen = (int64_t)1 << j;
  c->band_rows = H;
  c->band_nb1 = nb1;
  return DBD_OK;
}

// B as the kernels see it: with its band index when one was built
void attach_band(dbd_ctx* c) {
  if (c->band_rows > 0) {
    c->B.hslot = c->band_slot.as<int32_t>();
    c->B.band = c->band.as<int32_t>();
    c->B.nb1 = c->band_nb1;
  } else {
    c->B.hslot = nullptr;
    c->B.band = nullptr;
    c->B.nb1 = 0;
  }
}

// Row analysis, int64 prefix, blocks, reference criteria, symbolic bins.
dbd_status plan(dbd_ctx* c) {
  if (!c->ops) return set_error(DBD_ERR_STATE, "dbd_plan: operands not set");
  const int64_t nr = c->A.nrows;
  c->nr = nr;
  c->nwg = bin_tiles(nr);
  const int BR = c->o.block_rows;
  c->nb = nr == 0 ? 0 : (nr - 1) / BR + 1;
  hipStream_t s = c->st;
  CHK(c->P.ensure(8 * (size_t)(nr + 1)));
  CHK(c->Pptr.ensure(8 * (size_t)(nr + 1)));
  CHK(c->nnz.ensure(8 * (size_t)(nr + 1)));
  CHK(c->Cptr.ensure(8 * (size_t)(nr + 1)));
  CHK(c->rows_sym.ensure(4 * (size_t)(nr + 1)));
  CHK(c->rows_num.ensure(4 * (size_t)(nr + 1)));
  const size_t ncnt = (size_t)kMaxCats * (size_t)c->nwg + 1;
  CHK(c->cnt.ensure(8 * ncnt));
  CHK(c->offs_sym.ensure(8 * (ncnt + 1)));
  CHK(c->offs_num.ensure(8 * (ncnt + 1)));
  CHK(c->tmp.ensure(8 * (size_t)scan_tmp_elems((int64_t)ncnt + nr + 1)));
  CHK(c->blk.ensure(8 * (size_t)(std::max(c->nb, (int64_t)kMaxCats) + 2)));
  CHK(c->err.ensure(64));  // error word, then a sink the numeric kernels write degenerate rows to
  CHK(c->bdesc_s.ensure(8 * (size_t)(c->A.nnz + 1)));
  CHK(c->bdesc_l.ensure(4 * (size_t)(c->A.nnz + 1)));
  c->A.bstart = c->bdesc_s.as<int64_t>();
  c->A.blen = c->bdesc_l.as<int32_t>();
  if (c->same) c->B = c->A;
  launch_fill_i32(c->err.as<int32_t>(), 1, 0, s);

  launch_row_products(c->A, c->B, c->P.as<int64_t>(), c->nnz.as<int64_t>(), (int64_t*)c->A.bstart,
                      (int32_t*)c->A.blen, s);
  launch_exclusive_scan(c->P.as<int64_t>(), c->Pptr.as<int64_t>(), nr, c->tmp.as<int64_t>(), s);
  launch_gather_strided(c->Pptr.as<int64_t>(), nr, BR, c->nb, c->blk.as<int64_t>(), s);
  c->h_blk.assign((size_t)c->nb + 1, 0);
  HIPCHK(hipMemcpyAsync(c->h_blk.data(), c->blk.p, 8 * (size_t)(c->nb + 1), hipMemcpyDeviceToHost, s));

  // symbolic bins
  const CatTable& ct = c->tb.sym_ct;
  const int64_t n_off = (int64_t)ct.n * c->nwg;
  launch_bin_count(c->P.as<int64_t>(), nr, c->B.ncols, ct, c->cnt.as<int64_t>(), s);
  launch_exclusive_scan(c->cnt.as<int64_t>(), c->offs_sym.as<int64_t>(), n_off, c->tmp.as<int64_t>(), s);
  launch_bin_scatter(c->P.as<int64_t>(), nr, c->B.ncols, ct, c->offs_sym.as<int64_t>(), c->rows_sym.as<int32_t>(), s);
  DevBuf& g = c->g_rng;
  CHK(g.ensure(8 * (kMaxCats + 2)));
  launch_gather_strided(c->offs_sym.as<int64_t>(), n_off, c->nwg, ct.n, g.as<int64_t>(), s);
  c->h_sym_tot.assign((size_t)ct.n + 1, 0);
  HIPCHK(hipMemcpyAsync(c->h_sym_tot.data(), g.p, 8 * (size_t)(ct.n + 1), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));

  // block table + reference criteria
  c->h_blockP.assign((size_t)c->nb, 0);
  for (int64_t b = 0; b < c->nb; ++b) c->h_blockP[(size_t)b] = c->h_blk[(size_t)b + 1] - c->h_blk[(size_t)b];
  dbd_plan_info& in = c->info;
  memset(&in, 0, sizeof(in));
  block_criteria(c->h_blockP.data(), c->nb, BR, &in);
  in.nrows = nr;
  in.ncols = c->B.ncols;
  in.nnz = -1;
  in.n_sym_cats = ct.n;
  in.n_num_cats = c->tb.num_ct.n;
  for (int i = 0; i < ct.n; ++i) {
    in.sym_rows[i] = c->h_sym_tot[(size_t)i + 1] - c->h_sym_tot[(size_t)i];
    in.sym_limit[i] = ct.lim[i];
  }
  for (int i = 0; i < c->tb.num_ct.n; ++i) {
    in.num_rows[i] = -1;
    in.num_limit[i] = c->tb.num_ct.lim[i];
  }
  // long rows (the last symbolic category) read B's band index
  if (c->h_sym_tot[(size_t)ct.n] - c->h_sym_tot[(size_t)ct.n - 1] > 0 && !c->band_ready) CHK(build_band(c));
  attach_band(c);
  CHK(assign_hub_slots(c));
  c->planned = true;
  c->symb = false;
  return DBD_OK;
}

// window-kernel workspace: 16 bytes per A nonzero (cursor state of long rows)
dbd_status ensure_cursor(dbd_ctx* c) { return c->cursor.ensure(16 * (size_t)(c->A.nnz + 1)); }

// Symbolic pass over all rows -> nnz_i -> Cptr (device, ctx-owned) -> numeric bins.
dbd_status symbolic(dbd_ctx* c) {
  if (!c->planned) return set_error(DBD_ERR_STATE, "symbolic before plan");
  hipStream_t s = c->st;
  const CatTable& ct = c->tb.sym_ct;
  const int64_t* offs = c->offs_sym.as<int64_t>();
  for (int k = 1; k < ct.n; ++k) {
    const int64_t n = c->h_sym_tot[(size_t)k + 1] - c->h_sym_tot[(size_t)k];
    if (n <= 0) continue;
    if (c->tb.sym_k[k].T == 0) CHK(ensure_cursor(c));
    const bool hubs = c->hub_n > 0;
    launch_symbolic(c->tb.sym_k[k], c->rows_sym.as<int32_t>(), offs + (int64_t)k * c->nwg,
                    offs + (int64_t)(k + 1) * c->nwg, n, 0, c->A, c->B, c->nnz.as<int64_t>(),
                    c->cursor.as<int64_t>(), c->err.as<int>(), s, hubs ? c->hub_slot.as<int32_t>() : nullptr,
                    hubs ? c->hub_rng.as<int32_t>() : nullptr, hubs ? hub_ranges(c->B.ncols) : 0);
  }
  launch_exclusive_scan(c->nnz.as<int64_t>(), c->Cptr.as<int64_t>(), c->nr, c->tmp.as<int64_t>(), s);
  const int BR = c->o.block_rows;
  launch_gather_strided(c->Cptr.as<int64_t>(), c->nr, BR, c->nb, c->blk.as<int64_t>(), s);
  std::vector<int64_t> hb((size_t)c->nb + 1, 0);
  HIPCHK(hipMemcpyAsync(hb.data(), c->blk.p, 8 * (size_t)(c->nb + 1), hipMemcpyDeviceToHost, s));
  // numeric bins by exact nnz_i
  const CatTable& nt = c->tb.num_ct;
  const int64_t n_off = (int64_t)nt.n * c->nwg;
  const int32_t* hub_rows = c->hub_n > 0 ? c->hub_slot.as<int32_t>() : nullptr;
  launch_bin_count(c->nnz.as<int64_t>(), c->nr, INT64_MAX, nt, c->cnt.as<int64_t>(), s, hub_rows);
  launch_exclusive_scan(c->cnt.as<int64_t>(), c->offs_num.as<int64_t>(), n_off, c->tmp.as<int64_t>(), s);
  launch_bin_scatter(c->nnz.as<int64_t>(), c->nr, INT64_MAX, nt, c->offs_num.as<int64_t>(), c->rows_num.as<int32_t>(),
                     s, hub_rows);
  c->h_offs_num.assign((size_t)n_off + 1, 0);
  HIPCHK(hipMemcpyAsync(c->h_offs_num.data(), c->offs_num.p, 8 * (size_t)(n_off + 1), hipMemcpyDeviceToHost, s));
  CHK(check_err(c, "symbolic"));  // synchronises the stream
  c->h_blockC.assign((size_t)c->nb, 0);
  for (int64_t b = 0; b < c->nb; ++b) c->h_blockC[(size_t)b] = hb[(size_t)b + 1] - hb[(size_t)b];
  c->info.nnz = c->nb > 0 ? hb[(size_t)c->nb] : 0;
  for (int k = 0; k < nt.n; ++k)
    c->info.num_rows[k] = c->h_offs_num[(size_t)(k + 1) * c->nwg] - c->h_offs_num[(size_t)k * c->nwg];
  c->symb = true;
  return DBD_OK;
}

// Numeric kernels for rows of blocks [b0, b1) into Cidx/Cval at Cptr - cbase.
dbd_status numeric_blocks(dbd_ctx* c, int64_t b0, int64_t b1, const int64_t* Cptr, int64_t cbase, int32_t* Cidx,
                          void* Cval) {
  const int BR = c->o.block_rows;
  const int64_t r0 = b0 * BR, r1 = std::min(c->nr, b1 * BR);
  const int64_t w0 = r0 / 256, w1 = (r1 + 255) / 256;  // bin tiles (BR is a multiple of 256)
  const CatTable& nt = c->tb.num_ct;
  const int64_t* offs = c->offs_num.as<int64_t>();
  // few duplicate products overall (nnzC >= 90% of the products): rows of distinct products
  // take the numeric kernels' register path
  const bool distinct = c->info.nnz >= 0 && c->info.nnz >= c->info.products - c->info.products / 10;
  for (int k = 1; k < nt.n; ++k) {
    const int64_t i0 = (int64_t)k * c->nwg + w0, i1 = (int64_t)k * c->nwg + w1;
    const int64_t n = c->h_offs_num[(size_t)i1] - c->h_offs_num[(size_t)i0];
    if (n <= 0) continue;
    if (c->tb.num_k[k].T == kHubShape) {  // rows with hub slots (hub_n > 0): the range kernels apply
      CHK(hub_numeric(c, c->h_offs_num[(size_t)i0], n, Cptr, cbase, Cidx, Cval));
      continue;
    }
    if (c->tb.num_k[k].T == 0 || c->tb.num_k[k].T == kHubShape) CHK(ensure_cursor(c));
    launch_numeric(c->o.dtype, c->tb.num_k[k], c->rows_num.as<int32_t>(), offs + i0, offs + i1, n, 0, c->A, c->B,
                   Cptr, cbase, Cidx, Cval, c->cursor.as<int64_t>(), c->err.as<int>(), c->st, distinct);
  }
  HIPCHK(hipGetLastError());  // a launch the runtime refused
  return DBD_OK;
}

struct Chunk {
  int64_t b0, b1, e0, e1;  // blocks and element range of C
};

// Group consecutive blocks so each chunk's output stays within `cap` elements.
// Chunks: runs of whole consecutive blocks whose C fits `cap` elements. Blocks marked in `cpu`
// (the hybrid mode's host blocks) are left out: chunks end before them and resume after.
dbd_status make_chunks(dbd_ctx* c, int64_t cap, std::vector<Chunk>& out, const std::vector<char>* cpu = nullptr) {
  out.clear();
  int64_t b = 0, e = 0;
  while (b < c->nb) {
    if (cpu && (*cpu)[(size_t)b]) {
      e += c->h_blockC[(size_t)b++];
      continue;
    }
    Chunk ch{b, b, e, e};
    while (ch.b1 < c->nb) {
      if (cpu && (*cpu)[(size_t)ch.b1]) break;
      const int64_t add = c->h_blockC[(size_t)ch.b1];
      if (ch.b1 > ch.b0 && (ch.e1 - ch.e0) + add > cap) break;
      if (add > cap)
        return set_error(DBD_ERR_BUDGET,
                         "Not Enough Memory for Buffer: block %lld needs %lld output elements, chunk capacity %lld",
                         (long long)ch.b1, (long long)add, (long long)cap);
      ch.e1 += add;
      ++ch.b1;
    }
    out.push_back(ch);
    b = ch.b1;
    e = ch.e1;
  }
  return DBD_OK;
}

int64_t fixed_device_bytes(dbd_ctx* c) {
  int64_t s = c->a_bytes + (c->same ? 0 : c->b_bytes);
  s += (int64_t)c->P.bytes + c->Pptr.bytes + c->nnz.bytes + c->Cptr.bytes + c->rows_sym.bytes + c->rows_num.bytes;
  s += (int64_t)c->cnt.bytes + c->offs_sym.bytes + c->offs_num.bytes + c->tmp.bytes + c->cursor.bytes;
  s += (int64_t)c->bdesc_s.bytes + c->bdesc_l.bytes;
  s += (int64_t)c->band_hist.bytes + c->band_flag.bytes + c->band_tmp.bytes + c->band_slot.bytes + c->band_row.bytes +
       c->band.bytes + c->hub_list.bytes + c->hub_cnt.bytes + c->hub_scnt.bytes + c->hub_rcnt.bytes + c->hub_tmp.bytes +
       c->hub_scol.bytes + c->hub_sval.bytes + c->hub_slot.bytes + c->hub_flag.bytes + c->hub_rng.bytes;
  return s;
}

dbd_status run(dbd_ctx* c, int mode, dbd_timing* tm) {
  clear_error();
  if (!c->ops) return set_error(DBD_ERR_STATE, "dbd_run: operands not set");
  if (mode < 0 || mode > 2) return set_error(DBD_ERR_INVALID, "dbd_run: mode must be 0, 1 or 2");
  CHK(ensure_events(c));
  if (c->o.tune && !c->tuned_set) {
    int32_t id = 0;
    CHK(dbd_autotune(c, &id));
    c->tuned_set = true;
  }
  dbd_timing t{};
  t.h2d_s = c->h2d_s;
  c->dev_result = c->host_result = false;
  hipStream_t s = c->st;
  const double t0 = now_s();
  CHK(plan(c));
  const double t1 = now_s();
  t.preprocess_s = t1 - t0;
  HIPCHK(hipEventRecord(c->ev_t0, s));
  CHK(symbolic(c));
  HIPCHK(hipEventRecord(c->ev_c0, s));  // end of the symbolic pass (it synchronised the stream)
  const int64_t nnzC = c->info.nnz;
  const size_t esz = 4 + c->vsz;
  const int64_t budget = c->o.device_budget_mib << 20;
  const int64_t hbudget = c->o.host_budget_mib << 20;
  const int64_t avail = budget - fixed_device_bytes(c);
  if (avail <= 0)
    return set_error(DBD_ERR_BUDGET, "Not Enough Memory for Dense Matrix for Result: device budget %lld MiB",
                     (long long)c->o.device_budget_mib);
  // hybrid mode: the blocks analyze_Block sends to the CPU (outcome 1) go to host threads
  const bool hybrid = c->o.hybrid && mode != DBD_MODE_COMPUTE;
  if (hybrid && !c->host_ops)
    return set_error(DBD_ERR_INVALID, "hybrid mode needs host operand views (its host threads read A and B)");
  std::vector<char> cpu_blk;
  std::vector<int64_t> cpu_list;
  int64_t cpu_prod = 0;
  if (hybrid) {
    cpu_blk.assign((size_t)c->nb, 0);
    for (int64_t b = 0; b < c->nb; ++b)
      if (dbd_analyze_block(c->info.criteria, c->h_blockP[(size_t)b] / c->o.block_rows) == 1) {
        cpu_blk[(size_t)b] = 1;
        cpu_list.push_back(b);
        cpu_prod += c->h_blockP[(size_t)b];
      }
  }
  std::vector<Chunk> chunks;
  const bool whole_dev = (mode == DBD_MODE_COMPUTE) && (int64_t)(nnzC * esz) <= avail;
  int64_t cap;
  if (whole_dev) {
    cap = std::max<int64_t>(nnzC, 1);
  } else {
    cap = avail / (int64_t)(2 * esz);
    if (mode != DBD_MODE_COMPUTE) {
      const int hint = c->o.chunks_hint > 0 ? c->o.chunks_hint : 8;
      cap = std::min<int64_t>(cap, std::max<int64_t>((nnzC + hint - 1) / hint, 1));
    }
    // largest block must fit
    int64_t mxb = 0;
    for (int64_t v : c->h_blockC) mxb = std::max(mxb, v);
    cap = std::max(cap, std::min(mxb, avail / (int64_t)(2 * esz)));
  }
  CHK(make_chunks(c, cap, chunks, hybrid ? &cpu_blk : nullptr));
  int64_t maxch = 1;
  for (auto& ch : chunks) maxch = std::max(maxch, ch.e1 - ch.e0);
  const int nbuf = (whole_dev || mode == DBD_MODE_COMPUTE) ? 1 : 2;
  for (int i = 0; i < nbuf; ++i) {
    CHK(c->out_idx[i].ensure(4 * (size_t)maxch));
    CHK(c->out_val[i].ensure(c->vsz * (size_t)maxch));
  }
  // host side
  bool whole_host = false;
  Writer wr;
  int nslots = 0;
  if (mode != DBD_MODE_COMPUTE) {
    CHK(c->h_cptr.ensure(8 * (size_t)(c->nr + 1)));
    HIPCHK(hipMemcpyAsync(c->h_cptr.p, c->Cptr.p, 8 * (size_t)(c->nr + 1), hipMemcpyDeviceToHost, c->st));
    t.d2h_bytes += 8.0 * (double)(c->nr + 1);
    whole_host = (mode == DBD_MODE_D2H) && c->o.retain_result && (int64_t)(nnzC * esz) <= hbudget;
    if (whole_host) {
      CHK(c->h_idx.ensure(4 * (size_t)std::max<int64_t>(nnzC, 1)));
      CHK(c->h_val.ensure(c->vsz * (size_t)std::max<int64_t>(nnzC, 1)));
    } else {
      int want = (mode == DBD_MODE_IO) ? std::max(2, c->o.writer_threads + 1) : 2;
      const int64_t per = maxch * (int64_t)esz;
      if (per * 2 > hbudget)
        return set_error(DBD_ERR_BUDGET, "Not Enough Memory for Buffer: host budget %lld MiB < two chunks of %lld B",
                         (long long)c->o.host_budget_mib, (long long)per);
      want = (int)std::min<int64_t>(want, hbudget / per);
      nslots = want;
      c->ring_idx.resize(nslots);
      c->ring_val.resize(nslots);
      for (int i = 0; i < nslots; ++i) {
        CHK(c->ring_idx[i].ensure(4 * (size_t)maxch));
        CHK(c->ring_val[i].ensure(c->vsz * (size_t)maxch));
      }
      while ((int)c->ev_slot.size() < nslots) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_slot.push_back(e);
      }
    }
    if (mode == DBD_MODE_IO) {
      std::string d = (c->o.out_dir && c->o.out_dir[0]) ? std::string(c->o.out_dir) + "/" : std::string();
      std::string pi = d + "Idx.output", pv = d + "Val.output";
      wr.fd_idx = open(pi.c_str(), O_CREAT | O_WRONLY | O_TRUNC, 0644);
      wr.fd_val = open(pv.c_str(), O_CREAT | O_WRONLY | O_TRUNC, 0644);
      if (wr.fd_idx < 0 || wr.fd_val < 0)  // the writer's destructor closes what did open
        return set_error(DBD_ERR_IO, "cannot open %s / %s: %s", pi.c_str(), pv.c_str(), strerror(errno));
      wr.vsz = c->vsz;
      if (ftruncate(wr.fd_idx, (off_t)(4 * nnzC)) != 0 || ftruncate(wr.fd_val, (off_t)(c->vsz * nnzC)) != 0)
        return set_error(DBD_ERR_IO, "cannot size output files: %s", strerror(errno));
      wr.start(nslots);
    }
  }
  // hybrid: host threads take the CPU blocks off a shared counter while the GPU runs its chunks;
  // each writes its rows at their final offsets (Cptr from the symbolic pass): into the host
  // result, or a buffer handed to the file writers, or (a D2H run that keeps no host result) a
  // scratch buffer that is dropped like the GPU chunks' ring slots
  std::vector<std::thread> cpu_th;
  std::atomic<int64_t> cpu_next{0};
  std::atomic<bool> cpu_bad{false}, cpu_stop{false};
  std::atomic<int64_t> cpu_done_ns{0};
  struct CpuJoin {
    std::vector<std::thread>& th;
    std::atomic<bool>& stop;
    ~CpuJoin() {
      stop = true;
      for (auto& x : th)
        if (x.joinable()) x.join();
    }
  } cpu_join{cpu_th, cpu_stop};
  const double cpu_t0 = now_s();
  if (hybrid && !cpu_list.empty()) {
    HIPCHK(hipStreamSynchronize(c->st));  // h_cptr is read by the host threads
    const int64_t BR = c->o.block_rows;
    const int64_t* cp = c->h_cptr.as<int64_t>();
    struct HostPart {
      std::vector<int32_t> idx;
      std::vector<char> val;
    };
    auto worker = [&, BR, cp]() {
      HostRowScratch sc;
      HostPart scratch;
      for (;;) {
        const int64_t q = cpu_next.fetch_add(1);
        if (q >= (int64_t)cpu_list.size() || cpu_stop || cpu_bad) break;
        const int64_t b = cpu_list[(size_t)q];
        const int64_t r0 = b * BR, r1 = std::min(c->nr, r0 + BR);
        const int64_t e0 = cp[r0], n = cp[r1] - e0;
        bool ok;
        if (whole_host) {
          ok = host_rows(c->hostA, c->hostB, r0, r1, cp, 0, c->h_idx.as<int32_t>(), c->h_val.p, sc);
        } else {
          auto part = std::make_shared<HostPart>();
          HostPart& hp = mode == DBD_MODE_IO ? *part : scratch;
          hp.idx.resize((size_t)std::max<int64_t>(n, 1));
          hp.val.resize(c->vsz * (size_t)std::max<int64_t>(n, 1));
          ok = host_rows(c->hostA, c->hostB, r0, r1, cp, e0, hp.idx.data(), hp.val.data(), sc);
          if (ok && mode == DBD_MODE_IO && n > 0) wr.push({-1, nullptr, hp.idx.data(), hp.val.data(), n, e0, part});
        }
        if (!ok) cpu_bad = true;
      }
      const int64_t ns = (int64_t)((now_s() - cpu_t0) * 1e9);
      int64_t prev = cpu_done_ns.load();
      while (prev < ns && !cpu_done_ns.compare_exchange_weak(prev, ns)) {
      }
    };
    const int nt = (int)std::min<int64_t>(c->o.cpu_threads > 0 ? c->o.cpu_threads : hw_threads(),
                                          (int64_t)cpu_list.size());
    for (int w = 0; w < nt; ++w) cpu_th.emplace_back(worker);
  }
  // chunk loop
  double comp_ms = 0, copy_ms = 0;
  for (size_t i = 0; i < chunks.size(); ++i) {
    const Chunk& ch = chunks[i];
    const int bi = (int)(i % (size_t)nbuf);
    if (mode != DBD_MODE_COMPUTE && i >= (size_t)nbuf) HIPCHK(hipStreamWaitEvent(s, c->ev_copy[bi], 0));
    int32_t* di = c->out_idx[bi].as<int32_t>();
    void* dvv = c->out_val[bi].p;
    CHK(numeric_blocks(c, ch.b0, ch.b1, c->Cptr.as<int64_t>(), ch.e0, di, dvv));
    if (mode == DBD_MODE_COMPUTE) continue;
    HIPCHK(hipEventRecord(c->ev_comp[bi], s));
    HIPCHK(hipStreamWaitEvent(c->cst, c->ev_comp[bi], 0));
    const int64_t n = ch.e1 - ch.e0;
    void *hi, *hv;
    int slot = -1;
    if (whole_host) {
      hi = c->h_idx.as<char>() + 4 * ch.e0;
      hv = c->h_val.as<char>() + c->vsz * ch.e0;
    } else {
      slot = (int)(i % (size_t)nslots);
      if (mode == DBD_MODE_IO) {
        wr.acquire(slot);
        if (wr.failed) break;
      }
      hi = c->ring_idx[slot].p;
      hv = c->ring_val[slot].p;
    }
    if (n > 0) {
      HIPCHK(hipMemcpyAsync(hi, di, 4 * (size_t)n, hipMemcpyDeviceToHost, c->cst));
      HIPCHK(hipMemcpyAsync(hv, dvv, c->vsz * (size_t)n, hipMemcpyDeviceToHost, c->cst));
    }
    t.d2h_bytes += (double)esz * (double)n;
    HIPCHK(hipEventRecord(c->ev_copy[bi], c->cst));
    if (mode == DBD_MODE_IO) {
      HIPCHK(hipEventRecord(c->ev_slot[slot], c->cst));
      wr.push({slot, c->ev_slot[slot], hi, hv, n, ch.e0, {}});
    }
  }
  HIPCHK(hipEventRecord(c->ev_t1, s));
  for (auto& x : cpu_th) x.join();  // the host blocks are pushed to the writers before they stop
  cpu_th.clear();
  if (cpu_bad) return set_error(DBD_ERR_INTERNAL, "hybrid mode: a host row's column count differs from the symbolic pass");
  HIPCHK(hipStreamSynchronize(c->cst));
  HIPCHK(hipStreamSynchronize(s));
  if (mode == DBD_MODE_IO) {
    wr.finish();
    const int ci = close(wr.fd_idx), cv = close(wr.fd_val);  // both, whatever the first returns
    const bool closed_ok = ci == 0 && cv == 0;
    wr.fd_idx = wr.fd_val = -1;
    if (!closed_ok) return set_error(DBD_ERR_IO, "closing the output files failed: %s", strerror(errno));
    if (wr.failed) return set_error(DBD_ERR_IO, "Error during File Write: %s", wr.why.c_str());
    // Ptr.output last, like write_Ptr (Z/writeFile.cpp:177-197)
    const double w0 = now_s();
    dbd_csr pc{};
    pc.nrows = c->nr;
    pc.ncols = c->B.ncols;
    pc.nnz = 0;
    pc.ptr = c->h_cptr.as<int64_t>();
    pc.location = DBD_HOST;
    pc.dtype = c->o.dtype;
    std::string d = (c->o.out_dir && c->o.out_dir[0]) ? std::string(c->o.out_dir) + "/" : std::string();
    std::string pp = d + "Ptr.output";
    int fd = open(pp.c_str(), O_CREAT | O_WRONLY | O_TRUNC, 0644);
    if (fd < 0) return set_error(DBD_ERR_IO, "cannot open %s", pp.c_str());
    bool ok;
    const bool p64 = c->o.ptr64_file || nnzC > INT32_MAX;
    if (p64) {
      ok = Writer::pw(fd, c->h_cptr.p, 8 * (size_t)(c->nr + 1), 0);
    } else {
      std::vector<int32_t> p32((size_t)c->nr + 1);
      const int64_t* hp = c->h_cptr.as<int64_t>();
      for (int64_t r = 0; r <= c->nr; ++r) p32[(size_t)r] = (int32_t)hp[r];
      ok = Writer::pw(fd, p32.data(), 4 * p32.size(), 0);
    }
    ok = (close(fd) == 0) && ok;
    if (!ok) return set_error(DBD_ERR_IO, "write to %s failed", pp.c_str());
    t.write_s = wr.busy_s + (now_s() - w0);
    t.io_bytes = wr.bytes + (p64 ? 8.0 : 4.0) * (double)(c->nr + 1);
  }
  CHK(check_err(c, "numeric"));
  float ms = 0, sym_ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1));
  HIPCHK(hipEventElapsedTime(&sym_ms, c->ev_t0, c->ev_c0));
  comp_ms = ms;
  t.symbolic_s = sym_ms * 1e-3;
  t.numeric_s = (ms - sym_ms) * 1e-3;
  (void)copy_ms;
  const double t2 = now_s();
  t.compute_gpu_s = comp_ms * 1e-3;
  t.kernel_exec_s = t2 - t1;
  t.total_s = t2 - t0;
  t.chunks = (int64_t)chunks.size();
  t.d2h_s = 0;
  t.cpu_s = cpu_done_ns.load() * 1e-9;
  t.cpu_blocks = (int64_t)cpu_list.size();
  t.cpu_products = cpu_prod;
  (void)cpu_t0;
  c->dev_result = whole_dev && c->o.retain_result && mode == DBD_MODE_COMPUTE;
  c->host_result = whole_host;
  if (tm) *tm = t;
  return DBD_OK;
}

dbd_status upload(dbd_ctx* c, const dbd_csr* M, Mat& m, DevBuf* up, int64_t& bytes) {
  if (M->dtype != c->o.dtype)
    return set_error(DBD_ERR_INVALID, "operand dtype %d differs from the context dtype %d", M->dtype, c->o.dtype);
  if (M->nrows < 0 || M->ncols < 0 || M->nnz < 0 || M->nrows >= INT32_MAX || M->ncols > INT32_MAX)
    return set_error(DBD_ERR_OVERFLOW, "matrix dimensions must be below 2^31");
  m.nrows = M->nrows;
  m.ncols = M->ncols;
  m.nnz = M->nnz;
  const size_t vs = dtype_size(M->dtype);
  bytes = (int64_t)(8 * (M->nrows + 1) + (4 + vs) * M->nnz);
  if (M->location == DBD_DEVICE) {
    m.ptr = M->ptr;
    m.idx = M->idx;
    m.val = M->val;
    return DBD_OK;
  }
  // host view: validate structure, then upload
  if (M->ptr[0] != 0 || M->ptr[M->nrows] != M->nnz)
    return set_error(DBD_ERR_INVALID, "row pointer must start at 0 and end at nnz");
  for (int64_t r = 0; r < M->nrows; ++r) {
    if (M->ptr[r + 1] < M->ptr[r]) return set_error(DBD_ERR_INVALID, "row pointer decreases at row %lld", (long long)r);
    for (int64_t i = M->ptr[r]; i < M->ptr[r + 1]; ++i) {
      const int32_t col = M->idx[i];
      if (col < 0 || col >= M->ncols)
        return set_error(DBD_ERR_INVALID, "column %d out of range in row %lld", col, (long long)r);
      if (i > M->ptr[r] && col <= M->idx[i - 1])
        return set_error(DBD_ERR_INVALID, "columns of row %lld are not strictly increasing", (long long)r);
    }
  }
  CHK(up[0].ensure(8 * (size_t)(M->nrows + 1)));
  CHK(up[1].ensure(4 * (size_t)std::max<int64_t>(M->nnz, 1)));
  CHK(up[2].ensure(vs * (size_t)std::max<int64_t>(M->nnz, 1)));
  HIPCHK(hipMemcpyAsync(up[0].p, M->ptr, 8 * (size_t)(M->nrows + 1), hipMemcpyHostToDevice, c->st));
  if (M->nnz > 0) {
    HIPCHK(hipMemcpyAsync(up[1].p, M->idx, 4 * (size_t)M->nnz, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(up[2].p, M->val, vs * (size_t)M->nnz, hipMemcpyHostToDevice, c->st));
  }
  m.ptr = up[0].as<int64_t>();
  m.idx = up[1].as<int32_t>();
  m.val = up[2].p;
  return DBD_OK;
}

std::string problem_class(dbd_ctx* c) {
  // coarse features: dtype, log2 of products per row, log2 ncols, log2 of nnz per B row
  auto lg = [](double x) { return x <= 1 ? 0 : (int)std::floor(std::log2(x)); };
  char buf[128];
  const double ppr = c->nr > 0 ? (double)c->info.products / (double)c->nr : 0;
  const double bpr = c->B.nrows > 0 ? (double)c->B.nnz / (double)c->B.nrows : 0;
  snprintf(buf, sizeof(buf), "dt%d_ppr%d_nc%d_bpr%d", c->o.dtype, lg(ppr), lg((double)c->B.ncols), lg(bpr));
  return buf;
}

}  // namespace

extern "C" {

dbd_status dbd_create(const dbd_opts* opts, dbd_ctx** out) {
  clear_error();
  if (!out) return set_error(DBD_ERR_INVALID, "dbd_create: out is NULL");
  *out = nullptr;
  dbd_opts o;
  dbd_default_opts(&o);
  if (opts) {
    o = *opts;
    if (o.block_rows == 0) o.block_rows = 4096;
    if (o.host_budget_mib == 0) o.host_budget_mib = 28 << 10;
    if (o.device_budget_mib == 0) o.device_budget_mib = 22 << 10;
    if (o.writer_threads == 0) o.writer_threads = 4;
    if (o.chunks_hint == 0) o.chunks_hint = 8;
  }
  if (o.block_rows <= 0 || o.block_rows % 256 != 0)
    return set_error(DBD_ERR_INVALID, "block_rows must be a positive multiple of 256 (got %d)", o.block_rows);
  if (o.dtype != DBD_F32 && o.dtype != DBD_F64) return set_error(DBD_ERR_INVALID, "dtype must be DBD_F32 or DBD_F64");
  if (o.host_budget_mib < 0 || o.device_budget_mib < 0) return set_error(DBD_ERR_INVALID, "negative budget");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return set_error(DBD_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  }
  dbd_ctx* c = new dbd_ctx();
  c->o = o;
  c->vsz = dtype_size(o.dtype);
  c->tb = candidate_tables()[0];
  if (hipGetDevice(&c->dev) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_error(DBD_ERR_HIP, "stream creation failed");
  }
  c->own_st = true;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->dev) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  *out = c;
  return DBD_OK;
}

dbd_status dbd_destroy(dbd_ctx* c) {
  if (!c) return DBD_OK;
  (void)hipStreamSynchronize(c->st);  // a null st is the default stream
  if (c->cst) (void)hipStreamSynchronize(c->cst);
  for (auto* b : {&c->P, &c->Pptr, &c->nnz, &c->Cptr, &c->rows_sym, &c->rows_num, &c->cnt, &c->offs_sym, &c->offs_num,
                  &c->tmp, &c->cursor, &c->blk, &c->err, &c->g_rng, &c->bdesc_s, &c->bdesc_l, &c->band_hist,
                  &c->band_flag, &c->band_tmp, &c->band_slot, &c->band_row, &c->band, &c->hub_list, &c->hub_cnt,
                  &c->hub_scnt, &c->hub_rcnt, &c->hub_tmp, &c->hub_scol, &c->hub_sval, &c->hub_slot, &c->hub_flag,
                  &c->hub_rng})
    b->release();
  for (int i = 0; i < 3; ++i) {
    c->upA[i].release();
    c->upB[i].release();
  }
  for (int i = 0; i < 2; ++i) {
    c->out_idx[i].release();
    c->out_val[i].release();
    if (c->ev_comp[i]) (void)hipEventDestroy(c->ev_comp[i]);
    if (c->ev_copy[i]) (void)hipEventDestroy(c->ev_copy[i]);
  }
  for (auto e : c->ev_slot) (void)hipEventDestroy(e);
  for (auto e : {c->ev_t0, c->ev_t1, c->ev_c0, c->ev_c1})
    if (e) (void)hipEventDestroy(e);
  c->h_cptr.release();
  c->h_idx.release();
  c->h_val.release();
  for (auto& b : c->ring_idx) b.release();
  for (auto& b : c->ring_val) b.release();
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->own_st && c->st) (void)hipStreamDestroy(c->st);
  if (c->cst) (void)hipStreamDestroy(c->cst);
  delete c;
  return DBD_OK;
}

// The caller's stream, taken as it is: NULL is the HIP default stream, the one a framework's
// default-stream work (torch's allocations, fills, elementwise ops) is ordered on.
dbd_status dbd_set_stream(dbd_ctx* c, void* stream) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  HIPCHK(hipStreamSynchronize(c->st));
  if (c->own_st) (void)hipStreamDestroy(c->st);
  c->st = (hipStream_t)stream;
  c->own_st = false;
  return DBD_OK;
}

dbd_status dbd_use_own_stream(dbd_ctx* c) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  if (c->own_st) return DBD_OK;
  HIPCHK(hipStreamSynchronize(c->st));
  hipStream_t st = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  c->st = st;
  c->own_st = true;
  return DBD_OK;
}

dbd_status dbd_set_operands(dbd_ctx* c, const dbd_csr* A, const dbd_csr* B) {
  clear_error();
  if (!c || !A) return set_error(DBD_ERR_INVALID, "dbd_set_operands: null argument");
  if (!B) B = A;
  if (A->ncols != B->nrows)
    return set_error(DBD_ERR_INVALID, "inner dimensions differ: A is %lld x %lld, B is %lld x %lld",
                     (long long)A->nrows, (long long)A->ncols, (long long)B->nrows, (long long)B->ncols);
  c->ops = false;
  c->planned = c->symb = false;
  c->tuned_set = c->table_explicit;
  c->band_ready = false;
  c->band_rows = 0;
  c->dev_result = c->host_result = false;
  const double t0 = now_s();
  CHK(upload(c, A, c->A, c->upA, c->a_bytes));
  c->same = (A == B) || (A->ptr == B->ptr && A->idx == B->idx && A->val == B->val);
  if (c->same) {
    c->B = c->A;
    c->b_bytes = c->a_bytes;
  } else {
    CHK(upload(c, B, c->B, c->upB, c->b_bytes));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  c->h2d_s = (A->location == DBD_HOST || B->location == DBD_HOST) ? now_s() - t0 : 0.0;
  c->host_ops = A->location == DBD_HOST && B->location == DBD_HOST;
  c->hostA = *A;
  c->hostB = *B;
  c->ops = true;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (c->graph) {
    (void)hipGraphDestroy(c->graph);
    c->graph = nullptr;
  }
  return DBD_OK;
}

dbd_status dbd_plan(dbd_ctx* c, dbd_plan_info* info) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  CHK(plan(c));
  if (info) *info = c->info;
  return DBD_OK;
}

dbd_status dbd_get_plan_info(dbd_ctx* c, dbd_plan_info* info) {
  if (!c || !info) return set_error(DBD_ERR_INVALID, "null argument");
  if (!c->planned) return set_error(DBD_ERR_STATE, "not planned");
  *info = c->info;
  return DBD_OK;
}

dbd_status dbd_get_row_products(dbd_ctx* c, int64_t* hP) {
  clear_error();
  if (!c || !hP) return set_error(DBD_ERR_INVALID, "null argument");
  if (!c->planned) return set_error(DBD_ERR_STATE, "not planned");
  HIPCHK(hipMemcpyAsync(hP, c->P.p, 8 * (size_t)c->nr, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return DBD_OK;
}

dbd_status dbd_get_blocks(dbd_ctx* c, int64_t* bptr, int64_t* bprod) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  if (!c->planned) return set_error(DBD_ERR_STATE, "not planned");
  for (int64_t b = 0; b <= c->nb; ++b)
    if (bptr) bptr[b] = std::min<int64_t>(b * c->o.block_rows, c->nr);
  for (int64_t b = 0; b < c->nb; ++b)
    if (bprod) bprod[b] = c->h_blockP[(size_t)b];
  return DBD_OK;
}

dbd_status dbd_run(dbd_ctx* c, int32_t mode, dbd_timing* tm) {
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  return run(c, mode, tm);
}

dbd_status dbd_result(dbd_ctx* c, dbd_csr* C) {
  clear_error();
  if (!c || !C) return set_error(DBD_ERR_INVALID, "null argument");
  memset(C, 0, sizeof(*C));
  C->nrows = c->nr;
  C->ncols = c->B.ncols;
  C->nnz = c->info.nnz;
  C->dtype = c->o.dtype;
  if (c->dev_result) {
    C->ptr = c->Cptr.as<int64_t>();
    C->idx = c->out_idx[0].as<int32_t>();
    C->val = c->out_val[0].p;
    C->location = DBD_DEVICE;
    return DBD_OK;
  }
  if (c->host_result) {
    C->ptr = c->h_cptr.as<int64_t>();
    C->idx = c->h_idx.as<int32_t>();
    C->val = c->h_val.p;
    C->location = DBD_HOST;
    return DBD_OK;
  }
  return set_error(DBD_ERR_STATE, "no retained result (run COMPUTE or D2H with retain_result=1 within budget)");
}

dbd_status dbd_copy_result(dbd_ctx* c, int64_t* h_ptr, int32_t* h_idx, void* h_val) {
  clear_error();
  dbd_csr C;
  CHK(dbd_result(c, &C));
  const hipMemcpyKind k = C.location == DBD_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
  if (h_ptr) HIPCHK(hipMemcpyAsync(h_ptr, C.ptr, 8 * (size_t)(C.nrows + 1), k, c->st));
  if (h_idx && C.nnz > 0) HIPCHK(hipMemcpyAsync(h_idx, C.idx, 4 * (size_t)C.nnz, k, c->st));
  if (h_val && C.nnz > 0) HIPCHK(hipMemcpyAsync(h_val, C.val, c->vsz * (size_t)C.nnz, k, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return DBD_OK;
}

dbd_status dbd_symbolic(dbd_ctx* c, int64_t* d_Cptr, int64_t* nnz_out) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  if (!c->planned) CHK(plan(c));
  CHK(symbolic(c));
  if (d_Cptr)
    HIPCHK(hipMemcpyAsync(d_Cptr, c->Cptr.p, 8 * (size_t)(c->nr + 1), hipMemcpyDeviceToDevice, c->st));
  if (nnz_out) *nnz_out = c->info.nnz;
  return DBD_OK;
}

dbd_status dbd_numeric(dbd_ctx* c, const int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval) {
  clear_error();
  if (!c || !d_Cptr || (c->info.nnz > 0 && (!d_Cidx || !d_Cval))) return set_error(DBD_ERR_INVALID, "null argument");
  if (!c->symb) return set_error(DBD_ERR_STATE, "dbd_numeric before dbd_symbolic");
  launch_fill_i32(c->err.as<int32_t>(), 1, 0, c->st);
  CHK(numeric_blocks(c, 0, c->nb, d_Cptr, 0, d_Cidx, d_Cval));
  return check_err(c, "numeric");
}

dbd_status dbd_backward(dbd_ctx* c, const int64_t* d_Cptr, const int32_t* d_Cidx, const void* d_dC, void* d_dA,
                        void* d_dB) {
  clear_error();
  if (!c || !d_Cptr || !d_Cidx || !d_dC) return set_error(DBD_ERR_INVALID, "null argument");
  if (!c->symb) return set_error(DBD_ERR_STATE, "dbd_backward needs the forward plan (run dbd_symbolic first)");
  hipStream_t s = c->st;
  launch_fill_i32(c->err.as<int32_t>(), 1, 0, s);
  if (d_dA) launch_fill_bytes0(d_dA, (int64_t)c->vsz * c->A.nnz, s);
  const CatTable& nt = c->tb.num_ct;
  const int64_t* offs = c->offs_num.as<int64_t>();
  for (int k = 1; k < nt.n; ++k) {
    // long and hub rows both take the window backward kernel, which keeps a cursor per A entry;
    // the two categories lie next to each other in the bins and share one launch, which takes
    // its rows from the end (the hub rows, the heaviest, first)
    const bool with_hub = c->tb.num_k[k].T == 0 && k + 1 < nt.n && c->tb.num_k[k + 1].T == kHubShape;
    const int64_t i0 = (int64_t)k * c->nwg, i1 = (int64_t)(k + (with_hub ? 2 : 1)) * c->nwg;
    const int64_t n = c->h_offs_num[(size_t)i1] - c->h_offs_num[(size_t)i0];
    const KernelShape ks = c->tb.num_k[k];
    if (with_hub) ++k;
    if (n <= 0) continue;
    if (ks.T == 0 || ks.T == kHubShape) CHK(ensure_cursor(c));
    launch_backward(c->o.dtype, ks, c->rows_num.as<int32_t>(), offs + i0, offs + i1, n, 0, c->A, c->B, d_Cptr,
                    d_Cidx, d_dC, d_dA, d_dB, c->cursor.as<int64_t>(), c->err.as<int>(), s);
  }
  return check_err(c, "backward");
}

}  // extern "C"

// ---- whole pipeline without host synchronisation (batched graph, no-allocation entry) -----
static int persistent_grid(int num_cu, const KernelShape& k) {
  if (k.T == 0) return num_cu * 2;
  return num_cu * (k.NT <= 64 ? 8 : 2);
}

// The whole pipeline (row analysis, bins, symbolic, scan, bins, numeric) on stream s with
// persistent category kernels that read their row ranges from device memory: no host
// synchronisation and no allocation, so it can be captured in a graph. If nnz(C) exceeds `cap`,
// the numeric kernels drain empty ranges and the capacity bit is set in the error word. Hub rows
// take their range kernels (w.hub_fn, buckets sized beforehand) when w.hub_slot is set, else they
// stay in the long category.
void dbd::enqueue_pipeline(const Work& w, const Tables& tb, Mat A, const Mat& B, int dtype, int num_cu,
                             int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval, int64_t cap, hipStream_t s) {
  const int64_t nr = w.nr;
  A.bstart = w.bstart;
  A.blen = w.blen;
  launch_fill_i32(w.err, 1, 0, s);
  launch_row_products(A, B, w.P, w.nnz, w.bstart, w.blen, s);
  const CatTable& ct = tb.sym_ct;
  launch_bin_count(w.P, nr, B.ncols, ct, w.cnt, s);
  launch_exclusive_scan(w.cnt, w.offs_sym, (int64_t)ct.n * w.nwg, w.tmp, s);
  launch_bin_scatter(w.P, nr, B.ncols, ct, w.offs_sym, w.rows_sym, s);
  for (int k = 1; k < ct.n; ++k)
    launch_symbolic(tb.sym_k[k], w.rows_sym, w.offs_sym + (int64_t)k * w.nwg, w.offs_sym + (int64_t)(k + 1) * w.nwg, 0,
                    persistent_grid(num_cu, tb.sym_k[k]), A, B, w.nnz, w.cursor, w.err, s, w.hub_slot, w.hub_rng,
                    w.nrng);
  launch_exclusive_scan(w.nnz, d_Cptr, nr, w.tmp, s);
  const CatTable& nt = tb.num_ct;
  launch_bin_count(w.nnz, nr, INT64_MAX, nt, w.cnt, s, w.hub_slot);
  launch_exclusive_scan(w.cnt, w.offs_num, (int64_t)nt.n * w.nwg, w.tmp, s);
  launch_bin_scatter(w.nnz, nr, INT64_MAX, nt, w.offs_num, w.rows_num, s, w.hub_slot);
  launch_capacity_guard(d_Cptr + nr, cap, w.offs_num, (int64_t)nt.n * w.nwg + 1, w.err, s);
  for (int k = 1; k < nt.n; ++k) {
    if (tb.num_k[k].T == kHubShape) continue;  // the range kernels below (empty without hub slots)
    launch_numeric(dtype, tb.num_k[k], w.rows_num, w.offs_num + (int64_t)k * w.nwg, w.offs_num + (int64_t)(k + 1) * w.nwg,
                   0, persistent_grid(num_cu, tb.num_k[k]), A, B, d_Cptr, 0, d_Cidx, d_Cval, w.cursor, w.err, s, false);
  }
  if (w.hub_slot && w.hub_fn) w.hub_fn(w.hub_arg, d_Cptr, d_Cidx, d_Cval, s);
}

// The batch's pipeline on the context's workspace. C was sized by the symbolic pass of
// dbd_batch_prepare; should the operands' structure have changed since, a larger output is
// refused on the device instead of overrunning C.
static dbd_status batch_enqueue(dbd_ctx* c, hipStream_t s, int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval) {
  Work w;
  w.P = c->P.as<int64_t>();
  w.nnz = c->nnz.as<int64_t>();
  w.cnt = c->cnt.as<int64_t>();
  w.offs_sym = c->offs_sym.as<int64_t>();
  w.offs_num = c->offs_num.as<int64_t>();
  w.tmp = c->tmp.as<int64_t>();
  w.bstart = c->bdesc_s.as<int64_t>();
  w.cursor = c->cursor.as<int64_t>();
  w.rows_sym = c->rows_sym.as<int32_t>();
  w.rows_num = c->rows_num.as<int32_t>();
  w.blen = c->bdesc_l.as<int32_t>();
  w.err = c->err.as<int>();
  w.nr = c->nr;
  w.nwg = c->nwg;
  if (c->bh_n > 0) {
    w.hub_slot = c->hub_slot.as<int32_t>();
    w.hub_rng = c->hub_rng.as<int32_t>();
    w.nrng = hub_ranges(c->B.ncols);
    w.hub_arg = c;
    w.hub_fn = [](void* arg, const int64_t* Cptr, int32_t* Cidx, void* Cval, hipStream_t st) {
      dbd_ctx* cc = static_cast<dbd_ctx*>(arg);
      hub_head(cc, cc->bh_pos0, cc->bh_n, st);
      hub_tail(cc, cc->bh_pos0, cc->bh_n, cc->bh_nshort, Cptr, 0, Cidx, Cval, st);
    };
  }
  enqueue_pipeline(w, c->tb, c->A, c->B, c->o.dtype, c->num_cu, d_Cptr, d_Cidx, d_Cval, c->batch_cap, s);
  HIPCHK(hipGetLastError());  // a launch the runtime refused
  return DBD_OK;
}

extern "C" {

dbd_status dbd_batch_prepare(dbd_ctx* c, const dbd_csr* A_cat, const dbd_csr* B_cat, int64_t* nnz_bound) {
  clear_error();
  if (!c || !A_cat) return set_error(DBD_ERR_INVALID, "null argument");
  CHK(dbd_set_operands(c, A_cat, B_cat));
  CHK(plan(c));
  CHK(ensure_cursor(c));
  // one eager symbolic pass: the exact size of C, which the caller allocates (the graph then
  // recomputes the same counts on every replay)
  CHK(symbolic(c));
  c->batch_cap = c->info.nnz;
  // hub rows: their place in the numeric bins (the graph bins the same way) and the workspace of
  // their range kernels, buckets included, sized here so the graph neither allocates nor waits
  c->bh_n = c->bh_pos0 = c->bh_nshort = 0;
  const CatTable& nt = c->tb.num_ct;
  if (c->hub_n > 0 && c->tb.num_k[nt.n - 1].T == kHubShape) {
    const int64_t p0 = c->h_offs_num[(size_t)(nt.n - 1) * c->nwg], p1 = c->h_offs_num[(size_t)nt.n * c->nwg];
    if (p1 > p0) {
      CHK(hub_reserve(c, p1 - p0));
      hub_head(c, p0, p1 - p0, c->st);
      CHK(hub_buckets(c, p1 - p0, &c->bh_nshort));
      c->bh_pos0 = p0;
      c->bh_n = p1 - p0;
    }
  }
  if (nnz_bound) *nnz_bound = c->info.nnz;
  return DBD_OK;
}

dbd_status dbd_batch_launch(dbd_ctx* c, int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval) {
  clear_error();
  if (!c || !d_Cptr || !d_Cidx || !d_Cval) return set_error(DBD_ERR_INVALID, "null argument");
  if (!c->planned) return set_error(DBD_ERR_STATE, "dbd_batch_launch before dbd_batch_prepare");
  const void* key[3] = {d_Cptr, d_Cidx, d_Cval};
  if (!c->gexec || memcmp(key, c->g_key, sizeof(key)) != 0) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    if (!c->st) return set_error(DBD_ERR_STATE, "dbd_batch_launch: the default stream cannot be captured; set another");
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
    dbd_status st = batch_enqueue(c, c->st, d_Cptr, d_Cidx, d_Cval);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->st, &g);
    if (st != DBD_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess || !g) return set_error(DBD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    memcpy(c->g_key, key, sizeof(key));
  }
  HIPCHK(hipGraphLaunch(c->gexec, c->st));
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->err.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->symb = false;
  if (e & kErrCapacityBit)
    return set_error(DBD_ERR_BUDGET, "batch: C needs more than the %lld entries dbd_batch_prepare reported "
                     "(operand structure changed since); nothing was written", (long long)c->batch_cap);
  if (e != 0) return set_error(DBD_ERR_INTERNAL, "batch: device consistency check failed (flags 0x%x)", e);
  return DBD_OK;
}

}  // extern "C"

extern "C" {
int32_t dbd_table_count(void) { return (int32_t)candidate_tables().size(); }

dbd_status dbd_table_knobs(int32_t id, dbd_table_desc* out) {
  clear_error();
  if (!out) return set_error(DBD_ERR_INVALID, "null argument");
  if (id < 0 || id >= (int32_t)candidate_tables().size())
    return set_error(DBD_ERR_INVALID, "table id %d out of range [0, %d)", id, (int)candidate_tables().size());
  const Tables& t = candidate_tables()[(size_t)id];
  out->g_small = t.g_small;
  out->g_big = t.g_big;
  out->load = t.load;
  out->top = t.top;
  out->hub_min_cols = t.hub_min_cols;
  return DBD_OK;
}

dbd_status dbd_set_table(dbd_ctx* c, int32_t id) {
  clear_error();
  if (!c) return set_error(DBD_ERR_INVALID, "null ctx");
  if (id < 0 || id >= (int32_t)candidate_tables().size())
    return set_error(DBD_ERR_INVALID, "table id %d out of range [0, %d)", id, (int)candidate_tables().size());
  c->tb = candidate_tables()[(size_t)id];
  c->planned = c->symb = false;
  c->tuned_set = c->table_explicit = true;  // an explicit choice is not overridden by opts.tune
  return DBD_OK;
}

dbd_status dbd_autotune(dbd_ctx* c, int32_t* chosen) {
  clear_error();
  if (!c || !c->ops) return set_error(DBD_ERR_STATE, "dbd_autotune: operands not set");
  CHK(ensure_events(c));
  CHK(plan(c));
  const std::string key = problem_class(c);
  int hit = -1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) hit = it->second;
  }
  if (hit >= 0) {  // re-plan: the bins and limits must be the chosen table's
    c->tb = candidate_tables()[(size_t)hit];
    CHK(plan(c));
    if (chosen) *chosen = hit;
    return DBD_OK;
  }
  // Coordinate descent over the knobs from the default table: each knob in turn takes the value
  // whose table runs symbolic + numeric fastest on the real operands, the others held at the best
  // so far (1 + 1 + 1 + 1 + 1 + 2 = 7 tables timed instead of all 48).
  DevBuf ci, cv;
  std::map<int, double> timed;
  auto time_table = [&](int id, double* ms_out) -> dbd_status {
    auto it = timed.find(id);
    if (it != timed.end()) {
      *ms_out = it->second;
      return DBD_OK;
    }
    c->tb = candidate_tables()[(size_t)id];
    CHK(plan(c));
    double tot = 0;
    for (int rep = 0; rep < 2; ++rep) {
      HIPCHK(hipEventRecord(c->ev_t0, c->st));
      CHK(symbolic(c));
      CHK(ci.ensure(4 * (size_t)std::max<int64_t>(c->info.nnz, 1)));
      CHK(cv.ensure(c->vsz * (size_t)std::max<int64_t>(c->info.nnz, 1)));
      CHK(numeric_blocks(c, 0, c->nb, c->Cptr.as<int64_t>(), 0, ci.as<int32_t>(), cv.p));
      HIPCHK(hipEventRecord(c->ev_t1, c->st));
      HIPCHK(hipEventSynchronize(c->ev_t1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1));
      if (rep > 0) tot += ms;
    }
    CHK(check_err(c, "autotune"));
    timed[id] = tot;
    *ms_out = tot;
    return DBD_OK;
  };
  int best = 0;
  double best_ms = 0;
  CHK(time_table(best, &best_ms));
  for (int k = 0; k < 5; ++k) {
    const int base = best;
    for (int x = 0; x < kKnobRadix[k]; ++x) {
      const int id = table_with_knob(base, k, x);
      double ms = 0;
      CHK(time_table(id, &ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = id;
      }
    }
  }
  ci.release();
  cv.release();
  c->tb = candidate_tables()[(size_t)best];
  CHK(plan(c));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  if (chosen) *chosen = best;
  return DBD_OK;
}

}  // extern "C"
