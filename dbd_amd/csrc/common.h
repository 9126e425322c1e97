// Shared host-side helpers of libdbd: thread-local error message, status helpers,
// a small std::thread parallel-for (no OpenMP runtime inside the library).
#pragma once
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dbd.h"

namespace dbd {

dbd_status set_error(dbd_status st, const char* fmt, ...);
void clear_error();

// Reference block criteria and categories (calculate_Criteria + analyze_Block) from per-block
// product totals: fills num_blocks, block_rows, products, max_block_products, criteria and
// blocks_{gpu_only,cpu_pref,any} of info. The one implementation the engine and
// dbd_plan_blocks share.
void block_criteria(const int64_t* block_products, int64_t nb, int32_t block_rows, dbd_plan_info* info);

// Host multiply of the hybrid mode: rows [r0, r1) of C = A*B (host CSR views of one dtype) into
// idx/val at positions Cptr[r] - base, columns in increasing order (Gustavson's row-wise method
// with a dense accumulator and a touched-column list, the scheme of compute_CPU,
// Z/computeCPU.cpp:10-98, with the first touch decided by structure, not by |acc| < 1e-4).
// Returns false if a row's column count differs from Cptr's (the GPU's symbolic count).
struct HostRowScratch {
  std::vector<double> acc;      // ncols
  std::vector<int64_t> stamp;   // ncols: last row that touched the column, -1 never
  std::vector<int32_t> cols;    // touched columns of the current row
};
bool host_rows(const dbd_csr& A, const dbd_csr& B, int64_t r0, int64_t r1, const int64_t* Cptr, int64_t base,
               int32_t* idx, void* val, HostRowScratch& s);

inline int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 64) n = 64;
  return (int)n;
}

// Static partition of [0, n) over up to `threads` workers; f(begin, end, worker).
template <typename F>
void parallel_for(int64_t n, F&& f, int threads = 0) {
  if (threads <= 0) threads = hw_threads();
  if (n <= 0) return;
  if (threads > n) threads = (int)n;
  if (threads <= 1 || n < 4096) {
    f((int64_t)0, n, 0);
    return;
  }
  std::vector<std::thread> ts;
  ts.reserve(threads);
  for (int w = 0; w < threads; ++w) {
    int64_t b = n * w / threads, e = n * (w + 1) / threads;
    ts.emplace_back([&, b, e, w]() { f(b, e, w); });
  }
  for (auto& t : ts) t.join();
}

inline size_t dtype_size(int32_t dt) { return dt == DBD_F64 ? 8 : 4; }

// splitmix64: counter-based stream so generated data is independent of thread count.
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed, uint64_t stream) : s(splitmix64(seed * 0x9E3779B97F4A7C15ull ^ splitmix64(stream))) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

}  // namespace dbd
