// Host-only parts of libdbd: error state, block planning (reference criteria), matrix file
// formats, output files, and seeded generators. No HIP calls here, so everything in this
// file runs (and is tested) without a GPU.
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <numeric>
#include <string>
#include <unordered_set>
#include <vector>

#include "common.h"

namespace dbd {

static thread_local std::string g_err;

dbd_status set_error(dbd_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}
void clear_error() { g_err.clear(); }

}  // namespace dbd

using namespace dbd;

extern "C" {

int32_t dbd_abi_version(void) { return DBD_ABI_VERSION; }
const char* dbd_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace dbd {

// calculate_Criteria (Z/distributeWorkload.cpp:249-264) and analyze_Block per block. The
// reference accumulates (avg - x)^2 in int (which overflows for large blocks) into a float;
// here the square is formed in int64 and the float accumulation is kept.
void block_criteria(const int64_t* block_products, int64_t nb, int32_t block_rows, dbd_plan_info* info) {
  int64_t total = 0, mx = 0;
  for (int64_t b = 0; b < nb; ++b) {
    total += block_products[b];
    mx = std::max(mx, block_products[b]);
  }
  info->num_blocks = nb;
  info->block_rows = block_rows;
  info->products = total;
  info->max_block_products = mx;
  int64_t criteria = 0, g = 0, c = 0, a = 0;
  if (nb > 0) {
    const int64_t avg = (total / nb) / block_rows;
    float sd = 0.0f;
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t d = avg - block_products[b] / block_rows;
      sd += (float)(d * d);
    }
    sd /= (float)nb;
    sd = sqrtf(sd);
    criteria = (int64_t)(1.28f * sd) + avg;
    for (int64_t b = 0; b < nb; ++b) {
      switch (dbd_analyze_block(criteria, block_products[b] / block_rows)) {
        case 0: ++g; break;
        case 1: ++c; break;
        default: ++a; break;
      }
    }
  }
  info->criteria = criteria;
  info->blocks_gpu_only = g;
  info->blocks_cpu_pref = c;
  info->blocks_any = a;
}

bool host_rows(const dbd_csr& A, const dbd_csr& B, int64_t r0, int64_t r1, const int64_t* Cptr, int64_t base,
               int32_t* idx, void* val, HostRowScratch& s) {
  if ((int64_t)s.acc.size() < B.ncols) {
    s.acc.assign((size_t)B.ncols, 0.0);
    s.stamp.assign((size_t)B.ncols, -1);
  }
  const bool f64 = A.dtype == DBD_F64;
  for (int64_t r = r0; r < r1; ++r) {
    s.cols.clear();
    for (int64_t ka = A.ptr[r]; ka < A.ptr[r + 1]; ++ka) {
      const int32_t k = A.idx[ka];
      const double a = f64 ? static_cast<const double*>(A.val)[ka] : (double)static_cast<const float*>(A.val)[ka];
      for (int64_t kb = B.ptr[k]; kb < B.ptr[k + 1]; ++kb) {
        const int32_t j = B.idx[kb];
        const double b = f64 ? static_cast<const double*>(B.val)[kb] : (double)static_cast<const float*>(B.val)[kb];
        if (s.stamp[(size_t)j] != r) {
          s.stamp[(size_t)j] = r;
          s.acc[(size_t)j] = a * b;
          s.cols.push_back(j);
        } else {
          s.acc[(size_t)j] += a * b;
        }
      }
    }
    const int64_t o = Cptr[r] - base;
    if ((int64_t)s.cols.size() != Cptr[r + 1] - Cptr[r]) return false;
    std::sort(s.cols.begin(), s.cols.end());
    for (size_t q = 0; q < s.cols.size(); ++q) {
      const int32_t j = s.cols[q];
      idx[o + (int64_t)q] = j;
      if (f64)
        static_cast<double*>(val)[o + (int64_t)q] = s.acc[(size_t)j];
      else
        static_cast<float*>(val)[o + (int64_t)q] = (float)s.acc[(size_t)j];
    }
  }
  return true;
}

}  // namespace dbd

extern "C" {

void dbd_default_opts(dbd_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->block_rows = 4096;
  o->dtype = DBD_F64;
  o->host_budget_mib = 28 << 10;
  o->device_budget_mib = 22 << 10;
  o->writer_threads = 4;
  o->chunks_hint = 8;
  o->retain_result = 1;
}

// analyze_Block, Z/distributeWorkload.cpp:266-270.
int32_t dbd_analyze_block(int64_t criteria, int64_t avg) {
  if (avg > criteria) return 0;
  if (avg <= 64) return 1;
  return 2;
}

// divide_Blocks (Z/main.cpp:259-296) + calculate_Criteria (Z/distributeWorkload.cpp:249-264).
dbd_status dbd_plan_blocks(const int64_t* P, int64_t nr, int32_t block_rows, int64_t* block_ptr,
                           int64_t* block_products, dbd_plan_info* info) {
  if (nr < 0 || block_rows <= 0 || (nr > 0 && (!P || !block_ptr || !block_products)))
    return set_error(DBD_ERR_INVALID, "dbd_plan_blocks: bad arguments");
  const int64_t nb = nr == 0 ? 0 : (nr - 1) / block_rows + 1;
  int64_t total = 0, mx = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t r0 = b * block_rows, r1 = std::min(nr, r0 + block_rows);
    int64_t s = 0;
    for (int64_t r = r0; r < r1; ++r) s += P[r];
    block_ptr[b] = r0;
    block_products[b] = s;
    total += s;
    mx = std::max(mx, s);
  }
  if (nb > 0) block_ptr[nb] = nr;
  if (info) {
    info->nrows = nr;
    block_criteria(block_products, nb, block_rows, info);
  }
  return DBD_OK;
}

dbd_status dbd_csr_free(dbd_csr* M) {
  if (!M) return DBD_OK;
  free(M->ptr);
  free(M->idx);
  free(M->val);
  M->ptr = nullptr;
  M->idx = nullptr;
  M->val = nullptr;
  return DBD_OK;
}

}  // extern "C"

namespace {

// Allocate a host CSR owned by the library (malloc, freed by dbd_csr_free).
dbd_status alloc_csr(dbd_csr* out, int64_t nr, int64_t nc, int64_t nnz, int32_t dtype) {
  memset(out, 0, sizeof(*out));
  out->nrows = nr;
  out->ncols = nc;
  out->nnz = nnz;
  out->dtype = dtype;
  out->location = DBD_HOST;
  out->ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nr + 1));
  out->idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
  out->val = malloc(dtype_size(dtype) * (size_t)std::max<int64_t>(nnz, 1));
  if (!out->ptr || !out->idx || !out->val) {
    dbd_csr_free(out);
    return set_error(DBD_ERR_NOMEM, "host allocation of a %lld x %lld CSR with %lld nonzeros failed",
                     (long long)nr, (long long)nc, (long long)nnz);
  }
  return DBD_OK;
}

inline void store_val(void* v, int32_t dt, int64_t i, double x) {
  if (dt == DBD_F64)
    ((double*)v)[i] = x;
  else
    ((float*)v)[i] = (float)x;
}
inline double load_val(const void* v, int32_t dt, int64_t i) {
  return dt == DBD_F64 ? ((const double*)v)[i] : (double)((const float*)v)[i];
}

// Build CSR from per-row generated column lists: rows[r] sorted distinct columns.
// `gen(r, cols, vals)` fills one row; rows are produced in parallel and then packed.
template <typename Gen>
dbd_status build_rows(int64_t nr, int64_t nc, int32_t dtype, dbd_csr* out, Gen&& gen) {
  std::vector<std::vector<int32_t>> cols(nr);
  std::vector<std::vector<double>> vals(nr);
  parallel_for(nr, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) gen(r, cols[r], vals[r]);
  });
  int64_t nnz = 0;
  for (int64_t r = 0; r < nr; ++r) nnz += (int64_t)cols[r].size();
  dbd_status st = alloc_csr(out, nr, nc, nnz, dtype);
  if (st != DBD_OK) return st;
  out->ptr[0] = 0;
  for (int64_t r = 0; r < nr; ++r) out->ptr[r + 1] = out->ptr[r] + (int64_t)cols[r].size();
  parallel_for(nr, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      int64_t o = out->ptr[r];
      for (size_t j = 0; j < cols[r].size(); ++j) {
        out->idx[o + (int64_t)j] = cols[r][j];
        store_val(out->val, dtype, o + (int64_t)j, vals[r][j]);
      }
      std::vector<int32_t>().swap(cols[r]);
      std::vector<double>().swap(vals[r]);
    }
  });
  return DBD_OK;
}

// Floyd's sampling of k distinct integers in [0, n), returned sorted.
void sample_distinct(Rng& rng, int64_t n, int64_t k, std::vector<int32_t>& out) {
  out.clear();
  if (k <= 0) return;
  if (k >= n) {
    out.resize(n);
    std::iota(out.begin(), out.end(), 0);
    return;
  }
  if (k <= 64) {
    for (int64_t j = n - k; j < n; ++j) {
      int32_t t = (int32_t)rng.below((uint64_t)j + 1);
      if (std::find(out.begin(), out.end(), t) != out.end()) t = (int32_t)j;
      out.push_back(t);
    }
  } else {
    std::unordered_set<int32_t> s;
    s.reserve((size_t)k * 2);
    for (int64_t j = n - k; j < n; ++j) {
      int32_t t = (int32_t)rng.below((uint64_t)j + 1);
      if (!s.insert(t).second) {
        t = (int32_t)j;
        s.insert(t);
      }
      out.push_back(t);
    }
  }
  std::sort(out.begin(), out.end());
}

double normal(Rng& rng) {
  double u1 = rng.uniform(), u2 = rng.uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Inverse-CDF table of a Zipf(s) law on [lo, hi].
struct Zipf {
  int64_t lo;
  std::vector<double> cdf;
  double mean = 0;
  Zipf(double s, int64_t lo_, int64_t hi) : lo(lo_) {
    double acc = 0, m = 0;
    cdf.resize((size_t)(hi - lo + 1));
    for (int64_t k = lo; k <= hi; ++k) {
      double w = pow((double)k, -s);
      acc += w;
      m += w * (double)k;
      cdf[(size_t)(k - lo)] = acc;
    }
    for (auto& c : cdf) c /= acc;
    mean = m / acc;
  }
  int64_t draw(Rng& rng) const {
    double u = rng.uniform();
    size_t i = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    if (i >= cdf.size()) i = cdf.size() - 1;
    return lo + (int64_t)i;
  }
};

}  // namespace

extern "C" {

dbd_status dbd_gen_er(int64_t nr, int64_t nc, int32_t deg, uint64_t seed, int32_t dtype, dbd_csr* out) {
  clear_error();
  if (!out || nr < 0 || nc <= 0 || deg < 0 || nc > INT32_MAX || nr > INT32_MAX)
    return set_error(DBD_ERR_INVALID, "dbd_gen_er: bad arguments");
  return build_rows(nr, nc, dtype, out, [&](int64_t r, std::vector<int32_t>& cols, std::vector<double>& vals) {
    Rng rng(seed, (uint64_t)r);
    sample_distinct(rng, nc, std::min<int64_t>(deg, nc), cols);
    vals.resize(cols.size());
    for (auto& v : vals) v = normal(rng);
  });
}

dbd_status dbd_gen_web(int64_t n, double zs, int32_t dmin, int32_t dmax, int64_t band, double p_local,
                       uint64_t seed, int32_t dtype, dbd_csr* out) {
  clear_error();
  if (!out || n <= 0 || n > INT32_MAX || dmin < 0 || dmax < dmin || band < 0)
    return set_error(DBD_ERR_INVALID, "dbd_gen_web: bad arguments");
  Zipf z(zs, std::max(dmin, 1), std::max(dmax, 1));
  return build_rows(n, n, dtype, out, [&](int64_t r, std::vector<int32_t>& cols, std::vector<double>& vals) {
    Rng rng(seed, (uint64_t)r);
    int64_t d = std::min<int64_t>(z.draw(rng), n);
    const int64_t lo = std::max<int64_t>(0, r - band), hi = std::min<int64_t>(n - 1, r + band);
    std::unordered_set<int32_t> seen;
    cols.clear();
    int64_t attempts = 0;
    while ((int64_t)cols.size() < d && attempts < 64 * d + 64) {
      ++attempts;
      int32_t t;
      if (rng.uniform() < p_local)
        t = (int32_t)(lo + (int64_t)rng.below((uint64_t)(hi - lo + 1)));
      else
        t = (int32_t)rng.below((uint64_t)n);
      if (seen.insert(t).second) cols.push_back(t);
    }
    std::sort(cols.begin(), cols.end());
    vals.resize(cols.size());
    for (auto& v : vals) v = rng.uniform();
  });
}

dbd_status dbd_gen_skewed(int64_t n, double mean_deg, double zs, uint64_t seed, int32_t dtype, dbd_csr* out) {
  clear_error();
  if (!out || n <= 0 || n > INT32_MAX || mean_deg <= 0)
    return set_error(DBD_ERR_INVALID, "dbd_gen_skewed: bad arguments");
  Zipf z(zs, 1, n);
  const double scale = mean_deg / z.mean;
  return build_rows(n, n, dtype, out, [&](int64_t r, std::vector<int32_t>& cols, std::vector<double>& vals) {
    Rng rng(seed, (uint64_t)r);
    int64_t d = (int64_t)llround((double)z.draw(rng) * scale);
    d = std::max<int64_t>(1, std::min<int64_t>(d, n));
    sample_distinct(rng, n, d, cols);
    vals.resize(cols.size());
    for (auto& v : vals) v = rng.uniform();
  });
}

// R-MAT without noise: each edge descends `scale` levels picking quadrant a / b / c / d.
dbd_status dbd_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                        int32_t add_identity, int32_t row_normalize, int32_t dtype, dbd_csr* out) {
  clear_error();
  if (!out || scale < 1 || scale > 30 || ef < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
    return set_error(DBD_ERR_INVALID, "dbd_gen_rmat: bad arguments");
  const int64_t n = (int64_t)1 << scale;
  const int64_t m = n * (int64_t)ef;
  const int64_t tot = m + (add_identity ? n : 0);
  std::vector<uint64_t> keys((size_t)tot);
  parallel_for(m, [&](int64_t b0, int64_t e0, int) {
    for (int64_t e = b0; e < e0; ++e) {
      Rng rng(seed, (uint64_t)e);
      uint64_t row = 0, col = 0;
      for (int l = 0; l < scale; ++l) {
        const double u = rng.uniform();
        row <<= 1;
        col <<= 1;
        if (u < a) {
        } else if (u < a + b) {
          col |= 1;
        } else if (u < a + b + c) {
          row |= 1;
        } else {
          row |= 1;
          col |= 1;
        }
      }
      keys[(size_t)e] = (row << 32) | col;
    }
  });
  if (add_identity)
    for (int64_t i = 0; i < n; ++i) keys[(size_t)(m + i)] = ((uint64_t)i << 32) | (uint64_t)i;
  // LSD radix sort on the 2*scale significant bits, 16 bits per pass
  {
    std::vector<uint64_t> tmp(keys.size());
    const int passes[] = {0, 16, 32, 48};
    for (int sh : passes) {
      const int bits = (sh < 32) ? std::min(16, scale - sh) : std::min(16, scale - (sh - 32));
      if (bits <= 0) continue;
      std::vector<int64_t> cnt(65537, 0);
      for (uint64_t k : keys) ++cnt[((k >> sh) & 0xFFFF) + 1];
      for (int i = 0; i < 65536; ++i) cnt[i + 1] += cnt[i];
      for (uint64_t k : keys) tmp[(size_t)cnt[(k >> sh) & 0xFFFF]++] = k;
      keys.swap(tmp);
    }
  }
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int64_t nnz = (int64_t)keys.size();
  dbd_status st = alloc_csr(out, n, n, nnz, dtype);
  if (st != DBD_OK) return st;
  std::fill(out->ptr, out->ptr + n + 1, 0);
  for (uint64_t k : keys) ++out->ptr[(k >> 32) + 1];
  for (int64_t i = 0; i < n; ++i) out->ptr[i + 1] += out->ptr[i];
  parallel_for(nnz, [&](int64_t b0, int64_t e0, int) {
    for (int64_t i = b0; i < e0; ++i) {
      const uint64_t k = keys[(size_t)i];
      out->idx[i] = (int32_t)(k & 0xFFFFFFFFu);
      if (!row_normalize) {
        const uint64_t h = splitmix64(seed ^ splitmix64(k));
        store_val(out->val, dtype, i, 1.0 - (double)(h >> 11) * (1.0 / 9007199254740992.0));
      }
    }
  });
  if (row_normalize) {
    parallel_for(n, [&](int64_t b0, int64_t e0, int) {
      for (int64_t r = b0; r < e0; ++r) {
        const int64_t len = out->ptr[r + 1] - out->ptr[r];
        for (int64_t i = out->ptr[r]; i < out->ptr[r + 1]; ++i) store_val(out->val, dtype, i, 1.0 / (double)len);
      }
    });
  }
  return DBD_OK;
}

// ---- file formats ---------------------------------------------------------------------

// The reference's cache (Z/readInput.cpp:166-175) holds FP32 values. FP64 reads keep their own
// cache, `<file>.csr64`: the same int32 header, row pointer and columns, then FP64 values, so an
// FP64 run never multiplies values that went through float.
static dbd_status write_cache(const char* path, const dbd_csr* M, bool f64) {
  if (!path || !M || M->location != DBD_HOST) return set_error(DBD_ERR_INVALID, "dbd_write_csr_cache: bad arguments");
  if (M->nrows >= INT32_MAX || M->ncols > INT32_MAX || M->nnz > INT32_MAX)
    return set_error(DBD_ERR_OVERFLOW, "the .csr cache format is int32; matrix too large");
  FILE* fp = fopen(path, "wb");
  if (!fp) return set_error(DBD_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  int32_t hdr[3] = {(int32_t)M->nrows, (int32_t)M->ncols, (int32_t)M->nnz};
  bool ok = fwrite(hdr, 4, 3, fp) == 3;
  std::vector<int32_t> p32((size_t)M->nrows + 1);
  for (int64_t i = 0; i <= M->nrows; ++i) p32[(size_t)i] = (int32_t)M->ptr[i];
  ok = ok && fwrite(p32.data(), 4, p32.size(), fp) == p32.size();
  ok = ok && fwrite(M->idx, 4, (size_t)M->nnz, fp) == (size_t)M->nnz;
  if (f64) {
    std::vector<double> v((size_t)M->nnz);
    for (int64_t i = 0; i < M->nnz; ++i) v[(size_t)i] = load_val(M->val, M->dtype, i);
    ok = ok && fwrite(v.data(), 8, v.size(), fp) == v.size();
  } else {
    std::vector<float> v32((size_t)M->nnz);
    for (int64_t i = 0; i < M->nnz; ++i) v32[(size_t)i] = (float)load_val(M->val, M->dtype, i);
    ok = ok && fwrite(v32.data(), 4, v32.size(), fp) == v32.size();
  }
  ok = (fclose(fp) == 0) && ok;
  if (!ok) return set_error(DBD_ERR_IO, "write to %s failed", path);
  return DBD_OK;
}

dbd_status dbd_write_csr_cache(const char* path, const dbd_csr* M) {
  clear_error();
  return write_cache(path, M, false);
}

static dbd_status read_cache(const char* path, int32_t dtype, bool f64, dbd_csr* out) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return set_error(DBD_ERR_IO, "cannot open %s", path);
  int32_t hdr[3];
  if (fread(hdr, 4, 3, fp) != 3 || hdr[0] < 0 || hdr[1] < 0 || hdr[2] < 0) {
    fclose(fp);
    return set_error(DBD_ERR_INVALID, "%s: bad .csr header", path);
  }
  dbd_status st = alloc_csr(out, hdr[0], hdr[1], hdr[2], dtype);
  if (st != DBD_OK) {
    fclose(fp);
    return st;
  }
  std::vector<int32_t> p32((size_t)hdr[0] + 1);
  std::vector<float> v32(f64 ? 0 : (size_t)hdr[2]);
  std::vector<double> v64(f64 ? (size_t)hdr[2] : 0);
  bool ok = fread(p32.data(), 4, p32.size(), fp) == p32.size();
  ok = ok && fread(out->idx, 4, (size_t)hdr[2], fp) == (size_t)hdr[2];
  ok = ok && (f64 ? fread(v64.data(), 8, v64.size(), fp) == v64.size()
                  : fread(v32.data(), 4, v32.size(), fp) == v32.size());
  fclose(fp);
  if (!ok) {
    dbd_csr_free(out);
    return set_error(DBD_ERR_INVALID, "%s: truncated .csr file", path);
  }
  for (size_t i = 0; i < p32.size(); ++i) out->ptr[i] = p32[i];
  for (int64_t i = 0; i < hdr[2]; ++i) store_val(out->val, dtype, i, f64 ? v64[(size_t)i] : (double)v32[(size_t)i]);
  if (out->ptr[0] != 0 || out->ptr[hdr[0]] != hdr[2]) {
    dbd_csr_free(out);
    return set_error(DBD_ERR_INVALID, "%s: inconsistent .csr row pointer", path);
  }
  return DBD_OK;
}

// make_CSR conventions (Z/readInput.cpp:48-176): the banner line is skipped unless the name
// contains ".dat"; "symmetric" in it mirrors off-diagonal entries; '%'/'#' lines are comments;
// dims are header+1 because indices are used as given (1-based files leave row/col 0 empty);
// a missing value is 1.0; fields split on tab, comma or space. Rows are ordered by a stable
// sort like the reference's radix sort; additionally columns are sorted within a row and
// duplicates summed, which the kernels require.
dbd_status dbd_read_matrix(const char* path, int32_t dtype, int32_t use_cache, dbd_csr* out) {
  clear_error();
  if (!path || !out) return set_error(DBD_ERR_INVALID, "dbd_read_matrix: bad arguments");
  std::string p(path);
  const bool f64 = dtype == DBD_F64;
  const std::string cache_path = p + (f64 ? ".csr64" : ".csr");
  if (use_cache && access(cache_path.c_str(), F_OK) == 0) return read_cache(cache_path.c_str(), dtype, f64, out);
  FILE* fp = fopen(path, "rt");
  if (!fp) return set_error(DBD_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  char line[4096];
  bool symmetric = false;
  if (p.find(".dat") == std::string::npos) {
    if (!fgets(line, sizeof(line), fp)) {
      fclose(fp);
      return set_error(DBD_ERR_INVALID, "Wrong Format.");
    }
    if (strstr(line, "symmetric")) symmetric = true;
  }
  const char* delims = "\t, \r\n";
  int64_t nr = -1, nc = -1, ne = -1;
  while (fgets(line, sizeof(line), fp)) {
    if (line[0] == '%' || line[0] == '#') continue;
    char* save = nullptr;
    char* t1 = strtok_r(line, delims, &save);
    if (!t1) continue;
    char* t2 = strtok_r(nullptr, delims, &save);
    char* t3 = t2 ? strtok_r(nullptr, delims, &save) : nullptr;
    if (!t2 || !t3) {
      fclose(fp);
      return set_error(DBD_ERR_INVALID, "Wrong Format.");
    }
    nr = atoll(t1) + 1;
    nc = atoll(t2) + 1;
    ne = atoll(t3);
    break;
  }
  if (nr <= 0 || nc <= 0 || ne < 0 || nr > INT32_MAX || nc > INT32_MAX) {
    fclose(fp);
    return set_error(DBD_ERR_INVALID, "Wrong Format.");
  }
  struct E {
    int32_t r, c;
    double v;
  };
  std::vector<E> es;
  es.reserve((size_t)(symmetric ? 2 * ne : ne));
  while (fgets(line, sizeof(line), fp)) {
    if (line[0] == '%' || line[0] == '#') continue;
    char* save = nullptr;
    char* t1 = strtok_r(line, delims, &save);
    if (!t1) continue;
    char* t2 = strtok_r(nullptr, delims, &save);
    if (!t2) {
      fclose(fp);
      return set_error(DBD_ERR_INVALID, "Wrong Format.");
    }
    char* t3 = strtok_r(nullptr, delims, &save);
    const int64_t r = atoll(t1), c = atoll(t2);
    const double v = t3 ? atof(t3) : 1.0;
    if (r < 0 || r >= nr || c < 0 || c >= nc) {
      fclose(fp);
      return set_error(DBD_ERR_INVALID, "%s: entry (%lld, %lld) outside %lld x %lld", path, (long long)r,
                       (long long)c, (long long)nr, (long long)nc);
    }
    es.push_back({(int32_t)r, (int32_t)c, v});
    if (symmetric && r != c) es.push_back({(int32_t)c, (int32_t)r, v});
  }
  fclose(fp);
  std::stable_sort(es.begin(), es.end(), [](const E& x, const E& y) { return x.r != y.r ? x.r < y.r : x.c < y.c; });
  // merge duplicates
  size_t w = 0;
  for (size_t i = 0; i < es.size(); ++i) {
    if (w > 0 && es[w - 1].r == es[i].r && es[w - 1].c == es[i].c)
      es[w - 1].v += es[i].v;
    else
      es[w++] = es[i];
  }
  es.resize(w);
  dbd_status st = alloc_csr(out, nr, nc, (int64_t)es.size(), dtype);
  if (st != DBD_OK) return st;
  std::fill(out->ptr, out->ptr + nr + 1, 0);
  for (const E& e : es) ++out->ptr[e.r + 1];
  for (int64_t i = 0; i < nr; ++i) out->ptr[i + 1] += out->ptr[i];
  for (size_t i = 0; i < es.size(); ++i) {
    out->idx[i] = es[i].c;
    store_val(out->val, dtype, (int64_t)i, es[i].v);
  }
  if (use_cache && out->nnz <= INT32_MAX) {
    write_cache(cache_path.c_str(), out, f64);  // best effort, like the reference (:166-175)
    clear_error();
  }
  return DBD_OK;
}

static bool write_all(int fd, const void* buf, size_t n, off_t off) {
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

// Ptr.output / Idx.output / Val.output (Z/writeFile.cpp:19-20, :177-197). Ptr.output is int32
// when nnz(C) < 2^31 (the reference's width) unless ptr64, else int64.
dbd_status dbd_write_outputs(const char* dir, const dbd_csr* C, int32_t ptr64) {
  clear_error();
  if (!C || C->location != DBD_HOST) return set_error(DBD_ERR_INVALID, "dbd_write_outputs: host CSR required");
  std::string d = (dir && dir[0]) ? std::string(dir) + "/" : std::string();
  const bool p64 = ptr64 || C->nnz > INT32_MAX;
  struct F {
    const char* name;
    const void* p;
    size_t n;
  };
  std::vector<int32_t> p32;
  const void* pp = C->ptr;
  size_t pbytes = sizeof(int64_t) * (size_t)(C->nrows + 1);
  if (!p64) {
    p32.resize((size_t)C->nrows + 1);
    for (int64_t i = 0; i <= C->nrows; ++i) p32[(size_t)i] = (int32_t)C->ptr[i];
    pp = p32.data();
    pbytes = 4 * p32.size();
  }
  F fs[3] = {{"Ptr.output", pp, pbytes},
             {"Idx.output", C->idx, 4 * (size_t)C->nnz},
             {"Val.output", C->val, dtype_size(C->dtype) * (size_t)C->nnz}};
  for (auto& f : fs) {
    std::string path = d + f.name;
    int fd = open(path.c_str(), O_CREAT | O_WRONLY | O_TRUNC, 0644);
    if (fd < 0) return set_error(DBD_ERR_IO, "cannot open %s: %s", path.c_str(), strerror(errno));
    bool ok = write_all(fd, f.p, f.n, 0);
    ok = (close(fd) == 0) && ok;
    if (!ok) return set_error(DBD_ERR_IO, "write to %s failed: %s", path.c_str(), strerror(errno));
  }
  return DBD_OK;
}

}  // extern "C"
