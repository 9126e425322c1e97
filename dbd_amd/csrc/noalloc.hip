// Entry point that allocates nothing: the batched form's capturable pipeline (enqueue_pipeline)
// with all of its workspace in one fixed array of device memory that the library holds per GPU,
// for callers that cannot allocate device memory (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc). No counterpart in the reference.
//
// Calls share their GPU's array and it keeps the last call's contents; the pipeline writes every
// word it reads (the error word is cleared by a fill kernel first), so nothing is carried over.
// Calls that overlap on one GPU (two streams, two threads) race on the array.
#include <mutex>

#include "common.h"
#include "kernels.h"
#include "pipeline.h"

using namespace dbd;

constexpr size_t kNoAllocScratchBytes = DBD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device. The array's
// name in host code is only a stand-in for the symbol; every host use goes through this address.
dbd_status scratch_base(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(DBD_ERR_INVALID, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(DBD_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                       hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return DBD_OK;
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

// Workspace bytes of enqueue_pipeline for nr rows and nnzA nonzeros of A, carved from `base`
// when it is not null.
size_t carve_work(int64_t nr, int64_t nnzA, unsigned char* base, Work* w) {
  const int64_t nwg = bin_tiles(nr);
  const size_t ncnt = (size_t)kMaxCats * (size_t)nwg + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + off : nullptr;
    off += round256(bytes);
    return p;
  };
  unsigned char* P = take(8 * (size_t)(nr + 1));
  unsigned char* nnz = take(8 * (size_t)(nr + 1));
  unsigned char* rows_sym = take(4 * (size_t)(nr + 1));
  unsigned char* rows_num = take(4 * (size_t)(nr + 1));
  unsigned char* cnt = take(8 * ncnt);
  unsigned char* offs_sym = take(8 * (ncnt + 1));
  unsigned char* offs_num = take(8 * (ncnt + 1));
  unsigned char* tmp = take(8 * (size_t)scan_tmp_elems((int64_t)ncnt + nr + 1));
  unsigned char* bstart = take(8 * (size_t)(nnzA + 1));
  unsigned char* blen = take(4 * (size_t)(nnzA + 1));
  unsigned char* cursor = take(16 * (size_t)(nnzA + 1));
  unsigned char* err = take(64);  // error word, then the sink degenerate rows are written to
  if (w) {
    w->P = (int64_t*)P;
    w->nnz = (int64_t*)nnz;
    w->rows_sym = (int32_t*)rows_sym;
    w->rows_num = (int32_t*)rows_num;
    w->cnt = (int64_t*)cnt;
    w->offs_sym = (int64_t*)offs_sym;
    w->offs_num = (int64_t*)offs_num;
    w->tmp = (int64_t*)tmp;
    w->bstart = (int64_t*)bstart;
    w->blen = (int32_t*)blen;
    w->cursor = (int64_t*)cursor;
    w->err = (int*)err;
    w->nr = nr;
    w->nwg = nwg;
  }
  return off;
}

const Tables& noalloc_tables() {
  static const Tables t = default_tables(16, 64, 0.5);
  return t;
}
}  // namespace

extern "C" {

size_t dbd_noalloc_scratch_bytes(int64_t nrows, int64_t nnz_a) {
  if (nrows < 0 || nnz_a < 0) return 0;
  return carve_work(nrows, nnz_a, nullptr, nullptr);
}

dbd_status dbd_spgemm_noalloc(const dbd_csr* A, const dbd_csr* B, void* hip_stream, int64_t* d_Cptr, int32_t* d_Cidx,
                              void* d_Cval, int64_t capacity, int64_t* nnz_out) {
  clear_error();
  if (!A || !d_Cptr || capacity < 0) return set_error(DBD_ERR_INVALID, "dbd_spgemm_noalloc: bad arguments");
  if (!B) B = A;
  if (A->location != DBD_DEVICE || B->location != DBD_DEVICE)
    return set_error(DBD_ERR_INVALID, "dbd_spgemm_noalloc: operands must be device views");
  if (A->dtype != B->dtype || (A->dtype != DBD_F32 && A->dtype != DBD_F64))
    return set_error(DBD_ERR_INVALID, "dbd_spgemm_noalloc: A and B must share dtype DBD_F32 or DBD_F64");
  if (A->ncols != B->nrows)
    return set_error(DBD_ERR_INVALID, "inner dimensions differ: A is %lld x %lld, B is %lld x %lld",
                     (long long)A->nrows, (long long)A->ncols, (long long)B->nrows, (long long)B->ncols);
  if (A->nrows < 0 || A->nrows >= INT32_MAX || B->ncols > INT32_MAX || A->nnz < 0 || B->nnz < 0)
    return set_error(DBD_ERR_OVERFLOW, "matrix dimensions must be below 2^31");
  if (capacity > 0 && (!d_Cidx || !d_Cval)) return set_error(DBD_ERR_INVALID, "null output arrays");
  const size_t need = carve_work(A->nrows, A->nnz, nullptr, nullptr);
  if (need > kNoAllocScratchBytes)
    return set_error(DBD_ERR_BUDGET,
                     "dbd_spgemm_noalloc: the product needs %zu bytes of workspace, more than the fixed %zu "
                     "(26 bytes per row of A and 28 per nonzero of A, roughly); use the context API",
                     need, kNoAllocScratchBytes);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  CHK(scratch_base(dev, &base));
  int num_cu = 256;
  if (hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || num_cu <= 0)
    num_cu = 256;
  Work w;
  carve_work(A->nrows, A->nnz, base, &w);
  Mat a{}, b{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.nnz = A->nnz;
  a.ptr = A->ptr;
  a.idx = A->idx;
  a.val = A->val;
  b.nrows = B->nrows;
  b.ncols = B->ncols;
  b.nnz = B->nnz;
  b.ptr = B->ptr;
  b.idx = B->idx;
  b.val = B->val;
  hipStream_t s = (hipStream_t)hip_stream;
  if (A->nrows == 0) {
    launch_fill_i64(d_Cptr, 1, 0, s);
    if (nnz_out) *nnz_out = 0;
    HIPCHK(hipStreamSynchronize(s));
    return DBD_OK;
  }
  enqueue_pipeline(w, noalloc_tables(), a, b, A->dtype, num_cu, d_Cptr, d_Cidx, d_Cval, capacity, s);
  HIPCHK(hipGetLastError());
  int64_t nnz = 0;
  int e = 0;
  HIPCHK(hipMemcpyAsync(&nnz, d_Cptr + A->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&e, w.err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (nnz_out) *nnz_out = nnz;
  if (e & kErrCapacityBit)
    return set_error(DBD_ERR_BUDGET, "dbd_spgemm_noalloc: C has %lld entries, capacity %lld; nothing was written",
                     (long long)nnz, (long long)capacity);
  if (e != 0) return set_error(DBD_ERR_INTERNAL, "dbd_spgemm_noalloc: device consistency check failed (flags 0x%x)", e);
  return DBD_OK;
}

}  // extern "C"
