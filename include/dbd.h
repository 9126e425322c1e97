/*
 * dbd.h - C ABI of the MI355X-native Dynamic Block Distributor (CSR SpGEMM, C = A*B).
 *
 * Every entry point is plain C: opaque handle, plain pointers and sizes, a status code.
 * No torch or HIP types cross this boundary (the optional stream is passed as void*).
 *
 * Reference interfaces replaced (paths relative to the reference repository; "Z/" is the
 * member dbd/src/ of DBD_FULL.zip, "S/" is dbd/for_small_datasets/src/):
 *   - process globals + argv budgets          Z/main.cpp:20-51, :152-187      -> dbd_opts, dbd_create
 *   - read_CSR / make_CSR (.mtx, .csr cache)  Z/readInput.cpp:18-176         -> dbd_read_matrix, dbd_write_csr_cache
 *   - initialize_A / initialize_B (H2D)        S/preProc.cu:28-46             -> dbd_set_operands
 *   - get_UpperBound + divide_Blocks           S/preProc.cu:142-156, Z/main.cpp:259-310 -> dbd_plan
 *   - calculate_Criteria / analyze_Block       Z/distributeWorkload.cpp:249-270 -> dbd_plan_blocks, dbd_analyze_block
 *   - distribute_Workload + compute_GPU        Z/distributeWorkload.cpp:25-223, S/computeGPU.cu:15-98 -> dbd_run
 *   - write_File / write_Ptr                   Z/writeFile.cpp:13-197         -> dbd_run(DBD_MODE_IO), dbd_write_outputs
 *   - free_Resource / free_GPU                 Z/main.cpp:123-150, S/computeGPU.cu:101-132 -> dbd_destroy
 *
 * Errors: every function returns dbd_status; dbd_last_error() gives a thread-local message.
 * The library never calls exit() (the reference exits on bad argv, Z/main.cpp:177-184).
 *
 * Threading: a dbd_ctx is used by one caller thread at a time. Internal writer threads and
 * HIP streams are hidden. dbd_set_stream lets a framework (torch) supply its stream.
 */
#ifndef DBD_H
#define DBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DBD_ABI_VERSION 3

typedef enum dbd_status {
  DBD_OK = 0,
  DBD_ERR_INVALID = 1,  /* bad argument or malformed input ("Invalid Arguments", "Wrong Format") */
  DBD_ERR_HIP = 2,      /* a HIP runtime call failed */
  DBD_ERR_NOMEM = 3,    /* allocation failed ("Cannot Allocate ...") */
  DBD_ERR_BUDGET = 4,   /* host/device budget too small ("Not Enough Memory ...") */
  DBD_ERR_IO = 5,       /* file open/read/write failed */
  DBD_ERR_OVERFLOW = 6, /* index does not fit its type ("NE in Block Exceeds INTEGER Bound.") */
  DBD_ERR_STATE = 7,    /* call out of order (e.g. numeric before plan) */
  DBD_ERR_INTERNAL = 8  /* a device-side consistency check failed */
} dbd_status;

typedef enum dbd_dtype { DBD_F32 = 0, DBD_F64 = 1 } dbd_dtype;

/* The three reference binaries (Z/Makefile:17-30): bin/compute, bin/d2h, bin/SSpMM. */
typedef enum dbd_mode { DBD_MODE_COMPUTE = 0, DBD_MODE_D2H = 1, DBD_MODE_IO = 2 } dbd_mode;

typedef enum dbd_location { DBD_HOST = 0, DBD_DEVICE = 1 } dbd_location;

/* CSR matrix view. Replaces `struct csr` (S/SSpMM.hpp:36-44), widened to int64 row pointers
 * and FP32/FP64 values. Column indices within a row must be strictly increasing. */
typedef struct dbd_csr {
  int64_t nrows;
  int64_t ncols;
  int64_t nnz;
  int64_t* ptr;     /* nrows + 1 */
  int32_t* idx;     /* nnz */
  void* val;        /* nnz elements of `dtype` */
  int32_t dtype;    /* dbd_dtype */
  int32_t location; /* dbd_location */
} dbd_csr;

/* Options. Defaults (field = 0) follow the reference: 4096-row blocks (S/SSpMM.hpp:16-22),
 * 28 GiB host / 22 GiB device budgets (Z/main.cpp:153-155). */
typedef struct dbd_opts {
  int32_t block_rows;        /* ROWS_IN_BLOCK; a multiple of 256; 0 -> 4096 */
  int32_t dtype;             /* dbd_dtype of the product */
  int64_t host_budget_mib;   /* cpu_MemBound in MiB (GiB << 10); 0 -> 28 GiB */
  int64_t device_budget_mib; /* gpu_MemBound in MiB; 0 -> 22 GiB */
  int32_t tune;              /* 0: default category table; 1: autotuned table (cached per problem class) */
  int32_t verbose;           /* 1: print the reference's progress lines to stdout */
  const char* out_dir;       /* DBD_MODE_IO: directory of Ptr/Idx/Val.output (NULL -> cwd) */
  int32_t ptr64_file;        /* 1: always write Ptr.output as int64 (else int32 when nnz(C) < 2^31) */
  int32_t writer_threads;    /* pinned ring depth in DBD_MODE_IO is writer_threads + 1 chunk
                                slots (0 -> 4); each output file has one writer thread */
  int32_t chunks_hint;       /* minimum number of output chunks in D2H/IO modes; 0 -> 8 */
  int32_t retain_result;     /* 1: keep whole C (device in COMPUTE mode, host in D2H mode) if it fits */
  int32_t hybrid;            /* 1: CPU co-execution in D2H and IO modes (Z/distributeWorkload.cpp:69-144):
                                blocks whose analyze_Block outcome is 1 (CPU preferred) are multiplied
                                by host threads into the host result or the files, the GPU does the
                                rest. Needs host operand views that stay valid during dbd_run.
                                Ignored in COMPUTE mode (C is kept on the device there). 0: off */
  int32_t cpu_threads;       /* host threads of the hybrid mode (0 -> the hardware's thread count) */
} dbd_opts;

#define DBD_MAX_CATS 16

/* Result of planning. `products` is the reference's neC (Z/main.cpp:294); block fields mirror
 * block_Info / max_CalulationAmount / calculate_Criteria (Z/main.cpp:276-296,
 * Z/distributeWorkload.cpp:249-270). */
typedef struct dbd_plan_info {
  int64_t nrows, ncols, num_blocks, block_rows;
  int64_t products;           /* sum_i P_i, P_i = sum_{k in A_i} len(B_k) */
  int64_t max_block_products; /* max over blocks of sum of P_i */
  int64_t criteria;           /* reference GPU-only threshold (products per row) */
  int64_t blocks_gpu_only;    /* analyze_Block == 0 */
  int64_t blocks_cpu_pref;    /* analyze_Block == 1 */
  int64_t blocks_any;         /* analyze_Block == 2 */
  int64_t nnz;                /* nnz(C); -1 until the symbolic phase ran */
  int32_t n_sym_cats, n_num_cats;
  int64_t sym_rows[DBD_MAX_CATS]; /* rows per symbolic category (category 0 = empty rows) */
  int64_t num_rows[DBD_MAX_CATS]; /* rows per numeric category (-1 until symbolic ran) */
  int64_t sym_limit[DBD_MAX_CATS];/* upper key bound of each symbolic category */
  int64_t num_limit[DBD_MAX_CATS];/* upper key bound of each numeric category */
} dbd_plan_info;

/* Timings in seconds, labelled like the reference's report (Z/main.cpp:100-106). */
typedef struct dbd_timing {
  double preprocess_s;  /* "Preprocessing": row analysis + blocks + categories */
  double compute_gpu_s; /* "Compute GPU": symbolic + numeric kernels (device time) */
  double write_s;       /* "Write File": writer threads' busy time */
  double kernel_exec_s; /* "Kernel Execution": the distributed run after preprocessing */
  double total_s;       /* "Total Execution" = preprocess + kernel_exec */
  double h2d_s;         /* operand upload (excluded from the windows, as in Z/main.cpp:75-79) */
  double d2h_s;         /* time the copy engine spent on D2H of C */
  double d2h_bytes;     /* bytes copied D2H */
  double io_bytes;      /* bytes written to files */
  int64_t chunks;       /* output chunks used */
  double symbolic_s;    /* device time of the symbolic pass (row counts, scan, numeric bins) */
  double numeric_s;     /* device time from the end of the symbolic pass to the last numeric chunk */
  double cpu_s;         /* "Compute CPU": wall time of the hybrid mode's host threads */
  int64_t cpu_blocks;   /* blocks the host threads computed (hybrid mode) */
  int64_t cpu_products; /* their products (the CPU's share of the reference's "Elements") */
} dbd_timing;

typedef struct dbd_ctx dbd_ctx;

/* ---- library / context ------------------------------------------------------------ */
int32_t dbd_abi_version(void);
const char* dbd_last_error(void);
void dbd_default_opts(dbd_opts* opts);
dbd_status dbd_create(const dbd_opts* opts, dbd_ctx** out);
dbd_status dbd_destroy(dbd_ctx* ctx);
/* Use a caller-owned hipStream_t (passed as void*) for all device work. NULL is the HIP
 * default stream, where torch's default-stream work runs: pass it as is to stay ordered with it.
 * dbd_use_own_stream returns to a stream the context owns (the state after dbd_create). The
 * batched form's graph capture needs a stream other than the default one. */
dbd_status dbd_set_stream(dbd_ctx* ctx, void* hip_stream);
dbd_status dbd_use_own_stream(dbd_ctx* ctx);

/* ---- operands + plan + run (the reference pipeline) ------------------------------ */
/* Host views are uploaded (timed as h2d_s); device views are borrowed and must outlive use.
 * B may equal A (C = A*A, the reference's one-file argv form, Z/main.cpp:153-158). */
dbd_status dbd_set_operands(dbd_ctx* ctx, const dbd_csr* A, const dbd_csr* B);
/* Row analysis (P_i), int64 prefix, blocks, reference block criteria, symbolic categories. */
dbd_status dbd_plan(dbd_ctx* ctx, dbd_plan_info* info);
/* Whole pipeline in one of the three modes; plans first if needed. */
dbd_status dbd_run(dbd_ctx* ctx, int32_t mode, dbd_timing* timing);
/* Retained result: device CSR after COMPUTE (retain_result=1), host CSR after D2H.
 * The arrays stay owned by ctx until the next run or dbd_destroy. */
dbd_status dbd_result(dbd_ctx* ctx, dbd_csr* C);
/* Copy the retained result into caller host arrays (nrows+1 int64, nnz int32, nnz values; any
 * may be NULL to skip). Counterpart of the reference's result buffers being read by its writer
 * (Z/writeFile.cpp:58-154); here for callers that want C in their own memory. */
dbd_status dbd_copy_result(dbd_ctx* ctx, int64_t* host_ptr, int32_t* host_idx, void* host_val);
dbd_status dbd_get_plan_info(dbd_ctx* ctx, dbd_plan_info* info);
/* Per-row product counts P_i (host array of nrows) and block table (num_blocks + 1 entries
 * of block_ptr and num_blocks of block_products) from the last plan. */
dbd_status dbd_get_row_products(dbd_ctx* ctx, int64_t* host_P);
dbd_status dbd_get_blocks(dbd_ctx* ctx, int64_t* block_ptr, int64_t* block_products);

/* ---- device two-phase API for framework interop (torch) -------------------------- */
/* After dbd_plan: exact row pointer of C into caller memory (device, nrows+1 int64). */
dbd_status dbd_symbolic(dbd_ctx* ctx, int64_t* d_Cptr, int64_t* nnz_out);
/* Fill C's column indices and values (device) for the Cptr produced by dbd_symbolic. */
dbd_status dbd_numeric(dbd_ctx* ctx, const int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval);
/* Masked gradients: dA (nnz(A) values) = (dC * B^T) on pattern(A), dB (nnz(B) values) =
 * (A^T * dC) on pattern(B). Reuses the numeric categories of the forward plan. dA/dB may be
 * NULL to skip. dB is accumulated with atomics, so the caller zero-fills it. */
dbd_status dbd_backward(dbd_ctx* ctx, const int64_t* d_Cptr, const int32_t* d_Cidx, const void* d_dC,
                        void* d_dA, void* d_dB);

/* ---- batched form: many independent C_i = A_i * B_i in one hipGraph launch -------- */
/* The batch is given as block-diagonal concatenations (A_cat rows = all A_i rows, its column
 * indices already offset into B_cat's rows; B_cat column indices offset per matrix), so one
 * plan covers every member. Persistent category kernels read their work ranges from device
 * memory, so the whole pipeline is captured once and replayed with one hipGraphLaunch. */
/* dbd_batch_prepare plans the batch and runs its symbolic pass once: *nnz_out is the exact
 * nnz(C_cat), the number of column indices and values the caller allocates. dbd_batch_launch
 * replays the captured pipeline into Cptr (nrows+1), Cidx and Cval (nnz_out entries each); if the
 * operands' structure changed so that C would need more, nothing is written and DBD_ERR_BUDGET
 * is returned. */
dbd_status dbd_batch_prepare(dbd_ctx* ctx, const dbd_csr* A_cat, const dbd_csr* B_cat, int64_t* nnz_out);
dbd_status dbd_batch_launch(dbd_ctx* ctx, int64_t* d_Cptr, int32_t* d_Cidx, void* d_Cval);
/* Autotuner: search the candidate category tables on the operands set in ctx (coordinate
 * descent over the knobs of dbd_table_desc from table 0, timing symbolic + numeric of each table
 * it visits) and keep the fastest for this problem class (cached in-process). Returns the chosen
 * table's id. */
dbd_status dbd_autotune(dbd_ctx* ctx, int32_t* chosen);
/* The knobs a candidate table is built from. Table 0 is the default table. */
typedef struct dbd_table_desc {
  int32_t g_small;      /* lanes per B row in hash tables of up to 1024 slots (16 or 64) */
  int32_t g_big;        /* lanes per B row in larger tables (16 or 64) */
  double load;          /* hash-table load limit: a table of T slots takes rows of <= load*T keys */
  int32_t top;          /* slots of the largest numeric LDS table (4096 or 8192; symbolic 2x);
                           rows beyond it take the long-row windows */
  int64_t hub_min_cols; /* long rows of more than 1024 A entries take the numeric hub range
                           kernels when B has at least this many columns (0: always, the
                           default; 2^21: wide problems only; INT64_MAX: never) */
} dbd_table_desc;
/* Candidate category tables (every combination of the knobs above: 48) the autotuner chooses
 * from; dbd_set_table selects one explicitly (results do not depend on the choice). */
int32_t dbd_table_count(void);
dbd_status dbd_table_knobs(int32_t table_id, dbd_table_desc* out);
dbd_status dbd_set_table(dbd_ctx* ctx, int32_t table_id);

/* ---- entry point that allocates nothing ------------------------------------------- */
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): C = A*B for device views
 * A and B (B NULL -> A*A) on the caller's stream (NULL: the default stream), with the whole
 * pipeline's workspace in one fixed array of DBD_NOALLOC_SCRATCH_BYTES of device memory that the
 * library holds on each GPU it runs on until the process ends; no hipMalloc, hipFree or other
 * allocating call is made. The caller passes Cptr (nrows+1 int64) and Cidx / Cval with
 * `capacity` entries; *nnz_out receives nnz(C). If nnz(C) > capacity nothing is written to
 * Cidx / Cval and DBD_ERR_BUDGET is returned (*nnz_out says how much is needed). A product whose
 * workspace exceeds the array (dbd_noalloc_scratch_bytes(nrows, nnz(A)); about 26 bytes per row
 * and 28 per nonzero of A) is refused with DBD_ERR_BUDGET before any device work. Calls share
 * their GPU's array: calls that overlap on one GPU race. Synchronises the stream before returning. */
#define DBD_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
size_t dbd_noalloc_scratch_bytes(int64_t nrows, int64_t nnz_a);
dbd_status dbd_spgemm_noalloc(const dbd_csr* A, const dbd_csr* B, void* hip_stream, int64_t* d_Cptr,
                              int32_t* d_Cidx, void* d_Cval, int64_t capacity, int64_t* nnz_out);

/* ---- host-only logic (no device needed) ------------------------------------------ */
/* Blocks from per-row products (divide_Blocks, Z/main.cpp:259-296) and reference criteria
 * (Z/distributeWorkload.cpp:249-270). block_ptr has num_blocks+1 entries, block_products
 * num_blocks; info->criteria / blocks_* are filled. */
dbd_status dbd_plan_blocks(const int64_t* row_products, int64_t nrows, int32_t block_rows,
                           int64_t* block_ptr, int64_t* block_products, dbd_plan_info* info);
/* analyze_Block (Z/distributeWorkload.cpp:266-270): 0 GPU-only, 1 CPU-preferred, 2 any. */
int32_t dbd_analyze_block(int64_t criteria, int64_t avg_products_per_row);
/* Read .mtx / .dat text (Z/readInput.cpp:48-176 conventions) or a <file>.csr cache; the
 * result is host memory owned by the library (free with dbd_csr_free). */
dbd_status dbd_read_matrix(const char* path, int32_t dtype, int32_t use_cache, dbd_csr* out);
/* Write the reference's .csr cache format (int32 nr, nc, ne; int32 ptr; int32 idx; f32 val). */
dbd_status dbd_write_csr_cache(const char* path, const dbd_csr* M);
dbd_status dbd_csr_free(dbd_csr* M);
/* Write Ptr.output / Idx.output / Val.output for a host CSR (Z/writeFile.cpp formats). */
dbd_status dbd_write_outputs(const char* dir, const dbd_csr* C, int32_t ptr64);

/* ---- seeded synthetic generators (host, deterministic for any thread count) -------- */
/* Erdos-Renyi: exactly `deg` distinct uniform columns per row, values N(0,1). */
dbd_status dbd_gen_er(int64_t nrows, int64_t ncols, int32_t deg, uint64_t seed, int32_t dtype, dbd_csr* out);
/* Web-like graph: out-degree Zipf(s) truncated to [dmin, dmax]; each target within +-band of
 * the row with probability p_local, else uniform; distinct per row; values U[0,1). */
dbd_status dbd_gen_web(int64_t n, double zipf_s, int32_t dmin, int32_t dmax, int64_t band, double p_local,
                       uint64_t seed, int32_t dtype, dbd_csr* out);
/* R-MAT scale/edge factor/(a,b,c), duplicates merged, values U(0,1]; add_identity adds the
 * diagonal; row_normalize divides each row by its sum. */
dbd_status dbd_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                        int32_t add_identity, int32_t row_normalize, int32_t dtype, dbd_csr* out);
/* Square matrix with rows whose lengths follow Zipf(s) scaled to mean `mean_deg` (>=1),
 * uniform distinct columns, values U(0,1). */
dbd_status dbd_gen_skewed(int64_t n, double mean_deg, double zipf_s, uint64_t seed, int32_t dtype, dbd_csr* out);

#ifdef __cplusplus
}
#endif
#endif /* DBD_H */
