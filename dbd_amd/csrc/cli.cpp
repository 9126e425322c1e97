// Drop-in command line of the reference binaries (Z/Makefile:17-30): one source built four
// times. DBD_CLI_MODE 0 -> compute (kernel only), 1 -> d2h (kernel + D2H), 2 -> SSpMM
// (kernel + D2H + file I/O); DBD_CLI_DEBUG adds the statistics of the reference's DEBUG build
// (Z/main.cpp:110-117).
//
// argv forms follow Z/main.cpp:152-184:
//   <exe> A | <exe> A B | <exe> cpuGiB gpuGiB A | <exe> cpuGiB gpuGiB A B
// Budgets are GiB, stored as MiB (<< 10). Optional flags may precede them:
//   --block-rows N   rows per block (reference ROWS_IN_BLOCK; default 4096)
//   --fp64 | --fp32  value type (default fp32, the reference's)
//   --out-dir DIR    where Ptr/Idx/Val.output go (default: current directory)
//   --ptr64          always write Ptr.output as int64
//   --tune           autotune category tables for this input
//   --writers N      file writer threads
//   --no-cache       do not read or write the <input>.csr cache
//   --hybrid         CPU co-execution: CPU-preferred blocks on host threads (d2h, SSpMM)
//   --cpu-threads N  host threads of --hybrid (default: all)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/dbd.h"

#ifndef DBD_CLI_MODE
#define DBD_CLI_MODE 2
#endif

static void usage() {
  puts("Invalid Arguments.");
  puts("Usage:\t<1> Executable File Name, Input File 1");
  puts("\t<2> Executable File Name, Input File 1, Input File 2");
  puts("\t<3> Executable File Name, CPU DRAM Size, GPU DRAM Size, Input File 1");
  puts("\t<4> Executable File Name, CPU DRAM Size, GPU DRAM Size, Input File 1, Input File 2\n");
  exit(-1);
}

static int fail(const char* what) {
  fprintf(stderr, "%s: %s\n", what, dbd_last_error());
  return 1;
}

int main(int argc, char** argv) {
  puts("\nLarge-Scaled Sparse Matrix - Sparse Matrix Multiplication. (L3SpMM)\n");
  dbd_opts o;
  dbd_default_opts(&o);
  o.dtype = DBD_F32;
  int use_cache = 1, tune = 0;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--block-rows" && i + 1 < argc) {
      o.block_rows = atoi(argv[++i]);
    } else if (a == "--fp64") {
      o.dtype = DBD_F64;
    } else if (a == "--fp32") {
      o.dtype = DBD_F32;
    } else if (a == "--out-dir" && i + 1 < argc) {
      o.out_dir = argv[++i];
    } else if (a == "--ptr64") {
      o.ptr64_file = 1;
    } else if (a == "--tune") {
      tune = 1;
    } else if (a == "--writers" && i + 1 < argc) {
      o.writer_threads = atoi(argv[++i]);
    } else if (a == "--no-cache") {
      use_cache = 0;
    } else if (a == "--hybrid") {
      o.hybrid = 1;
    } else if (a == "--cpu-threads" && i + 1 < argc) {
      o.cpu_threads = atoi(argv[++i]);
    } else if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
      usage();
    } else {
      pos.push_back(a);
    }
  }
  std::string fa, fb;
  if (pos.size() == 1) {
    fa = fb = pos[0];
  } else if (pos.size() == 2) {
    fa = pos[0];
    fb = pos[1];
  } else if (pos.size() == 3 || pos.size() == 4) {
    o.host_budget_mib = (int64_t)atoi(pos[0].c_str()) << 10;
    o.device_budget_mib = (int64_t)atoi(pos[1].c_str()) << 10;
    fa = pos[2];
    fb = pos.size() == 4 ? pos[3] : pos[2];
  } else {
    usage();
  }
  if (o.host_budget_mib <= 0 || o.device_budget_mib <= 0) usage();
  std::string pre = o.out_dir ? std::string(o.out_dir) + "/" : std::string();
  remove((pre + "Ptr.output").c_str());
  remove((pre + "Idx.output").c_str());
  remove((pre + "Val.output").c_str());

  dbd_csr A{}, B{};
  if (dbd_read_matrix(fa.c_str(), o.dtype, use_cache, &A) != DBD_OK) return fail("read A");
  printf("Input: %s\n\tNumber of Rows:\t\t%lld\n\tNumber of Columns:\t%lld\n\tNumber of Edges:\t%lld\n\n", fa.c_str(),
         (long long)A.nrows, (long long)A.ncols, (long long)A.nnz);
  const bool same = (fa == fb);
  if (!same) {
    if (dbd_read_matrix(fb.c_str(), o.dtype, use_cache, &B) != DBD_OK) return fail("read B");
    printf("Input: %s\n\tNumber of Rows:\t\t%lld\n\tNumber of Columns:\t%lld\n\tNumber of Edges:\t%lld\n\n",
           fb.c_str(), (long long)B.nrows, (long long)B.ncols, (long long)B.nnz);
  }
  dbd_ctx* ctx = nullptr;
  if (dbd_create(&o, &ctx) != DBD_OK) return fail("create");
  if (dbd_set_operands(ctx, &A, same ? &A : &B) != DBD_OK) return fail("operands");
  if (tune) {
    int32_t id = -1;
    if (dbd_autotune(ctx, &id) != DBD_OK) return fail("autotune");
    printf("Autotuned category table: %d\n", id);
  }
  dbd_plan_info pi{};
  if (dbd_plan(ctx, &pi) != DBD_OK) return fail("plan");
  printf("Divide Matrix A into %lld Blocks (Block Size: %lld)\n\n", (long long)pi.num_blocks, (long long)pi.block_rows);
#ifdef DBD_CLI_DEBUG
  // block analysis before the run, as distribute_Workload prints it (Z/distributeWorkload.cpp:35-44)
  printf("Top 10%% (Higher than %lld): %lld\n", (long long)pi.criteria, (long long)pi.blocks_gpu_only);
  printf("Lower than Two Warp Size (64): %lld\n\n", (long long)pi.blocks_cpu_pref);
#endif
  dbd_timing t{};
  if (dbd_run(ctx, DBD_CLI_MODE, &t) != DBD_OK) return fail("run");
  if (dbd_get_plan_info(ctx, &pi) != DBD_OK) return fail("plan info");
  printf("------------ AVG TIME ------------\n");
  printf("Preprocessing\t\t%f s\n", t.preprocess_s);
  printf("Compute CPU\t\t%f s\n", t.cpu_s);
  printf("Compute GPU\t\t%f s\n", t.compute_gpu_s);
  printf("Write File\t\t%f s\n", t.write_s);
  printf("Kernel Execution\t%f s\n", t.kernel_exec_s);
  printf("Total Execution\t\t%f s\n", t.total_s);
  const long long nbk = (long long)pi.num_blocks, np_ = (long long)pi.products;
  const long long cb = (long long)t.cpu_blocks, cpr = (long long)t.cpu_products;
  printf(" - Workload Distribution -\nCPUs\tBlocks:   %lld (%.2f %%)\n\tElements: %lld (%.2f %%)\n", cb,
         nbk ? 100.0 * (double)cb / (double)nbk : 0.0, cpr, np_ ? 100.0 * (double)cpr / (double)np_ : 0.0);
  printf("GPUs\tBlocks:   %lld (%.2f %%)\n\tElements: %lld (%.2f %%)\n\n", nbk - cb,
         nbk ? 100.0 * (double)(nbk - cb) / (double)nbk : 0.0, np_ - cpr,
         np_ ? 100.0 * (double)(np_ - cpr) / (double)np_ : 0.0);
#ifdef DBD_CLI_DEBUG
  // Z/main.cpp:110-116. Blocks reach the files one chunk per write (each chunk is a run of
  // consecutive blocks written with one pwrite per file, the I/O-binding writer's coalescing,
  // Z/writeFile_binding.cpp:60-168), so the blocks merged into another block's write are
  // num_blocks - chunks.
  puts("\n - DEBUG -");
  // (host blocks of the hybrid mode are written one block per write)
  const long long merged = t.chunks > 0 ? (long long)pi.num_blocks - (long long)t.cpu_blocks - (long long)t.chunks : 0;
  printf("Merged %lld Blocks while Writing (%f)\n", merged,
         pi.num_blocks > 0 ? (double)merged / (double)pi.num_blocks : 0.0);
  printf("Computed Total %lld Elements (%lld Bytes = %lld MebiBytes)\n", (long long)pi.products,
         (long long)pi.products * 4, (long long)(pi.products >> 18));
  printf("Actual Total %lld Elements (%lld Bytes = %lld MebiBytes)\n", (long long)pi.nnz, (long long)pi.nnz * 4,
         (long long)(pi.nnz >> 18));
  printf("Reduced to %.2f %% After Merge\n\n", pi.products ? 100.0 * (double)pi.nnz / (double)pi.products : 0.0);
#endif
  const double gf = t.total_s > 0 ? 2.0 * (double)pi.products / t.total_s * 1e-9 : 0.0;
  printf("{\"mode\": %d, \"products\": %lld, \"nnz\": %lld, \"total_s\": %.6f, \"gflops\": %.3f, \"d2h_bytes\": %.0f, "
         "\"io_bytes\": %.0f}\n",
         DBD_CLI_MODE, (long long)pi.products, (long long)pi.nnz, t.total_s, gf, t.d2h_bytes, t.io_bytes);
  dbd_destroy(ctx);
  dbd_csr_free(&A);
  if (!same) dbd_csr_free(&B);
  return 0;
}
