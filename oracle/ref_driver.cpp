// Test driver around the reference's own CPU executor (compute_CPU, member dbd/src/computeCPU.cpp
// of DBD_FULL.zip), compiled unmodified with -DCOMPUTE. TEST INFRASTRUCTURE ONLY (checks the
// oracle); it defines the process globals that the reference's main.cpp / distributeWorkload.cpp
// would define (S/SSpMM.hpp:69-124) and drives the executor thread block by block through its
// threadArgs handshake (Z/distributeWorkload.cpp:145-188).
//
// The executor keeps the reference semantics: FP32 values, first touch detected by
// |acc| < 1e-4 (Z/computeCPU.cpp:48), columns in first-touch order, results per row at
// upper-bound offsets within the block (Z/computeCPU.cpp:38-40). The driver copies each row out
// in that order; the test sorts rows before comparing.
#include <pthread.h>
#include <string.h>
#include <unistd.h>

#include <atomic>
#include <vector>

#include "SSpMM.hpp"

// globals of S/SSpMM.hpp used by compute_CPU in COMPUTE mode
float elapsed_Time[7] = {0};
csr A, B;
int* h_Upp = nullptr;
int* block_Ptr = nullptr;
int* block_Info = nullptr;
int* ptrC = nullptr;
int* h_IdxC = nullptr;
float* h_ValC = nullptr;
float* cpu_Dense = nullptr;
threadArgs* cpu_Args = nullptr;
long int cpu_BufLoc = 0;
// declared in the header, not used by compute_CPU in COMPUTE mode
int *cpu_BufIdx1 = nullptr, *gpu_BufIdx1 = nullptr, *cpu_BufIdx2 = nullptr, *gpu_BufIdx2 = nullptr;
float *cpu_BufVal1 = nullptr, *gpu_BufVal1 = nullptr, *cpu_BufVal2 = nullptr, *gpu_BufVal2 = nullptr;
int* cpu_Ptr_for_Collecting = nullptr;
storedBlock* block_Queue = nullptr;

static volatile bool* working_flag() { return (volatile bool*)&cpu_Args->working; }

// C = A*B with the reference CPU executor, `rows_in_block` rows per block.
// Outputs: cnt[nr] (ptrC counts), idx/val (per row, first-touch order) at out_off[row].
extern "C" int dbdref_spgemm(int nr, int nc_a, int nc_b, const int* Aptr, const int* Aidx, const float* Aval,
                             const int* Bptr, const int* Bidx, const float* Bval, int rows_in_block, int* cnt,
                             long long* out_off, int* out_idx, float* out_val) {
  A.nr = nr;
  A.nc = nc_a;
  A.ne = Aptr[nr];
  A.ptr = (int*)Aptr;
  A.idx = (int*)Aidx;
  A.val = (float*)Aval;
  B.nr = nc_a;
  B.nc = nc_b;
  B.ne = Bptr[nc_a];
  B.ptr = (int*)Bptr;
  B.idx = (int*)Bidx;
  B.val = (float*)Bval;
  // the executor indexes its dense accumulator with A.nc (Z/computeCPU.cpp:35), so size it
  // for B's columns: the reference only runs square A*A or A.nc >= B.nc inputs this way
  const int nc = nc_a > nc_b ? nc_a : nc_b;
  A.nc = nc;
  std::vector<int> upp(nr + 1, 0);
  for (int i = 0; i < nr; ++i) {
    long long s = 0;
    for (int a = Aptr[i]; a < Aptr[i + 1]; ++a) s += Bptr[Aidx[a] + 1] - Bptr[Aidx[a]];
    upp[i + 1] = upp[i] + (int)s;
  }
  h_Upp = upp.data();
  const int nb = (nr - 1) / rows_in_block + 1;
  std::vector<int> bptr(nb + 1);
  int maxb = 1;
  for (int b = 0; b <= nb; ++b) bptr[b] = b * rows_in_block < nr ? b * rows_in_block : nr;
  for (int b = 0; b < nb; ++b) maxb = std::max(maxb, upp[bptr[b + 1]] - upp[bptr[b]]);
  block_Ptr = bptr.data();
  std::vector<int> pc(nr + 1, 0);
  ptrC = pc.data();
  std::vector<int> ti(maxb);
  std::vector<float> tv(maxb);
  h_IdxC = ti.data();
  h_ValC = tv.data();
  std::vector<float> dense((size_t)nc * CPU_DENSE, 0.0f);
  cpu_Dense = dense.data();
  threadArgs args = {0, false, true, false};
  cpu_Args = &args;
  pthread_t th;
  if (pthread_create(&th, nullptr, compute_CPU, nullptr) != 0) return -1;
  for (int b = 0; b < nb; ++b) {
    cpu_Args->block_Index = b;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    *working_flag() = true;
    while (*working_flag()) usleep(50);
    std::atomic_thread_fence(std::memory_order_seq_cst);
    const int r0 = bptr[b];
    for (int r = r0; r < bptr[b + 1]; ++r) {
      const int n = pc[r + 1];
      const int src = upp[r] - upp[r0];
      cnt[r] = n;
      memcpy(out_idx + out_off[r], ti.data() + src, sizeof(int) * n);
      memcpy(out_val + out_off[r], tv.data() + src, sizeof(float) * n);
    }
  }
  cpu_Args->finished = true;
  pthread_join(th, nullptr);
  cpu_Args = nullptr;
  return 0;
}
