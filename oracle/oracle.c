/*
 * oracle.c - CPU restatement of the reference DBD SpGEMM path. TEST INFRASTRUCTURE ONLY:
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker / CPU baseline. The product (libdbd) never links or calls it.
 *
 * Reference citations (Z/ = member dbd/src/ of DBD_FULL.zip; S/ = dbd/for_small_datasets/src/):
 *   oracle_row_products     S/preProc.cu:170-179 (calculate_UpperBound) + :149-154 (prefix)
 *   oracle_blocks           Z/main.cpp:259-296 (divide_Blocks)
 *   oracle_criteria         Z/distributeWorkload.cpp:249-264 (calculate_Criteria)
 *   oracle_analyze_block    Z/distributeWorkload.cpp:266-270 (analyze_Block)
 *   oracle_spgemm_count/fill Z/computeCPU.cpp:32-63 (row-wise Gustavson with a dense
 *                           accumulator per thread), mkl/mkl.cpp:73-92 (C = A*B exported with
 *                           rows ordered, mkl_sparse_order :81), and the textbook SMMP of
 *                           library/cusplibrary/cusp/system/detail/sequential/multiply/csr_spgemm.h:39-153.
 *   oracle_backward         no reference counterpart (the masked gradients of C = A*B).
 *
 * Semantics (deliberate, documented differences from the reference's own CPU path):
 *   - FP64 accumulation; first touch is tracked by a marker array, never by |acc| < 1e-4
 *     (the reference heuristic, Z/computeCPU.cpp:48, which duplicates columns whose partial
 *     sum is near zero);
 *   - structural output: every (i, j) reached by a product is kept, numeric zeros included
 *     (CUSP's pass 2 drops exact zeros, csr_spgemm.h:135 - not restated);
 *   - columns sorted ascending in every row; int64 row pointers;
 *   - products are accumulated in the order k ascending within A_i, then B_k's order; the
 *     multiply is val(A) * val(B) (compiled with -ffp-contract=off: no fused multiply-add).
 *
 * Parity pinning of each function is done in tests/test_oracle.py (known-answer tests,
 * closed forms, brute force, scipy); see the test docstrings.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int nthreads_or_default(int t) { return t > 0 ? t : omp_get_max_threads(); }

/* calculate_UpperBound: P_i = sum over k in A_i of len(B_k). */
void oracle_row_products(int64_t nr, const int64_t* Aptr, const int32_t* Aidx, const int64_t* Bptr, int64_t* P) {
  for (int64_t i = 0; i < nr; ++i) {
    int64_t s = 0;
    for (int64_t a = Aptr[i]; a < Aptr[i + 1]; ++a) {
      const int32_t k = Aidx[a];
      s += Bptr[k + 1] - Bptr[k];
    }
    P[i] = s;
  }
}

/* divide_Blocks: blocks of `br` rows, block_info[b] = sum of P over the block. Returns the
 * number of blocks ((nr - 1) / br + 1, as Z/main.cpp:265). */
int64_t oracle_blocks(const int64_t* P, int64_t nr, int32_t br, int64_t* block_info) {
  if (nr <= 0) return 0;
  const int64_t nb = (nr - 1) / br + 1;
  for (int64_t b = 0; b < nb; ++b) {
    int64_t s = 0;
    const int64_t r1 = (b + 1) * br < nr ? (b + 1) * br : nr;
    for (int64_t r = b * br; r < r1; ++r) s += P[r];
    block_info[b] = s;
  }
  return nb;
}

/* calculate_Criteria: avg = (neC / nb) / br (integer); sd = sqrt(mean((avg - info/br)^2)) in
 * float; criteria = (int)(1.28f * sd) + avg. The square is formed in 64-bit (the reference's
 * int product overflows for large blocks). */
int64_t oracle_criteria(const int64_t* block_info, int64_t nb, int32_t br) {
  if (nb <= 0) return 0;
  int64_t neC = 0;
  for (int64_t b = 0; b < nb; ++b) neC += block_info[b];
  const int64_t avg = (neC / nb) / br;
  float sd = 0.0f;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t d = avg - block_info[b] / br;
    sd += (float)(d * d);
  }
  sd /= (float)nb;
  sd = sqrtf(sd);
  return (int64_t)(1.28f * sd) + avg;
}

/* analyze_Block: 0 = GPU only (above criteria), 1 = CPU preferred (<= 64 products/row), 2 = any. */
int32_t oracle_analyze_block(int64_t criteria, int64_t avg) {
  if (avg > criteria) return 0;
  if (avg <= 64) return 1;
  return 2;
}

static void sort_i32(int32_t* a, int64_t n) {
  if (n < 32) {
    for (int64_t i = 1; i < n; ++i) {
      int32_t v = a[i];
      int64_t j = i - 1;
      while (j >= 0 && a[j] > v) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = v;
    }
    return;
  }
  /* LSD radix sort on 32-bit non-negative keys, 11 bits per pass */
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* src = a;
  int32_t* dst = tmp;
  for (int sh = 0; sh < 32; sh += 11) {
    int64_t cnt[2049];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; ++i) cnt[((uint32_t)src[i] >> sh & 2047u) + 1]++;
    for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
    for (int64_t i = 0; i < n; ++i) dst[cnt[(uint32_t)src[i] >> sh & 2047u]++] = src[i];
    int32_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(int32_t) * (size_t)n);
  free(tmp);
}

/* Symbolic count: Cptr (nr + 1) = exclusive prefix of |union of B_k columns over k in A_i|. */
int64_t oracle_spgemm_count(int64_t nr, int64_t ncols, const int64_t* Aptr, const int32_t* Aidx,
                            const int64_t* Bptr, const int32_t* Bidx, int64_t* Cptr, int threads) {
  threads = nthreads_or_default(threads);
  Cptr[0] = 0;
#pragma omp parallel num_threads(threads)
  {
    int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncols > 0 ? ncols : 1));
    for (int64_t j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < nr; ++i) {
      int64_t n = 0;
      for (int64_t a = Aptr[i]; a < Aptr[i + 1]; ++a) {
        const int32_t k = Aidx[a];
        for (int64_t b = Bptr[k]; b < Bptr[k + 1]; ++b) {
          const int32_t c = Bidx[b];
          if (mark[c] != (int32_t)i) {
            mark[c] = (int32_t)i;
            ++n;
          }
        }
      }
      Cptr[i + 1] = n;
    }
    free(mark);
  }
  for (int64_t i = 0; i < nr; ++i) Cptr[i + 1] += Cptr[i];
  return Cptr[nr];
}

/* Numeric fill for the Cptr from oracle_spgemm_count. Cabs (optional) receives the sum of
 * |a_ik * b_kj| per output entry, the scale of the floating-point error of any summation
 * order. */
void oracle_spgemm_fill(int64_t nr, int64_t ncols, const int64_t* Aptr, const int32_t* Aidx, const double* Aval,
                        const int64_t* Bptr, const int32_t* Bidx, const double* Bval, const int64_t* Cptr,
                        int32_t* Cidx, double* Cval, double* Cabs, int threads) {
  threads = nthreads_or_default(threads);
#pragma omp parallel num_threads(threads)
  {
    int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncols > 0 ? ncols : 1));
    double* acc = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
    double* aab = Cabs ? (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double)) : NULL;
    for (int64_t j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < nr; ++i) {
      int32_t* cols = Cidx + Cptr[i];
      int64_t n = 0;
      for (int64_t a = Aptr[i]; a < Aptr[i + 1]; ++a) {
        const int32_t k = Aidx[a];
        const double va = Aval[a];
        for (int64_t b = Bptr[k]; b < Bptr[k + 1]; ++b) {
          const int32_t c = Bidx[b];
          const double p = va * Bval[b];
          if (mark[c] != (int32_t)i) {
            mark[c] = (int32_t)i;
            cols[n++] = c;
            acc[c] = p;
            if (aab) aab[c] = fabs(p);
          } else {
            acc[c] += p;
            if (aab) aab[c] += fabs(p);
          }
        }
      }
      sort_i32(cols, n);
      for (int64_t m = 0; m < n; ++m) {
        Cval[Cptr[i] + m] = acc[cols[m]];
        if (aab) Cabs[Cptr[i] + m] = aab[cols[m]];
      }
    }
    free(mark);
    free(acc);
    free(aab);
  }
}

/* Rows `sel[0..ns)` only (for sampled checks of very large products). Output CSR over the
 * selected rows: Sptr (ns + 1), Sidx / Sval (and optionally Sabs, as in oracle_spgemm_fill)
 * sized by the caller from the returned count of a first call with Sidx == NULL. */
int64_t oracle_spgemm_rows(int64_t ns, const int32_t* sel, int64_t ncols, const int64_t* Aptr, const int32_t* Aidx,
                           const double* Aval, const int64_t* Bptr, const int32_t* Bidx, const double* Bval,
                           int64_t* Sptr, int32_t* Sidx, double* Sval, double* Sabs, int threads) {
  int64_t* Aptr_s = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns + 1));
  int64_t tot = 0;
  Aptr_s[0] = 0;
  for (int64_t r = 0; r < ns; ++r) {
    tot += Aptr[sel[r] + 1] - Aptr[sel[r]];
    Aptr_s[r + 1] = tot;
  }
  int32_t* Aidx_s = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tot > 0 ? tot : 1));
  double* Aval_s = (double*)malloc(sizeof(double) * (size_t)(tot > 0 ? tot : 1));
  for (int64_t r = 0; r < ns; ++r) {
    memcpy(Aidx_s + Aptr_s[r], Aidx + Aptr[sel[r]], sizeof(int32_t) * (size_t)(Aptr_s[r + 1] - Aptr_s[r]));
    if (Aval) memcpy(Aval_s + Aptr_s[r], Aval + Aptr[sel[r]], sizeof(double) * (size_t)(Aptr_s[r + 1] - Aptr_s[r]));
  }
  const int64_t nnz = oracle_spgemm_count(ns, ncols, Aptr_s, Aidx_s, Bptr, Bidx, Sptr, threads);
  if (Sidx && Sval) oracle_spgemm_fill(ns, ncols, Aptr_s, Aidx_s, Aval_s, Bptr, Bidx, Bval, Sptr, Sidx, Sval, Sabs, threads);
  free(Aptr_s);
  free(Aidx_s);
  free(Aval_s);
  return nnz;
}

static int64_t find_col(const int32_t* cols, int64_t n, int32_t c) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (cols[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && cols[lo] == c) ? lo : -1;
}

/* Masked gradients of C = A * B for an upstream gradient dC on pattern(C):
 *   dA_ik = sum_{j in B_k} dC_ij * B_kj        (on pattern(A))
 *   dB_kj = sum_{i : k in A_i} A_ik * dC_ij    (on pattern(B))
 * FP64, rows in parallel (OpenMP): dA in B's order within each entry, dB by atomic adds (the
 * order of the terms of a dB entry varies, within FP64 rounding). Returns the number of products
 * whose column was missing from C's row (0 for a C produced by the forward). */
int64_t oracle_backward(int64_t nr, const int64_t* Aptr, const int32_t* Aidx, const double* Aval, const int64_t* Bptr,
                        const int32_t* Bidx, const double* Bval, const int64_t* Cptr, const int32_t* Cidx,
                        const double* dC, double* dA, double* dB, int64_t nnzB) {
  int64_t missing = 0;
  for (int64_t p = 0; p < nnzB; ++p) dB[p] = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : missing)
  for (int64_t i = 0; i < nr; ++i) {
    const int32_t* cols = Cidx + Cptr[i];
    const int64_t n = Cptr[i + 1] - Cptr[i];
    for (int64_t a = Aptr[i]; a < Aptr[i + 1]; ++a) {
      const int32_t k = Aidx[a];
      double s = 0.0;
      for (int64_t b = Bptr[k]; b < Bptr[k + 1]; ++b) {
        const int64_t m = find_col(cols, n, Bidx[b]);
        if (m < 0) {
          ++missing;
          continue;
        }
        const double g = dC[Cptr[i] + m];
        s += g * Bval[b];
#pragma omp atomic
        dB[b] += Aval[a] * g;
      }
      dA[a] = s;
    }
  }
  return missing;
}
