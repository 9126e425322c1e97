// gfx950 (CDNA4) kernels of the DBD SpGEMM path: C = A * B, CSR, int64 row pointers,
// int32 column indices, FP32 or FP64 values.
//
// Reference kernels replaced (S/ = dbd/for_small_datasets/src/ of the reference):
//   calculate_UpperBound  S/preProc.cu:170-179   -> k_row_products (no atomics, int64)
//   rowproduct_* + merge  S/computeGPU.cu:135-329 -> k_symbolic_hash, k_numeric_wave[_fast],
//                                                   k_numeric_block, k_*_long
//   collect               S/computeGPU.cu:331-341 -> not needed: exact offsets come from the
//                                                   symbolic pass, numeric writes in place.
// Differences on purpose: first touch is detected by a key compare-and-swap (the reference
// tests |acc| < 1e-4, S/computeGPU.cu:311-312), A's values are not truncated to int
// (S/computeGPU.cu:140), columns are emitted sorted, and every count is 64-bit.
//
// Layout of work: rows are binned by a per-row key (symbolic: min(P_i, ncols); numeric:
// nnz_i). Each bin has a kernel shape (T hash slots in LDS, NT threads per row, G lanes per
// A entry). Rows above the largest LDS table go to window kernels that sweep the row's
// column range in LDS-sized windows with one cursor per A entry.
#include "kernels.h"

#include <algorithm>
#include <climits>
#include <utility>
#include <mutex>
#include <map>

namespace dbd {

constexpr int kEmpty = -1;

// ------------------------------------------------------------------------------------
// wave / group primitives
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wait until no vector memory operation of this wave is outstanding (s_waitcnt vmcnt(0); the
// other counters at their maximum). Placed at the end of rarely taken branches that issue
// loads, so the compiler's count of outstanding loads on the common path stays exact.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// 64-bit readlane of a wave-uniform lane index.
__device__ __forceinline__ int64_t readlane64(int64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A "group" is NT threads that own one row. NT <= 64: a (sub-)wave, synchronised with
// wave_sync (several groups per workgroup). NT > 64: the whole workgroup (__syncthreads).
template <int NT>
__device__ __forceinline__ void gsync() {
  if constexpr (NT <= 64)
    wave_sync();
  else
    __syncthreads();
}

template <int NT>
struct Group {
  static constexpr int W = NT < 64 ? NT : 64;
  static constexpr int NW = NT > 64 ? NT / 64 : 1;

  template <typename T, typename Op>
  __device__ static T reduce(T v, Op op, T* sc) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, W));
    if constexpr (NT > 64) {
      const int t = threadIdx.x % NT;
      if ((t & 63) == 0) sc[t >> 6] = v;
      __syncthreads();
      v = sc[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) v = op(v, sc[i]);
      __syncthreads();
    }
    return v;
  }

  // Exclusive prefix of v over the group; total returned through `total`.
  template <typename T>
  __device__ static T exscan(T v, T& total, T* sc) {
    const int t = threadIdx.x % NT;
    const int lane = t % W;
    T x = v;
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
      T y = __shfl_up(x, o, W);
      if (lane >= o) x += y;
    }
    if constexpr (NT <= 64) {
      total = __shfl(x, W - 1, W);
      return x - v;
    } else {
      const int wid = t >> 6;
      if (lane == 63) sc[wid] = x;
      __syncthreads();
      T pre = 0, tot = 0;
#pragma unroll
      for (int i = 0; i < NW; ++i) {
        T s = sc[i];
        if (i < wid) pre += s;
        tot += s;
      }
      __syncthreads();
      total = tot;
      return pre + x - v;
    }
  }
};

struct OpMin {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct OpMax {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpOr {
  __device__ int operator()(int a, int b) const { return a | b; }
};
struct OpAdd {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a + b; }
};

// Whole-wave reduction of an int without LDS: DPP within rows of 16 lanes (quad swaps, half
// mirror, mirror), then the four row results through readlane. All 64 lanes must be active.
template <typename Op>
__device__ __forceinline__ int wave_reduce_dpp(int x, Op op) {
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, false));  // row_half_mirror
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x140, 0xF, 0xF, false));  // row_mirror
  const int r0 = __builtin_amdgcn_readlane(x, 0), r1 = __builtin_amdgcn_readlane(x, 16);
  const int r2 = __builtin_amdgcn_readlane(x, 32), r3 = __builtin_amdgcn_readlane(x, 48);
  return op(op(r0, r1), op(r2, r3));
}

// Inclusive prefix sum over the wave with DPP: shifts of 1-3, 4 and 8 lanes inside each row of
// 16 lanes, then the row totals through the row broadcasts of lanes 15 and 31. All 64 lanes
// must be active.
__device__ __forceinline__ int wave_incl_scan_dpp(int x) {
  int s = x + __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
  s += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);         // row_shr:2
  s += __builtin_amdgcn_update_dpp(0, x, 0x113, 0xF, 0xF, false);         // row_shr:3
  s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xE, false);         // row_shr:4, lanes 4-15
  s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xF, 0xC, false);         // row_shr:8, lanes 8-15
  s += __builtin_amdgcn_update_dpp(0, s, 0x142, 0xA, 0xF, false);         // row_bcast:15 -> rows 1, 3
  s += __builtin_amdgcn_update_dpp(0, s, 0x143, 0xC, 0xF, false);         // row_bcast:31 -> rows 2, 3
  return s;
}

// Mask of the G lanes (within this wave) that share this lane's subgroup.
template <int G>
__device__ __forceinline__ uint64_t subgroup_mask() {
  if constexpr (G >= 64) {
    return ~0ull;
  } else {
    const int first = (threadIdx.x & 63) & ~(G - 1);
    return ((1ull << G) - 1ull) << first;
  }
}

// Fibonacci hashing of a column into a power-of-two table.
template <int T>
__device__ __forceinline__ uint32_t hslot(int c) {
  constexpr int LOG = __builtin_ctz(T);
  return ((uint32_t)c * 2654435761u) >> (32 - LOG);
}

// Insert key c; returns 1 if this call created the entry, 0 if it was present, -1 if the
// table is full (cannot happen when the planner sized it; checked anyway, never spins).
template <int T>
__device__ __forceinline__ int hash_insert_key(int* hk, int c) {
  uint32_t h = hslot<T>(c);
#pragma unroll 1
  for (int probe = 0; probe < T; ++probe) {
    // compare-and-swap first: at load <= 1/2 most probes find the slot empty or holding c
    const int old = atomicCAS(hk + h, kEmpty, c);
    if (old == kEmpty) return 1;
    if (old == c) return 0;
    h = (h + 1) & (T - 1);
  }
  return -1;
}

template <typename V>
__device__ __forceinline__ void lds_add(V* p, V v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Insert-or-accumulate (c, v). Returns false only if the table is full.
template <int T, typename V>
__device__ __forceinline__ bool hash_insert_add(int* hk, V* hv, int c, V v) {
  uint32_t h = hslot<T>(c);
#pragma unroll 1
  for (int probe = 0; probe < T; ++probe) {
    const int old = atomicCAS(hk + h, kEmpty, c);
    if (old == kEmpty || old == c) {
      lds_add(hv + h, v);
      return true;
    }
    h = (h + 1) & (T - 1);
  }
  return false;
}

// Product sink of the hash kernels: tracks the column range, inserts keys (and accumulates
// values when VALS). With BATCH, batch() issues the first probe of all U products before
// looking at any result (U LDS compare-and-swaps in flight together); collisions then continue
// probing. Without it the products are inserted one after another.
template <int T, typename V, bool VALS, bool BATCH = true>
struct HashSink {
  int* K;
  V* H;
  int cmin = INT_MAX, cmax = -1, created = 0, bad = 0;
  __device__ __forceinline__ void probe_from(uint32_t h, int c, V v) {
#pragma unroll 1
    for (int probe = 1; probe < T; ++probe) {
      h = (h + 1) & (T - 1);
      const int old = atomicCAS(K + h, kEmpty, c);
      if (old == kEmpty || old == c) {
        created += (old == kEmpty);
        if (VALS) lds_add(H + h, v);
        return;
      }
    }
    bad = 1;
  }
  template <int U>
  __device__ __forceinline__ void batch(const int (&c)[U], const V (&v)[U]) {
    if constexpr (!BATCH) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c[u] >= 0) one(c[u], v[u]);
      return;
    }
    uint32_t h[U];
    int old[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      h[u] = hslot<T>(c[u]);
      old[u] = (c[u] >= 0) ? atomicCAS(K + h[u], kEmpty, c[u]) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c[u] < 0) continue;
      cmin = min(cmin, c[u]);
      cmax = max(cmax, c[u]);
      if (old[u] == kEmpty || old[u] == c[u]) {
        created += (old[u] == kEmpty);
        if (VALS) lds_add(H + h[u], v[u]);
      } else {
        probe_from(h[u], c[u], v[u]);
      }
    }
  }
  __device__ __forceinline__ void one(int c, V v) {
    cmin = min(cmin, c);
    cmax = max(cmax, c);
    const uint32_t h = hslot<T>(c);
    const int old = atomicCAS(K + h, kEmpty, c);
    if (old == kEmpty || old == c) {
      created += (old == kEmpty);
      if (VALS) lds_add(H + h, v);
    } else {
      probe_from(h, c, v);
    }
  }
};

// Lookup of a key known to be present; returns slot or -1.
template <int T>
__device__ __forceinline__ int hash_find(const int* hk, int c) {
  uint32_t h = hslot<T>(c);
#pragma unroll 1
  for (int probe = 0; probe < T; ++probe) {
    int cur = hk[h];
    if (cur == c) return (int)h;
    if (cur == kEmpty) return -1;
    h = (h + 1) & (T - 1);
  }
  return -1;
}

// Slots a lane owns in a one-wave row table of T slots (T/64 per lane): 16-byte chunks
// t, t + 64, ... (fewer slots per lane: the slots t*IPT ..), so a wide LDS access of the whole
// wave covers one contiguous block and hits every bank once.
template <int T>
__device__ __forceinline__ int slot_of(int t, int q) {
  constexpr int IPT = T / 64;
  constexpr int CW = IPT < 4 ? IPT : 4;
  return CW * ((q / CW) * 64 + t) + q % CW;
}

// Clear a one-wave row table (keys to kEmpty, values to 0) with contiguous 16-byte stores.
template <int T, typename V>
__device__ __forceinline__ void clear_table(int t, int* K, V* H) {
  if constexpr (T >= 256) {
#pragma unroll
    for (int c = 0; c < T / 256; ++c) reinterpret_cast<int4*>(K)[c * 64 + t] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
  } else {
#pragma unroll
    for (int q = 0; q < T / 64; ++q) K[q * 64 + t] = kEmpty;
  }
  if (H) {
    constexpr int CH = T * (int)sizeof(V) / 16;  // 16-byte chunks of values
#pragma unroll
    for (int c = t; c < CH; c += 64) reinterpret_cast<int4*>(H)[c] = make_int4(0, 0, 0, 0);
  }
}

// Descriptor of the A entry a lane holds: start and length of the selected B row, A's value.
template <typename V>
struct ADesc {
  int64_t b0;
  int len;
  V va;
};

// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time (for indices
// that must be immediates, such as DPP lane selects).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Which A entry of a 64-entry chunk a lane holds. With G = 16 the layout is transposed: lane
// 16r + j holds entry 4j + r, so in step S (entries 16S .. 16S+15) subgroup r's u-th entry,
// 16S + 4u + r, sits in lane 4S + u of the subgroup's own row of 16 lanes and reaches all of
// the row through one DPP row broadcast (no LDS traffic). With G = 64 lane e holds entry e and
// the wave takes it with readlane.
template <int G>
__device__ __forceinline__ int desc_slot(int lane) {
  static_assert(G == 16 || G == 64, "walk widths");
  return G == 16 ? 4 * (lane & 15) + (lane >> 4) : lane;
}

// Field of the entry at broadcast index J (G = 16: lane J of each row; G = 64: lane e) to all
// lanes that walk it.
template <int G, int J>
__device__ __forceinline__ int ebcast(int x, int e) {
  if constexpr (G == 16)
    return __builtin_amdgcn_update_dpp(0, x, 0x150 + J, 0xF, 0xF, false);  // row_newbcast:J
  else
    return __builtin_amdgcn_readlane(x, e);
}
template <int G, int J>
__device__ __forceinline__ int64_t ebcast64(int64_t x, int e) {
  const uint32_t lo = (uint32_t)ebcast<G, J>((int)(uint32_t)x, e);
  const uint32_t hi = (uint32_t)ebcast<G, J>((int)(uint32_t)((uint64_t)x >> 32), e);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <int G, int J>
__device__ __forceinline__ double ebcastv(double x, int e) {
  return __longlong_as_double(ebcast64<G, J>(__double_as_longlong(x), e));
}
template <int G, int J>
__device__ __forceinline__ float ebcastv(float x, int e) {
  return __int_as_float(ebcast<G, J>(__float_as_int(x), e));
}

// Lane `lane` loads the descriptor of A entry base + desc_slot(lane). These loads cover the
// chunk's entries in A's order (coalesced). Lanes past a1 load entry a1 - 1 instead, and the
// values are kept raw: the consumers (gather_step, consume_step) take only entries below the
// chunk's count. Unconditional loads keep the compiler's in-order count of outstanding loads
// exact (a branch around a load, or a select on its result right after it, makes the wave wait
// for the load at that point). a1 >= 1 for every row that reaches a kernel (such rows have
// products).
template <int G, bool WITHV, typename V>
__device__ __forceinline__ ADesc<V> load_desc(int64_t base, int64_t a1, int lane, const int64_t* __restrict__ bstart,
                                              const int32_t* __restrict__ blen, const V* __restrict__ Aval) {
  const int64_t j = base + desc_slot<G>(lane);
  const int64_t jj = j < a1 ? j : (a1 > 0 ? a1 - 1 : 0);
  ADesc<V> d;
  d.b0 = bstart[jj];
  d.len = blen[jj];
  d.va = WITHV ? Aval[jj] : V(0);
  return d;
}

// One gather step over a chunk of A entries: the wave's 64/G subgroups take U entries each
// (G = 16: step S, entries 16S .. 16S+15; G = 64: entries e0 .. e0+U-1) and issue the U B-row
// gathers (the first G entries of each B row). Only the loaded columns and B values are kept;
// they are masked and multiplied when the step is consumed.
template <int U, typename V>
struct Step {
  int c[U];   // B columns as loaded (masked when consumed)
  V bv[U];    // B values as loaded
};

// index of entry u of this lane's subgroup in the step
template <int G, int S>
__device__ __forceinline__ int step_entry(int e0, int u, int sg) {
  return G == 16 ? 16 * S + 4 * u + sg : e0 + u;
}

template <int G, int U, bool WITHV, typename V, int S = 0>
__device__ __forceinline__ void gather_step(Step<U, V>& st, int lane, int e0, int cnt, const ADesc<V>& d,
                                            const int32_t* __restrict__ Bidx, const V* __restrict__ Bval) {
  static_assert(G == 64 || U == 4, "with G = 16 a step is 4 subgroups x 4 entries");
  const int sg = lane / G, lg = lane % G;
  // unconditional loads: lanes past the B row re-read its last entry, entries without a B row
  // read entry 0 (B has entries whenever a row reaches a kernel). The loaded values are kept
  // raw: masking them here would make the wave wait for the loads right away.
  static_for<U>([&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const int e = step_entry<G, S>(e0, u, sg);
    const int er = min(e0 + u, 63);
    const int64_t pb = ebcast64<G, 4 * S + (u & 3)>(d.b0, er);
    const int lv = ebcast<G, 4 * S + (u & 3)>(d.len, er);
    const int pl = e < cnt ? lv : 0;
    const int64_t at = pl > 0 ? pb + min(lg, pl - 1) : 0;
    st.c[u] = Bidx[at];
    st.bv[u] = WITHV ? Bval[at] : V(0);
  });
}

// Hand a gathered step to the sink: f.batch(c[U], v[U]) for the first G entries of each B row
// (c < 0 marks no product), f.one(c, v) for the tails of B rows longer than G; v = A_ik * B_kj.
template <int G, int U, bool WITHV, typename V, typename F, int S = 0>
__device__ __forceinline__ void consume_step(const Step<U, V>& st, int lane, int e0, int cnt, const ADesc<V>& d,
                                             const int32_t* __restrict__ Bidx, const V* __restrict__ Bval, F& f) {
  const int sg = lane / G, lg = lane % G;
  int pl[U], c[U];
  V v[U], pv[U];
  bool tail = false;
  static_for<U>([&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const int e = step_entry<G, S>(e0, u, sg);
    const int er = min(e0 + u, 63);
    const int lv = ebcast<G, 4 * S + (u & 3)>(d.len, er);
    if constexpr (WITHV) pv[u] = ebcastv<G, 4 * S + (u & 3)>(d.va, er);
    else pv[u] = V(0);
    pl[u] = e < cnt ? lv : 0;
    const bool has = lg < pl[u];
    c[u] = has ? st.c[u] : -1;
    v[u] = (WITHV && has) ? pv[u] * st.bv[u] : V(0);
    tail |= pl[u] > G;
  });
  f.template batch<U>(c, v);
  if (__builtin_amdgcn_ballot_w64(tail)) {  // B rows longer than G: the rest of them, one by one
    static_for<U>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t pb = ebcast64<G, 4 * S + (u & 3)>(d.b0, min(e0 + u, 63));
      for (int q = G + lg; q < pl[u]; q += G) {
        const int64_t p = pb + q;
        f.one(Bidx[p], WITHV ? pv[u] * Bval[p] : V(0));
      }
    });
    wait_vmem();
  }
}

template <int G, int U, bool WITHV, typename V, typename F, int S>
__device__ __forceinline__ void walk_step(int lane, int e0, int cnt, const ADesc<V>& d,
                                          const int32_t* __restrict__ Bidx, const V* __restrict__ Bval, F& f) {
  Step<U, V> st;
  gather_step<G, U, WITHV, V, S>(st, lane, e0, cnt, d, Bidx, Bval);
  consume_step<G, U, WITHV, V, F, S>(st, lane, e0, cnt, d, Bidx, Bval, f);
}

// All products of `cnt` (<= 64) A entries of one chunk, from entry `from` on (a multiple of a
// step). Control flow around the broadcasts is wave-uniform.
template <int G, int U, bool WITHV, typename V, typename F>
__device__ __forceinline__ void walk_chunk(int lane, int cnt, const ADesc<V>& d, const int32_t* __restrict__ Bidx,
                                           const V* __restrict__ Bval, F& f, int from = 0) {
  if constexpr (G == 16) {  // steps are compile-time (their broadcast lanes are DPP immediates)
    if (from <= 0 && 0 < cnt) walk_step<G, U, WITHV, V, F, 0>(lane, 0, cnt, d, Bidx, Bval, f);
    if (from <= 16 && 16 < cnt) walk_step<G, U, WITHV, V, F, 1>(lane, 16, cnt, d, Bidx, Bval, f);
    if (from <= 32 && 32 < cnt) walk_step<G, U, WITHV, V, F, 2>(lane, 32, cnt, d, Bidx, Bval, f);
    if (from <= 48 && 48 < cnt) walk_step<G, U, WITHV, V, F, 3>(lane, 48, cnt, d, Bidx, Bval, f);
  } else {
    for (int e0 = from; e0 < cnt; e0 += U) walk_step<G, U, WITHV, V, F, 0>(lane, e0, cnt, d, Bidx, Bval, f);
  }
}

// All products of one row for a group of NT >= 64 threads: wave w handles A entries
// a0 + 64w, a0 + 64w + NT, ...; `first` holds the descriptors of the wave's first chunk when
// the caller prefetched them.
template <int NT, int G, int U, bool WITHV, typename V, typename F>
__device__ __forceinline__ void for_row_products_pre(int t, int64_t a0, int64_t a1, const ADesc<V>& first,
                                                     const int64_t* __restrict__ bstart,
                                                     const int32_t* __restrict__ blen, const V* __restrict__ Aval,
                                                     const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
                                                     F& f, const Step<U, V>* step0 = nullptr) {
  static_assert(NT >= 64, "wave-level walk");
  constexpr int STEP = G == 16 ? 16 : U;  // entries per step
  const int lane = t & 63, wv = t >> 6;
  int64_t base = a0 + 64 * wv;
  if (base < a1) {
    const int cnt = (int)min((int64_t)64, a1 - base);
    int from = 0;
    if (step0) {  // the caller already issued the first step's gathers
      consume_step<G, U, WITHV, V, F, 0>(*step0, lane, 0, cnt, first, Bidx, Bval, f);
      from = STEP;
    }
    if (from < cnt) {
      walk_chunk<G, U, WITHV, V>(lane, cnt, first, Bidx, Bval, f, from);
      if (step0) wait_vmem();
    }
  }
  if (base + NT < a1) {
    for (base += NT; base < a1; base += NT) {
      const ADesc<V> d = load_desc<G, WITHV, V>(base, a1, lane, bstart, blen, Aval);
      walk_chunk<G, U, WITHV, V>(lane, (int)min((int64_t)64, a1 - base), d, Bidx, Bval, f);
    }
    if (step0) wait_vmem();
  }
}

template <int NT, int G, int U, bool WITHV, typename V, typename F>
__device__ __forceinline__ void for_row_products(int t, int64_t a0, int64_t a1, const int64_t* __restrict__ bstart,
                                                 const int32_t* __restrict__ blen, const V* __restrict__ Aval,
                                                 const int32_t* __restrict__ Bidx, const V* __restrict__ Bval, F& f) {
  const ADesc<V> first = load_desc<G, WITHV, V>(a0 + 64 * (t >> 6), a1, t & 63, bstart, blen, Aval);
  for_row_products_pre<NT, G, U, WITHV, V>(t, a0, a1, first, bstart, blen, Aval, Bidx, Bval, f);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2). When the grid
// is a multiple of 8, virtual block ids are regrouped so that consecutive ids (consecutive bin
// slots, i.e. nearby rows) run on the same XCD.
__device__ __forceinline__ int64_t xcd_block() {
  const uint32_t b = blockIdx.x, n = gridDim.x;
  if (n % 8 != 0) return b;
  return (int64_t)(b % 8) * (n / 8) + b / 8;
}

// Rows handled by one wave, software-pipelined. The wave's rows are the bin slots s0, s0 +
// stride, ...; their ids and row pointers are loaded 64 rows at a time with vector loads (lane
// L holds row L of the batch) and read out with readlane, so no scalar loads are outstanding
// while the wave waits on LDS (scalar loads and LDS share one counter). The next row's A
// descriptors are issued at the top of each iteration and its first B gathers after the
// current row's products are in LDS, so each wait is for loads issued a good while earlier.
template <int G, bool WITHV, typename V>
struct RowPipe {
  int64_t slot, stride, end;   // slot of the current row (wave-uniform)
  int kb;                      // batch index of the next row (wave-uniform)
  int bR;                      // batch: row ids (per lane)
  int64_t bA0, bA1, bC0, bC1;  // batch: row pointers (per lane)
  int row, row1;               // current / next row (wave-uniform)
  int64_t a0, a1, c0, c1;      // current row: A range and output range
  int64_t na0, na1, nc0, nc1;  // next row
  ADesc<V> d, nd;              // descriptors of the current / next row
  bool more;                   // a next row exists

  __device__ __forceinline__ void load_batch(int64_t s0, int lane, const int32_t* __restrict__ rows,
                                             const int64_t* __restrict__ Aptr, const int64_t* __restrict__ Cptr) {
    int64_t ls = s0 + (int64_t)lane * stride;
    if (ls >= end) ls = end - 1;  // past the bin: repeat its last row (never used)
    const int r = rows[ls];
    bR = r;
    bA0 = Aptr[r];
    bA1 = Aptr[r + 1];
    if (Cptr) {
      bC0 = Cptr[r];
      bC1 = Cptr[r + 1];
    } else {
      bC0 = bC1 = 0;
    }
  }
  __device__ __forceinline__ void extract(int j, int& r, int64_t& x0, int64_t& x1, int64_t& y0, int64_t& y1) const {
    r = __builtin_amdgcn_readlane(bR, j);
    x0 = readlane64(bA0, j);
    x1 = readlane64(bA1, j);
    y0 = readlane64(bC0, j);
    y1 = readlane64(bC1, j);
  }
  __device__ __forceinline__ void start(int64_t s0, int64_t st, int64_t e, const int32_t* __restrict__ rows,
                                        const int64_t* __restrict__ Aptr, const int64_t* __restrict__ Cptr, int lane,
                                        const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
                                        const V* __restrict__ Aval) {
    slot = s0;
    stride = st;
    end = e;
    load_batch(s0, lane, rows, Aptr, Cptr);
    kb = 0;
    extract(0, row, a0, a1, c0, c1);
    d = load_desc<G, WITHV, V>(a0, a1, lane, bstart, blen, Aval);
  }
  // top of an iteration: pointers (a new batch every 64 rows) and descriptors of the next row;
  // past the last row the current row is repeated, so the loads stay unconditional
  __device__ __forceinline__ void prefetch(int lane, const int32_t* __restrict__ rows,
                                           const int64_t* __restrict__ Aptr, const int64_t* __restrict__ Cptr,
                                           const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
                                           const V* __restrict__ Aval) {
    more = slot + stride < end;
    if (more) {
      ++kb;
      if (kb == 64) {
        load_batch(slot + stride, lane, rows, Aptr, Cptr);
        kb = 0;
        wait_vmem();
      }
    }
    extract(kb, row1, na0, na1, nc0, nc1);
    nd = load_desc<G, WITHV, V>(na0, na1, lane, bstart, blen, Aval);
  }
  __device__ __forceinline__ bool advance() {
    if (!more) return false;
    slot += stride;
    row = row1;
    a0 = na0;
    a1 = na1;
    c0 = nc0;
    c1 = nc1;
    d = nd;
    return true;
  }
};

// Error bits written to the device error word.
constexpr int kErrTableFull = 1;
constexpr int kErrCountMismatch = 2;
constexpr int kErrWindow = 4;
constexpr int kErrMissingKey = 8;
constexpr int kErrCapacity = 16;  // output larger than the capacity the caller provided

// ------------------------------------------------------------------------------------
// fills, scan, gathers
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static int fill_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

void launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<int64_t>, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
}
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
}
void launch_fill_bytes0(void* p, int64_t nbytes, hipStream_t s) {
  // callers pass 4-byte multiples (value arrays, counters)
  launch_fill_i32((int32_t*)p, nbytes / 4, 0, s);
}

// Reduce-then-scan over tiles of 4096 int64 (256 threads x 16 items). No look-back and no
// memsets: safe to capture in a hipGraph and replay.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = kScanThreads * kScanItems;

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int64_t* __restrict__ in, int64_t n,
                                                              int64_t* __restrict__ part) {
  __shared__ int64_t sc[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    if (j < n) s += in[j];
  }
  s = Group<kScanThreads>::reduce(s, OpAdd(), sc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Single workgroup: exclusive scan of part[0..nb) in place; part[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_partials(int64_t* __restrict__ part, int64_t nb) {
  __shared__ int64_t sc[16];
  int64_t carry = 0;
  for (int64_t base = 0; base < nb; base += 1024) {
    int64_t j = base + threadIdx.x;
    int64_t v = (j < nb) ? part[j] : 0;
    int64_t tot;
    int64_t ex = Group<1024>::exscan(v, tot, sc);
    if (j < nb) part[j] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) part[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int64_t* in, int64_t* out, int64_t n,
                                                             const int64_t* __restrict__ part, int64_t nb) {
  __shared__ int64_t sc[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t loc[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    loc[i] = (j < n) ? in[j] : 0;
    s += loc[i];
  }
  int64_t tot;
  int64_t pre = Group<kScanThreads>::exscan(s, tot, sc) + part[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    if (j < n) out[j] = pre;
    pre += loc[i];
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = part[nb];
}

void launch_exclusive_scan(const int64_t* in, int64_t* out, int64_t n, int64_t* tmp, hipStream_t s) {
  if (n <= 0) {
    launch_fill_i64(out, 1, 0, s);
    return;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, tmp, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n, tmp, nb);
}

__global__ void k_gather_strided(const int64_t* __restrict__ ptr, int64_t n, int64_t stride, int64_t nb,
                                 int64_t* __restrict__ out) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = b * stride;
    if (r > n) r = n;
    out[b] = ptr[r];
  }
}

void launch_gather_strided(const int64_t* ptr, int64_t n, int64_t stride, int64_t nb, int64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_strided, dim3(fill_grid(nb + 1)), dim3(256), 0, s, ptr, n, stride, nb, out);
}

// Output capacity check inside a captured pipeline: when the exact output size (*total) exceeds
// the caller's capacity, flag it and empty every bin range (offs[0 .. n)), so the kernels that
// follow drain nothing and write nothing.
__global__ void k_capacity_guard(const int64_t* __restrict__ total, int64_t cap, int64_t* __restrict__ offs, int64_t n,
                                 int* __restrict__ err) {
  if (*total <= cap) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrCapacity);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    offs[i] = 0;
}

void launch_capacity_guard(const int64_t* total, int64_t cap, int64_t* offs, int64_t n, int* err, hipStream_t s) {
  hipLaunchKernelGGL(k_capacity_guard, dim3(fill_grid(n)), dim3(256), 0, s, total, cap, offs, n, err);
}

// ------------------------------------------------------------------------------------
// row analysis: P_i = sum over A_i of B row lengths (16 lanes per row, no atomics)
// ------------------------------------------------------------------------------------
constexpr int kRpLanes = 16;

__global__ __launch_bounds__(256) void k_row_products(int64_t nr, const int64_t* __restrict__ Aptr,
                                                      const int32_t* __restrict__ Aidx,
                                                      const int64_t* __restrict__ Bptr, int64_t* __restrict__ P,
                                                      int64_t* __restrict__ nnz, int64_t* __restrict__ bstart,
                                                      int32_t* __restrict__ blen) {
  const int l = threadIdx.x % kRpLanes;
  for (int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kRpLanes; row < nr;
       row += (int64_t)gridDim.x * (256 / kRpLanes)) {
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
    int64_t s = 0;
    for (int64_t ka = a0 + l; ka < a1; ka += kRpLanes) {
      const int k = Aidx[ka];
      const int64_t b0 = Bptr[k];
      const int64_t len = Bptr[k + 1] - b0;
      s += len;
      if (bstart) {
        bstart[ka] = b0;
        blen[ka] = (int32_t)len;
      }
    }
#pragma unroll
    for (int o = kRpLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kRpLanes);
    if (l == 0) {
      P[row] = s;
      if (s == 0) nnz[row] = 0;
    }
  }
}

void launch_row_products(const Mat& A, const Mat& B, int64_t* P, int64_t* nnz, int64_t* bstart, int32_t* blen,
                         hipStream_t s) {
  if (A.nrows <= 0) return;
  int64_t g = (A.nrows * kRpLanes + 255) / 256;
  if (g > 262144) g = 262144;
  hipLaunchKernelGGL(k_row_products, dim3((unsigned)g), dim3(256), 0, s, A.nrows, A.ptr, A.idx, B.ptr, P, nnz, bstart,
                     blen);
}

// ------------------------------------------------------------------------------------
// band index of the long B rows (for the long-row window kernels)
// ------------------------------------------------------------------------------------
// B rows of at least `minlen` entries ("heavy" rows) get a slot; for each, the position (from the
// row's start) of its first entry at or past every multiple of 4096 columns, over nb1 entries,
// the last of which (past every column) holds the row's length. A long-row window of aligned
// columns then reads its exact element range from two entries instead of walking the row.
__global__ void k_heavy_flag(const int64_t* __restrict__ Bptr, int64_t nrows, int64_t minlen, int64_t* __restrict__ flag) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += (int64_t)gridDim.x * blockDim.x)
    flag[k] = (Bptr[k + 1] - Bptr[k]) >= minlen ? 1 : 0;
}

__global__ void k_heavy_slot(const int64_t* __restrict__ flag_scan, const int64_t* __restrict__ Bptr, int64_t nrows,
                             int64_t minlen, int32_t* __restrict__ slot, int32_t* __restrict__ slot_row) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += (int64_t)gridDim.x * blockDim.x) {
    const bool heavy = (Bptr[k + 1] - Bptr[k]) >= minlen;
    const int32_t sl = (int32_t)flag_scan[k];
    slot[k] = heavy ? sl : -1;
    if (heavy) slot_row[sl] = (int32_t)k;
  }
}

// One workgroup per heavy row: entry q of band position b is written by the element that opens
// band b (the first element whose band is >= b), bands past the last element by the tail.
__global__ __launch_bounds__(256) void k_band_index(const int32_t* __restrict__ slot_row, int64_t nslots,
                                                    const int64_t* __restrict__ Bptr, const int32_t* __restrict__ Bidx,
                                                    int64_t nb1, int32_t* __restrict__ band) {
  for (int64_t sl = blockIdx.x; sl < nslots; sl += gridDim.x) {
    const int k = slot_row[sl];
    const int64_t b0 = Bptr[k];
    const int32_t L = (int32_t)(Bptr[k + 1] - b0);
    int32_t* out = band + sl * nb1;
    for (int32_t q = threadIdx.x; q < L; q += blockDim.x) {
      const int bq = Bidx[b0 + q] >> kBandShift;
      const int bp = q > 0 ? (Bidx[b0 + q - 1] >> kBandShift) : -1;
      for (int b = bp + 1; b <= bq; ++b) out[b] = q;
    }
    const int blast = L > 0 ? (Bidx[b0 + L - 1] >> kBandShift) : -1;
    for (int64_t b = blast + 1 + threadIdx.x; b < nb1; b += blockDim.x) out[b] = L;
  }
}

// hist[j] = number of B rows with at least 2^j entries, j = 0 .. 31 (hist zeroed by the caller)
__global__ void k_len_hist(const int64_t* __restrict__ Bptr, int64_t nrows, int64_t* __restrict__ hist) {
  __shared__ int h[32];
  if (threadIdx.x < 32) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t L = Bptr[k + 1] - Bptr[k];
    if (L > 0) atomicAdd(&h[63 - __clzll((unsigned long long)L)], 1);  // floor(log2 L)
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    int64_t c = 0;
    for (int j = threadIdx.x; j < 32; ++j) c += h[j];  // rows with floor(log2 L) >= threadIdx.x
    if (c) atomicAdd((unsigned long long*)&hist[threadIdx.x], (unsigned long long)c);
  }
}

void launch_len_hist(const int64_t* Bptr, int64_t nrows, int64_t* hist, hipStream_t s) {
  launch_fill_i64(hist, 32, 0, s);
  if (nrows <= 0) return;
  int64_t g = (nrows + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_len_hist, dim3((unsigned)g), dim3(256), 0, s, Bptr, nrows, hist);
}

int64_t band_count(int64_t ncols) {
  const int64_t nb = (ncols + (1 << kBandShift) - 1) >> kBandShift;
  return (nb + kBandPad - 1) / kBandPad * kBandPad + 1;
}

void launch_heavy_slots(const int64_t* Bptr, int64_t nrows, int64_t minlen, int64_t* flag, int64_t* tmp,
                        int32_t* slot, int32_t* slot_row, hipStream_t s) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_heavy_flag, dim3(fill_grid(nrows)), dim3(256), 0, s, Bptr, nrows, minlen, flag);
  launch_exclusive_scan(flag, flag, nrows, tmp, s);
  hipLaunchKernelGGL(k_heavy_slot, dim3(fill_grid(nrows)), dim3(256), 0, s, flag, Bptr, nrows, minlen, slot, slot_row);
}

void launch_band_index(const int32_t* slot_row, int64_t nslots, const int64_t* Bptr, const int32_t* Bidx, int64_t nb1,
                       int32_t* band, hipStream_t s) {
  if (nslots <= 0) return;
  const int64_t g = nslots < 65536 ? nslots : 65536;
  hipLaunchKernelGGL(k_band_index, dim3((unsigned)g), dim3(256), 0, s, slot_row, nslots, Bptr, Bidx, nb1, band);
}

// ------------------------------------------------------------------------------------
// binning: stable counting sort of row ids by category
// ------------------------------------------------------------------------------------
constexpr int kBinTile = 256;
int64_t bin_tiles(int64_t n) { return (n + kBinTile - 1) / kBinTile; }

__device__ __forceinline__ int category_of(int64_t k, const CatTable& ct) {
  if (k <= 0) return 0;
  for (int c = 1; c < ct.n - 1; ++c)
    if (k <= ct.lim[c]) return c;
  return ct.n - 1;
}

// Category of row i: by key. Numeric tables end with the long category (limit INT64_MAX, so no
// key goes past it) and the hub category; with hub_slot, a long row that has a hub slot (a
// symbolic long row of more than kHubEntries A entries, whose range counts the symbolic pass
// wrote) is a hub row. Without hub_slot the hub category stays empty.
__device__ __forceinline__ int row_category(int64_t k, const CatTable& ct, const int32_t* __restrict__ hub_slot,
                                            int64_t i) {
  const int c = category_of(k, ct);
  if (hub_slot && c == ct.n - 2 && hub_slot[i] >= 0) return ct.n - 1;
  return c;
}

__global__ __launch_bounds__(kBinTile) void k_bin_count(const int64_t* __restrict__ key, int64_t n, int64_t cap,
                                                        CatTable ct, int64_t* __restrict__ counts, int64_t nwg,
                                                        const int32_t* __restrict__ hub_slot) {
  __shared__ int hist[kMaxCats];
  if (threadIdx.x < kMaxCats) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBinTile + threadIdx.x;
  if (i < n) {
    int64_t k = key[i];
    if (k > cap) k = cap;
    atomicAdd(&hist[row_category(k, ct, hub_slot, i)], 1);
  }
  __syncthreads();
  if ((int)threadIdx.x < ct.n) counts[(int64_t)threadIdx.x * nwg + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kBinTile) void k_bin_scatter(const int64_t* __restrict__ key, int64_t n, int64_t cap,
                                                          CatTable ct, const int64_t* __restrict__ offs, int64_t nwg,
                                                          int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ hub_slot) {
  __shared__ int wc[kBinTile / 64][kMaxCats];
  const int64_t i = (int64_t)blockIdx.x * kBinTile + threadIdx.x;
  int c = -1;
  if (i < n) {
    int64_t k = key[i];
    if (k > cap) k = cap;
    c = row_category(k, ct, hub_slot, i);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int rank = 0;
  for (int cc = 0; cc < ct.n; ++cc) {
    const uint64_t b = __ballot(c == cc);
    if (c == cc) rank = __popcll(b & lt);
    if (lane == 0) wc[wid][cc] = __popcll(b);
  }
  __syncthreads();
  if (c >= 0) {
    for (int w = 0; w < wid; ++w) rank += wc[w][c];
    rows[offs[(int64_t)c * nwg + blockIdx.x] + rank] = (int32_t)i;
  }
}

void launch_bin_count(const int64_t* key, int64_t n, int64_t cap, const CatTable& ct, int64_t* counts,
                      hipStream_t s, const int32_t* hub_slot) {
  const int64_t nwg = bin_tiles(n);
  if (nwg == 0) return;
  hipLaunchKernelGGL(k_bin_count, dim3((unsigned)nwg), dim3(kBinTile), 0, s, key, n, cap, ct, counts, nwg, hub_slot);
}
void launch_bin_scatter(const int64_t* key, int64_t n, int64_t cap, const CatTable& ct, const int64_t* offs,
                        int32_t* rows, hipStream_t s, const int32_t* hub_slot) {
  const int64_t nwg = bin_tiles(n);
  if (nwg == 0) return;
  hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)nwg), dim3(kBinTile), 0, s, key, n, cap, ct, offs, nwg, rows,
                     hub_slot);
}

// ------------------------------------------------------------------------------------
// symbolic: nnz_i by LDS hash of keys
// ------------------------------------------------------------------------------------
template <int NT>
constexpr int wg_of() { return NT <= 64 ? 256 : NT; }

template <int T, int NT, int G>
__global__ __launch_bounds__(wg_of<NT>()) void k_symbolic_hash(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
    const int32_t* __restrict__ Bidx, int64_t* __restrict__ nnz, int* __restrict__ err) {
  constexpr int WG = wg_of<NT>();
  constexpr int GPB = WG / NT;
  static_assert(NT % G == 0 && T % NT == 0, "shape");
  __shared__ int hk[GPB][T];
  __shared__ int sc[GPB][NT > 64 ? NT / 64 : 1];
  const int t = threadIdx.x % NT;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / NT);  // wave-uniform: keep row state scalar
  int* K = hk[g];
  const int64_t beg = *pbeg, end = *pend;
  if constexpr (NT == 64) {
    const int64_t s0 = beg + xcd_block() * GPB + g;
    if (s0 >= end) return;
    RowPipe<G, false, float> rp;
    rp.start(s0, (int64_t)gridDim.x * GPB, end, rows, Aptr, nullptr, t, bstart, blen, (const float*)nullptr);
    Step<4, float> st0;  // first gather step of the current row, issued during the previous row
    gather_step<G, 4, false, float>(st0, t, 0, (int)min((int64_t)64, rp.a1 - rp.a0), rp.d, Bidx,
                                    (const float*)nullptr);
    // row counts are collected one row per lane and written 64 rows at a time; error bits of
    // all rows are reported once after the loop
    int crow = 0, ccnt = 0, jn = 0, flags = 0;
    for (;;) {
      rp.prefetch(t, rows, Aptr, nullptr, bstart, blen, (const float*)nullptr);
      clear_table<T, float>(t, K, nullptr);
      wave_sync();
      HashSink<T, float, false, false> sk{K, nullptr};
      for_row_products_pre<64, G, 4, false, float>(t, rp.a0, rp.a1, rp.d, bstart, blen, (const float*)nullptr, Bidx,
                                                   (const float*)nullptr, sk, &st0);
      // the next row's first gathers (past the last row: the current row again, unused)
      gather_step<G, 4, false, float>(st0, t, 0, (int)min((int64_t)64, rp.na1 - rp.na0), rp.nd, Bidx,
                                      (const float*)nullptr);
      const int cnt = wave_reduce_dpp(sk.created, OpAdd());
      flags |= sk.bad ? kErrTableFull : 0;
      crow = (t == jn) ? rp.row : crow;
      ccnt = (t == jn) ? cnt : ccnt;
      ++jn;
      if (jn == 64 || !rp.more) {
        if (t < jn) nnz[crow] = ccnt;
        jn = 0;
      }
      wave_sync();
      if (!rp.advance()) break;
    }
    flags = wave_reduce_dpp(flags, OpOr());
    if (t == 0 && flags) atomicOr(err, flags);
    return;
  }
  for (int64_t slot = beg + (int64_t)blockIdx.x * GPB + g; slot < end; slot += (int64_t)gridDim.x * GPB) {
    const int row = rows[slot];
#pragma unroll
    for (int s = t; s < T; s += NT) K[s] = kEmpty;
    gsync<NT>();
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
    HashSink<T, float, false, false> sk{K, nullptr};
    for_row_products<NT, G, 4, false, float>(t, a0, a1, bstart, blen, (const float*)nullptr, Bidx,
                                             (const float*)nullptr, sk);
    const int cnt = Group<NT>::reduce(sk.created, OpAdd(), sc[g]);
    const int bad = Group<NT>::reduce(sk.bad, OpMax(), sc[g]);
    if (t == 0) {
      nnz[row] = cnt;
      if (bad) atomicOr(err, kErrTableFull);
    }
    gsync<NT>();
  }
}

// Per-row records of the banded long-row windows, 16 bytes each (the workspace holds 16 bytes per
// A nonzero): an entry with a banded B row (start of its B row, its slot, its index in the row),
// and one without (position from its B row's start, index in the row, next column, entries left).
struct HRec {
  int64_t b0;
  int32_t slot;
  int32_t e;
};
struct LRec {
  int32_t prel;
  int32_t e;
  int32_t nc;
  int32_t rem;
};
static_assert(sizeof(HRec) == 16 && sizeof(LRec) == 16, "records share the 16-byte workspace slots");

// Cursor of one A entry of a long row in a column sweep: position in B and entries left of
// its B row, and the column at that position (INT_MAX when the row is exhausted).
struct LongState {
  int64_t p;
  int nc;
  int rem;
};

// Lanes per A entry for a long row of nA entries in a group of NT threads: as many as divide
// the group evenly over the entries (a power of two from 4 to 64).
template <int NT>
__device__ __forceinline__ int long_lanes(int64_t nA) {
  int g = 64;
  while (g > 4 && (int64_t)(NT / g) < nA) g >>= 1;
  return g;
}


// Span walks (long-row windows, symbolic and numeric): elements per window of an A entry whose B
// row has no band index entry; the window ends before the first element such a span leaves out.
// 64 (measured against 16 and 256): B rows shorter than the band index's smallest banded length
// rarely cut a window, and a span re-reads at most 64 elements.
constexpr int kSpanCap = 64;
constexpr int kSymSpanEntries = 64;  // symbolic span walks: rows of up to one wave of entries

// Rows too long for an LDS table: sweep the row's columns in windows of W bits.
// With B's band index, rows of up to kSymSpanEntries entries (no hub slot) walk
// element-parallel: every active entry's element span in the window (banded B rows: up to the
// band boundary at or past the window end; other B rows: at most kSpanCap elements, the
// window then ending before the first element left out) is flattened by a scan in wave 0, each
// thread of the workgroup takes U elements at a time, and the entry state (B row start and
// length, elements consumed, next column, band slot) lives in LDS. (A workgroup-wide version for
// rows of up to NT entries measured slower than the cursor walks below: its scans cost more
// barriers per window than the walks' chains.)
// Otherwise one cursor per A entry (cursor[] is indexed like A's nonzeros) walks its sorted B
// row, whose extent comes from the row analysis' descriptors; an entry with more elements in the
// window than one step of G continues UL steps at a time, their loads issued together.
template <int NT, int W, int G>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_symbolic_long(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
    const int32_t* __restrict__ Bidx, int64_t* __restrict__ nnz, int64_t* __restrict__ cursor,
    int* __restrict__ err, const int32_t* __restrict__ hub_slot, int32_t* __restrict__ rng_cnt, int nrng,
    const int32_t* __restrict__ Aidx, const int32_t* __restrict__ hslot, const int32_t* __restrict__ band,
    int64_t nb1) {
  constexpr int NWORD = W / 32;
  constexpr int WPT = NWORD / NT;
  constexpr int UL = 8;
  constexpr int U = 4;
  constexpr int RPW = W >> kHubRangeShift;          // hub ranges per window
  constexpr int RWORD = NWORD / RPW;                // bitmap words per hub range
  static_assert(NWORD % NT == 0 && W % kHubRange == 0 && NT % 64 == 0, "shape");
  __shared__ uint32_t bm[NWORD];
  __shared__ int sc[NT / 64 > 0 ? NT / 64 : 1];
  __shared__ long long scl[NT / 64 > 0 ? NT / 64 : 1];
  __shared__ int rc[RPW];
  // span-walk entry state and window tables (entry e in slot e)
  __shared__ int64_t eB0[kSymSpanEntries];
  __shared__ int eL[kSymSpanEntries], eP[kSymSpanEntries], eNC[kSymSpanEntries], eSL[kSymSpanEntries];
  __shared__ int sS[kSymSpanEntries], sN[kSymSpanEntries];
  __shared__ int s_T, s_lo;
  __shared__ long long s_hi;
  const int t = threadIdx.x, lg = t % G, lane = t & 63;
  const int64_t beg = *pbeg, end = *pend;
  for (int64_t slot = beg + blockIdx.x; slot < end; slot += gridDim.x) {
    const int row = rows[slot];
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
    // rows with a hub slot take windows aligned to W and record their columns per hub range
    const int hs = hub_slot ? hub_slot[row] : -1;
#pragma unroll
    for (int q = 0; q < WPT; ++q) bm[t + q * NT] = 0u;
    const int nA = (int)min(a1 - a0, (int64_t)kSymSpanEntries + 1);
    if (band != nullptr && hs < 0 && nA <= kSymSpanEntries) {  // uniform
      // ================= span walks (entries in slots 0 .. nA - 1, wave 0) =================
      long long prow_l = 0;
      if (t < nA) {
        const int64_t ka = a0 + t;
        const int64_t b0 = bstart[ka];
        const int L = blen[ka];
        eB0[t] = b0;
        eL[t] = L;
        eP[t] = 0;
        eSL[t] = hslot[Aidx[ka]];
        eNC[t] = L > 0 ? Bidx[b0] : INT_MAX;
        prow_l = L;
      }
      if (t < 64) {
        const int m = wave_reduce_dpp(t < nA ? eNC[t] : INT_MAX, OpMin());
        if (t == 0) s_lo = m;
      }
      const long long prow = Group<NT>::reduce(prow_l, OpAdd(), scl);  // also orders the writes
      int lo = s_lo;
      long long count = 0;
      long long windows = 0;
      while (lo != INT_MAX) {
        // ---- the active entries' element spans; the window ends before any element a capped
        // span leaves out ----
        if (t < 64) {
          const int64_t hi0 = min((int64_t)lo + W, (int64_t)INT_MAX);
          int len = 0, clip = INT_MAX;
          if (t < nA && (int64_t)eNC[t] < hi0) {
            const int p = eP[t], L = eL[t], sl = eSL[t];
            int q;
            if (sl >= 0) {
              q = band[(int64_t)sl * nb1 + min((hi0 + (1 << kBandShift) - 1) >> kBandShift, nb1 - 1)];
            } else {
              q = min(L, p + kSpanCap);
              if (q < L) clip = Bidx[eB0[t] + q];
            }
            len = q - p;
          }
          const int cm = wave_reduce_dpp(clip, OpMin());
          const int incl = wave_incl_scan_dpp(len);
          if (t < nA) {
            sS[t] = incl - len;
            sN[t] = 0;
          }
          if (t == 63) {
            s_T = incl;
            s_hi = min(hi0, (int64_t)cm);
          }
        }
        __syncthreads();
        const int T = s_T;
        const int64_t hi = s_hi;
        // ---- presence bits, element-parallel ----
        for (int f0 = 0; f0 < T; f0 += NT * U) {  // uniform
          int c[U], jj[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int f = f0 + u * NT + t;
            int j = 0;  // the last entry with sS[j] <= f
#pragma unroll
            for (int step = kSymSpanEntries / 2; step >= 1; step >>= 1)
              if (j + step < nA && sS[j + step] <= f) j += step;
            jj[u] = j;
            const int64_t at = f < T ? eB0[j] + eP[j] + (f - sS[j]) : 0;  // lanes past T: unused read
            c[u] = Bidx[at];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int f = f0 + u * NT + t;
            const int j = jj[u];
            const bool in = f < T && (int64_t)c[u] < hi;
            if (in) {
              const int o = c[u] - lo;
              atomicOr(&bm[o >> 5], 1u << (o & 31));
            }
            // an entry's elements below hi are a prefix of its span: the last such lane of each
            // wave records its length
            const int jn = __shfl_down(j, 1);
            const int inn = __shfl_down((int)in, 1);
            if (in && (lane == 63 || !inn || jn != j)) atomicMax(&sN[j], f - sS[j] + 1);
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
          count += __popc(bm[t + q * NT]);
          bm[t + q * NT] = 0u;
        }
        // ---- cursors advance past the window; the next window starts at the smallest next column ----
        if (t < 64) {
          int nl = INT_MAX;
          if (t < nA) {
            const int took = sN[t];
            nl = eNC[t];
            if (took > 0) {
              const int p = eP[t] + took;
              eP[t] = p;
              nl = p < eL[t] ? Bidx[eB0[t] + p] : INT_MAX;
              eNC[t] = nl;
            }
          }
          const int m = wave_reduce_dpp(nl, OpMin());
          if (t == 0) s_lo = m;
        }
        __syncthreads();  // also: the bitmap is clear
        lo = s_lo;
        if (++windows > prow) {  // uniform: no progress; reported, the row's count is not trusted
          if (t == 0) atomicOr(err, kErrWindow);
          break;
        }
      }
      count = Group<NT>::reduce(count, OpAdd(), scl);
      if (t == 0) nnz[row] = count;
      __syncthreads();
      continue;
    }
    int lo_l = INT_MAX;
    long long prow_l = 0;
    for (int64_t ka = a0 + t; ka < a1; ka += NT) {
      const int64_t b0 = bstart[ka];
      cursor[ka] = b0;
      prow_l += blen[ka];
      if (blen[ka] > 0) lo_l = min(lo_l, Bidx[b0]);
    }
    // every window takes at least the element at its first column, so a row of P_i products
    // needs at most P_i windows; more would mean a walk that stopped making progress
    const long long prow = Group<NT>::reduce(prow_l, OpAdd(), scl);
    int lo = Group<NT>::reduce(lo_l, OpMin(), sc);  // also orders the cursor writes
    long long count = 0;
    long long windows = 0;
    while (lo != INT_MAX) {
      if (hs >= 0) lo &= ~(W - 1);  // uniform
      const int64_t hi = min((int64_t)lo + W, (int64_t)INT_MAX);  // INT_MAX: past the B row
      for (int64_t ka = a0 + t / G; ka < a1; ka += NT / G) {
        int64_t p = cursor[ka];
        const int64_t b1 = bstart[ka] + blen[ka];
        int nin;
        {
          const int64_t q = p + lg;
          const int c = (q < b1) ? Bidx[q] : INT_MAX;
          const bool in = (int64_t)c < hi;
          if (in) {
            const int o = c - lo;
            atomicOr(&bm[o >> 5], 1u << (o & 31));
          }
          nin = __popcll(__ballot(in) & subgroup_mask<G>());
          p += nin;
        }
        // B rows with more elements in the window than one step: UL steps at a time. p grows by
        // G per full step and never passes b1 on this path, so the walk ends within len / G steps
        // whatever the sentinel column past the row compares as.
        while (nin == G && p < b1) {  // subgroup-uniform
          int cc[UL];
#pragma unroll
          for (int k = 0; k < UL; ++k) {
            const int64_t q = p + k * G + lg;
            cc[k] = (q < b1) ? Bidx[q] : INT_MAX;
          }
#pragma unroll
          for (int k = 0; k < UL; ++k) {
            if (nin == G) {
              const bool in = (int64_t)cc[k] < hi;
              if (in) {
                const int o = cc[k] - lo;
                atomicOr(&bm[o >> 5], 1u << (o & 31));
              }
              nin = __popcll(__ballot(in) & subgroup_mask<G>());
              p += nin;
            }
          }
        }
        if (lg == 0) cursor[ka] = p;
      }
      if (hs >= 0 && t < RPW) rc[t] = 0;
      __syncthreads();
      long long pc = 0;
#pragma unroll
      for (int q = 0; q < WPT; ++q) {
        const int wq = __popc(bm[t + q * NT]);
        pc += wq;
        if (hs >= 0 && wq) atomicAdd(&rc[(t + q * NT) / RWORD], wq);
        bm[t + q * NT] = 0u;
      }
      count += pc;
      if (hs >= 0) {  // uniform
        __syncthreads();
        const int r0 = lo >> kHubRangeShift;
        if (t < RPW && r0 + t < nrng) rng_cnt[(int64_t)hs * nrng + r0 + t] = rc[t];
      }
      int nl = INT_MAX;
      for (int64_t ka = a0 + t; ka < a1; ka += NT) {
        const int64_t p = cursor[ka];
        if (p < bstart[ka] + blen[ka]) nl = min(nl, Bidx[p]);
      }
      lo = Group<NT>::reduce(nl, OpMin(), sc);
      if (++windows > prow) {  // uniform: no progress; reported, the row's count is not trusted
        if (t == 0) atomicOr(err, kErrWindow);
        break;
      }
    }
    count = Group<NT>::reduce(count, OpAdd(), scl);
    if (t == 0) nnz[row] = count;
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// hub rows: rows with more A entries than the window kernels hold (kHubEntries)
// ------------------------------------------------------------------------------------
// A numeric window kernel visits every A entry of its row in every dense window of 4096 columns
// it sweeps, so a row's cost grows as nA * ncols / 4096 beside its products: over a few million
// columns a row of thousands of entries spends more on those visits than on its products, and
// the whole row is one workgroup's work. Hub rows are cut into ranges of kHubRange columns
// instead, one workgroup per (row, range), each accumulating its range in LDS (a presence bitmap
// and, for the numeric pass, dense values), so a heavy row spreads over the GPU and no product
// needs a global atomic:
//   - entries whose B row is in the band index (long B rows) read their range's element span
//     from two index entries; `k_hub_prepare` lists them per row (in A's order) and counts the
//     products of the other (short) entries per range; `k_hub_scatter` writes those products,
//     (column, a_ik * b_kj), into per-(row, range) buckets;
//   - the symbolic long kernel counts each range's columns (hub rows' windows are aligned to
//     whole ranges), a scan turns the counts into output offsets, and `k_hub_range` accumulates
//     the range and emits it in column order at its offset.
// Within a bucket the products' order follows LDS atomics, so sums of a column may associate
// differently from run to run (within the usual rounding bound).
constexpr int kHubRangeBands = kHubRange >> kBandShift;

template <int NT>
__global__ __launch_bounds__(NT) void k_hub_prepare(const int32_t* __restrict__ rows, const int64_t* __restrict__ Aptr,
                                                    const int32_t* __restrict__ Aidx,
                                                    const int64_t* __restrict__ bstart,
                                                    const int32_t* __restrict__ blen, const int32_t* __restrict__ Bidx,
                                                    const int32_t* __restrict__ hslot, int nrng,
                                                    int64_t* __restrict__ scnt, int32_t* __restrict__ hlist,
                                                    int32_t* __restrict__ hcount) {
  extern __shared__ int lcnt[];  // nrng counters
  __shared__ int sc[NT / 64];
  const int t = threadIdx.x;
  const int64_t j = blockIdx.x;
  const int row = rows[j];
  const int64_t a0 = Aptr[row], nA = Aptr[row + 1] - a0;
  for (int r = t; r < nrng; r += NT) lcnt[r] = 0;
  __syncthreads();
  int nH = 0;  // banded entries listed so far (uniform)
  for (int64_t e0 = 0; e0 < nA; e0 += NT) {  // uniform
    const int64_t e = e0 + t;
    int banded = 0;
    if (e < nA) {
      const int64_t ka = a0 + e;
      if (hslot[Aidx[ka]] >= 0) {
        banded = 1;
      } else {
        const int64_t b0 = bstart[ka];
        const int L = blen[ka];
        for (int q = 0; q < L; ++q) atomicAdd(&lcnt[Bidx[b0 + q] >> kHubRangeShift], 1);
      }
    }
    int tot;
    const int pre = Group<NT>::exscan(banded, tot, sc);
    if (banded) hlist[a0 + nH + pre] = (int32_t)e;
    nH += tot;
  }
  __syncthreads();
  for (int r = t; r < nrng; r += NT) scnt[j * nrng + r] = lcnt[r];
  if (t == 0) hcount[j] = nH;
}

template <typename V, int NT>
__global__ __launch_bounds__(NT) void k_hub_scatter(const int32_t* __restrict__ rows, const int64_t* __restrict__ Aptr,
                                                    const int32_t* __restrict__ Aidx,
                                                    const int64_t* __restrict__ bstart,
                                                    const int32_t* __restrict__ blen, const V* __restrict__ Aval,
                                                    const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
                                                    const int32_t* __restrict__ hslot, int nrng,
                                                    const int64_t* __restrict__ soff, int32_t* __restrict__ scol,
                                                    V* __restrict__ sval, const int* __restrict__ err) {
  extern __shared__ unsigned long long lcur[];  // nrng write positions
  if (*err & kErrCapacity) return;  // the buckets (or C) are smaller than this structure needs
  const int t = threadIdx.x;
  const int64_t j = blockIdx.x;
  const int row = rows[j];
  const int64_t a0 = Aptr[row], nA = Aptr[row + 1] - a0;
  for (int r = t; r < nrng; r += NT) lcur[r] = (unsigned long long)soff[j * nrng + r];
  __syncthreads();
  for (int64_t e = t; e < nA; e += NT) {
    const int64_t ka = a0 + e;
    if (hslot[Aidx[ka]] >= 0) continue;
    const int64_t b0 = bstart[ka];
    const int L = blen[ka];
    const V av = Aval[ka];
    for (int q = 0; q < L; ++q) {
      const int col = Bidx[b0 + q];
      const int64_t pos = (int64_t)atomicAdd(&lcur[col >> kHubRangeShift], 1ull);
      scol[pos] = col;
      sval[pos] = av * Bval[b0 + q];
    }
  }
}

// One (row, range) item per workgroup: the range's values, emitted at Cptr[row] - cbase +
// roff[item] - roff[first item of the row], with roff[item + 1] - roff[item] columns expected.
template <typename V, int NT>
__global__ __launch_bounds__(NT) void k_hub_range(
    const int32_t* __restrict__ rows, int nrng, const int64_t* __restrict__ Aptr, const int32_t* __restrict__ Aidx,
    const int64_t* __restrict__ bstart, const V* __restrict__ Aval, const int32_t* __restrict__ Bidx,
    const V* __restrict__ Bval, const int32_t* __restrict__ hslot, const int32_t* __restrict__ band, int64_t nb1,
    const int32_t* __restrict__ hlist, const int32_t* __restrict__ hcount, const int64_t* __restrict__ soff,
    const int32_t* __restrict__ scol, const V* __restrict__ sval,
    const int64_t* __restrict__ roff, const int64_t* __restrict__ Cptr, int64_t cbase, int32_t* __restrict__ Cidx,
    V* __restrict__ Cval, int* __restrict__ err) {
  constexpr int NWORD = kHubRange / 32;
  constexpr int NW = NT / 64;
  constexpr int U = 4;
  static_assert(NWORD <= NT, "one bitmap word per thread at most");
  __shared__ uint32_t bm[NWORD];
  __shared__ V dv[kHubRange];
  __shared__ int64_t hb0[NT];
  __shared__ V hva[NT];
  __shared__ int hS[NT];
  __shared__ int sc[NW];
  if (*err & kErrCapacity) return;  // the buckets or C are smaller than this structure needs
  const int t = threadIdx.x, lane = t & 63;
  const int64_t item = blockIdx.x;
  // a range without columns (by the symbolic count) holds no product: nothing to add or emit
  if (roff[item + 1] == roff[item]) return;
  const int64_t j = item / nrng;
  const int r = (int)(item - j * nrng);
  const int row = rows[j];
  const int64_t a0 = Aptr[row];
  const int nH = hcount[j];
  const int64_t c0 = (int64_t)r << kHubRangeShift;
  const int64_t w0 = c0 >> kBandShift;
  const int64_t w1 = min(w0 + kHubRangeBands, nb1 - 1);
  for (int q = t; q < NWORD; q += NT) bm[q] = 0u;
  for (int q = t; q < kHubRange; q += NT) dv[q] = V(0);
  __syncthreads();
  // products of the banded entries in this range: chunks of NT entries, each wave taking 64 of
  // them (strided, so the heaviest entries, which lie together in A's order, spread over the
  // waves) and gathering its entries' spans element-parallel on its own: a wave scan flattens
  // them and an element's owner is searched among 64 entries, with no workgroup barrier
  const int wv = t >> 6;
  int* whS = hS + (t & ~63);
  const int64_t* whb0 = hb0 + (t & ~63);
  const V* whva = hva + (t & ~63);
  for (int e0 = 0; e0 < nH; e0 += NT) {  // uniform
    const int e = e0 + wv + (NT / 64) * lane;
    int len = 0;
    if (e < nH) {
      const int64_t ka = a0 + hlist[a0 + e];
      const int32_t* bd = band + (int64_t)hslot[Aidx[ka]] * nb1;
      const int s0 = bd[w0];
      len = bd[w1] - s0;
      hb0[t] = bstart[ka] + s0;
      hva[t] = Aval[ka];
    }
    const int incl = wave_incl_scan_dpp(len);
    hS[t] = incl - len;  // lanes past the row's entries hold the wave total: never an owner
    const int T = __builtin_amdgcn_readlane(incl, 63);
    wave_sync();
    for (int f0 = 0; f0 < T; f0 += 64 * U) {  // wave-uniform
      int col[U];
      V bv[U], va[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int f = f0 + u * 64 + lane;
        int jj = 0;  // largest jj with whS[jj] <= f
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
          if (whS[jj + step] <= f) jj += step;  // jj + step <= 63
        const int64_t at = f < T ? whb0[jj] + (f - whS[jj]) : 0;  // lanes past T: unused read of B's first entry
        col[u] = Bidx[at];
        bv[u] = Bval[at];
        va[u] = whva[jj];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (f0 + u * 64 + lane < T) {
          const int o = (int)(col[u] - c0);
          atomicOr(&bm[o >> 5], 1u << (o & 31));
          lds_add(&dv[o], va[u] * bv[u]);
        }
    }
    wave_sync();  // this wave's tables are rewritten by its next chunk
  }
  // products of the short entries, from the range's bucket
  const int64_t q0 = soff[item], q1 = soff[item + 1];
  for (int64_t q = q0 + t; q < q1; q += NT) {
    const int o = (int)(scol[q] - c0);
    atomicOr(&bm[o >> 5], 1u << (o & 31));
    lds_add(&dv[o], sval[q]);
  }
  __syncthreads();
  const uint32_t w = t < NWORD ? bm[t] : 0u;
  const int pc = __popc(w);
  int tot;
  const int pre = Group<NT>::exscan(pc, tot, sc);
  const int64_t base = roff[item] - roff[j * nrng];
  const int64_t out = Cptr[row] - cbase + base;
  const int64_t c1 = Cptr[row + 1] - cbase;
  int flags = (roff[item + 1] - roff[item] != tot) ? kErrWindow : 0;
  uint32_t bits = w;
  int64_t pos = out + pre;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1u;
    const int o = 32 * t + b;
    if (pos < c1) {
      Cidx[pos] = (int)(c0 + o);
      Cval[pos] = dv[o];
    } else {
      flags |= kErrWindow;
    }
    ++pos;
  }
  flags = wave_reduce_dpp(flags, OpOr());
  if ((t & 63) == 0 && flags) atomicOr(err, flags);
}

// ------------------------------------------------------------------------------------
// numeric: accumulate in an LDS hash, sort by a monotone bucket map + in-bucket rank,
// emit coalesced
// ------------------------------------------------------------------------------------
// Rank distinct keys held in registers (IPT per lane, kEmpty = none, at most T/2 in all) in
// column order, using the T ints at K (free for the call) for the bucket counters (lower half)
// and the bucket permutation (upper half). The bucket map b(x) = floor(x * mul / 2^32),
// x = c - cmin (or b = x when the column range fits the buckets) is monotone, so a key's rank is
// the start of its bucket plus the number of smaller keys in the bucket; mul only has to keep b
// below NB, so it is formed in single precision and rounded down. LDS reads of all keys are
// issued before their results are used. rk[q] = rank of kk[q], -1 if empty. Returns the number
// of keys (more than T/2 means a table was overfilled; ranks are then not exact and the caller
// reports it).
template <int T, int IPT>
__device__ __forceinline__ int wave_rank_keys(int t, int* K, int cmin, int cmax, const int (&kk)[IPT],
                                              int (&rk)[IPT]) {
  constexpr int H2 = T / 2;
  constexpr int NB = H2 - 1;           // buckets; BC[NB] receives the total
  constexpr int PER = (H2 + 63) / 64;  // counters per lane in the scan
  static_assert(NB < 65536 && T <= 65536, "bucket and in-bucket index share one int");
  const uint32_t range = (uint32_t)(cmax - cmin) + 1u;
  const bool direct = range <= (uint32_t)NB;
  const uint32_t mul =
      direct ? 0u : (uint32_t)((float)NB * 4294967296.0f / (float)range * 0.99999f);
  int* BC = K;
  int* perm = K + H2;
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (t * PER + q < H2) BC[t * PER + q] = 0;
  wave_sync();
  int bi[IPT];  // bucket | (arrival order in the bucket) << 16
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t x = (uint32_t)(kk[q] - cmin);
    const int b = direct ? (int)x : (int)__umulhi(x, mul);
    bi[q] = b;
    if (kk[q] != kEmpty) bi[q] = b | (atomicAdd(&BC[b], 1) << 16);
  }
  wave_sync();
  int total;
  {
    int loc[PER];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = t * PER + q;
      loc[q] = (i < NB) ? BC[i] : 0;
      sum += loc[q];
    }
    const int incl = wave_incl_scan_dpp(sum);
    total = __builtin_amdgcn_readlane(incl, 63);
    int pre = incl - sum;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = t * PER + q;
      if (i < NB) BC[i] = pre;
      pre += loc[q];
    }
    if (t == 0) BC[NB] = total;
  }
  wave_sync();
  // bucket start and length of every live slot (reads first, then the permutation writes)
  int sl[IPT];  // start | length << 16
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    sl[q] = 0;
    if (kk[q] != kEmpty) {
      const int b = bi[q] & 0xFFFF;
      const int s0 = BC[b], s1 = min(BC[b + 1], H2);
      sl[q] = s0 | (max(s1 - s0, 0) << 16);
    }
  }
#pragma unroll
  for (int q = 0; q < IPT; ++q)
    if (kk[q] != kEmpty) {
      const int pos = (sl[q] & 0xFFFF) + (bi[q] >> 16);
      if (pos < H2) perm[pos] = kk[q];
    }
  wave_sync();
  // in-bucket rank: only buckets holding more than one key need comparisons; position j of
  // every such bucket is read for all slots at once
#pragma unroll
  for (int q = 0; q < IPT; ++q) rk[q] = (kk[q] != kEmpty) ? (sl[q] & 0xFFFF) : -1;
  for (int j = 0;; ++j) {
    bool more = false;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const int len = sl[q] >> 16;
      if (len > 1 && j < len) rk[q] += (perm[(sl[q] & 0xFFFF) + j] < kk[q]);
      more |= (len > 1 && j + 1 < len);
    }
    if (!__builtin_amdgcn_ballot_w64(more)) break;
  }
  return total;
}

// Sort the live entries of a group's row table (T slots, NT > 64 threads, at most T/2 live)
// by column and write them to Cidx/Cval[0 .. min(n, limit)); returns n. Buckets by the
// monotone map b(c) = floor((c - cmin) * NB / range), NB = T/2 counters in BC (NB + 1 ints),
// then the rank inside a bucket by comparison. The table's key and value LDS are reused.
// Leaves the table (K, H) unspecified and BC dirty; the caller clears them.
template <typename V, int T, int NT>
__device__ __forceinline__ int block_rank_emit(int t, int* K, V* H, int* BC, int* sc, int cmin, int cmax,
                                               int32_t* __restrict__ Cidx, V* __restrict__ Cval, int64_t limit) {
  constexpr int IPT = T / NT;
  constexpr int NB = T / 2;
  constexpr int PER = (NB + NT - 1) / NT;
  for (int s2 = t; s2 <= NB; s2 += NT) BC[s2] = 0;
  gsync<NT>();
  const uint32_t range = (uint32_t)(cmax - cmin) + 1u;
  const uint64_t mul = ((uint64_t)NB << 32) / (uint64_t)range;
  int ik[IPT], ib[IPT], io[IPT];
  V iv[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int s2 = t + j * NT;
    const int key = K[s2];
    ik[j] = key;
    iv[j] = H[s2];
    ib[j] = 0;
    io[j] = 0;
    if (key != kEmpty) {
      const int b = (int)(((uint64_t)(uint32_t)(key - cmin) * mul) >> 32);
      ib[j] = b;
      io[j] = atomicAdd(&BC[b], 1);
    }
  }
  gsync<NT>();
  {  // exclusive scan of bucket counts, PER contiguous buckets per thread
    int loc[PER];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = t * PER + q;
      loc[q] = (i < NB) ? BC[i] : 0;
      sum += loc[q];
    }
    int total;
    int pre = Group<NT>::exscan(sum, total, sc);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = t * PER + q;
      if (i < NB) BC[i] = pre;
      pre += loc[q];
    }
    if (t == 0) BC[NB] = total;
  }
  gsync<NT>();
  int* perm = K;  // the table keys are in registers now; reuse their LDS
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (ik[j] != kEmpty) perm[BC[ib[j]] + io[j]] = ik[j];
  gsync<NT>();
  int rk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    rk[j] = 0;
    if (ik[j] != kEmpty) {
      const int s0 = BC[ib[j]], s1 = BC[ib[j] + 1];
      int r = s0;
      for (int m = s0; m < s1; ++m) r += (perm[m] < ik[j]);
      rk[j] = r;
    }
  }
  gsync<NT>();
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (ik[j] != kEmpty) {
      K[rk[j]] = ik[j];
      H[rk[j]] = iv[j];
    }
  gsync<NT>();
  const int n = BC[NB];
  const int nw = (int)min((int64_t)n, limit);
  for (int m = t; m < nw; m += NT) {
    Cidx[m] = K[m];
    Cval[m] = H[m];
  }
  gsync<NT>();
  return n;
}

template <typename V, int T, int NT, int G, bool FAST = false>
__device__ __forceinline__ void numeric_hash_body(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
    const V* __restrict__ Aval, const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
    const int64_t* __restrict__ Cptr, int64_t cbase, int32_t* __restrict__ Cidx, V* __restrict__ Cval,
    int* __restrict__ err) {
  constexpr int WG = wg_of<NT>();
  constexpr int GPB = WG / NT;
  constexpr int IPT = T / NT;  // table slots per thread
  static_assert(NT % G == 0 && T % NT == 0, "shape");
  constexpr bool WAVE = (NT == 64);
  constexpr int NB = T / 2;  // buckets (workgroup path)
  __shared__ int hk[GPB][T];
  __shared__ V hv[GPB][T];
  __shared__ int bc[GPB][WAVE ? 1 : NB + 1];
  __shared__ int sc[GPB][NT > 64 ? NT / 64 : 1];
  const int t = threadIdx.x % NT;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / NT);  // wave-uniform
  int* K = hk[g];
  V* H = hv[g];
  const int64_t beg = *pbeg, end = *pend;
  if constexpr (WAVE) {
    // one wave per row, rows software-pipelined (RowPipe); the table is ranked by
    // wave_rank_keys while the next row's descriptors are in flight
    const int64_t s0 = beg + xcd_block() * GPB + g;
    if (s0 >= end) return;
    RowPipe<G, true, V> rp;
    rp.start(s0, (int64_t)gridDim.x * GPB, end, rows, Aptr, Cptr, t, bstart, blen, Aval);
    Step<4, V> st0;  // first gather step of the current row, issued during the previous row
    gather_step<G, 4, true, V>(st0, t, 0, (int)min((int64_t)64, rp.a1 - rp.a0), rp.d, Bidx, Bval);
    int flags = 0;   // error bits of all rows of this wave, reported once after the loop
    constexpr int STEP = G == 16 ? 16 : 4;  // A entries covered by the first gather step
    for (;;) {
      rp.prefetch(t, rows, Aptr, Cptr, bstart, blen, Aval);
      // Rows whose products are all distinct (P_i equals the symbolic count nnz_i) and are all
      // in the first gather step (at most STEP A entries, B rows of at most G entries) skip the
      // hash table: their products are ranked straight from registers.
      bool fast = false;
      if (FAST && rp.a1 - rp.a0 <= STEP) {
        const bool inrow = desc_slot<G>(t) < rp.a1 - rp.a0;
        const bool longb = __builtin_amdgcn_ballot_w64(inrow && rp.d.len > G) != 0;
        const int P = wave_reduce_dpp(inrow ? min(rp.d.len, G + 1) : 0, OpAdd());
        fast = !longb && (int64_t)P == rp.c1 - rp.c0 && P <= T / 2;
      }
      if (fast) {
        const int64_t nA = rp.a1 - rp.a0;
        const int64_t nexp = rp.c1 - rp.c0;
        const int sg = t / G, lg = t % G;
        const int cnt = (int)nA;
        int kc[4], rk[4];
        V kv[4];
        int lmin = INT_MAX, lmax = -1;
        static_for<4>([&](auto uc) {
          constexpr int u = decltype(uc)::value;
          const int e = step_entry<G, 0>(0, u, sg);
          const int lv = ebcast<G, u>(rp.d.len, u);
          const V pv = ebcastv<G, u>(rp.d.va, u);
          const bool has = e < cnt && lg < lv;
          kc[u] = has ? st0.c[u] : kEmpty;
          kv[u] = pv * st0.bv[u];
          lmin = has ? min(lmin, kc[u]) : lmin;
          lmax = has ? max(lmax, kc[u]) : lmax;
        });
        // the next row's first gathers, in flight while this row is ranked and stored
        gather_step<G, 4, true, V>(st0, t, 0, (int)min((int64_t)64, rp.na1 - rp.na0), rp.nd, Bidx, Bval);
        const int cmin = wave_reduce_dpp(lmin, OpMin());
        const int cmax = wave_reduce_dpp(lmax, OpMax());
        const int n = wave_rank_keys<T, 4>(t, K, cmin, cmax, kc, rk);
        flags |= ((int64_t)n != nexp) ? kErrCountMismatch : 0;
        wave_sync();
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (rk[u] >= 0) {
            K[rk[u]] = kc[u];
            H[rk[u]] = kv[u];
          }
        wave_sync();
        // fixed-count unconditional coalesced stores, as below
        const int64_t o0 = rp.c0 - cbase;
        const int nw = (int)min((int64_t)n, nexp);
        int32_t* di = nw > 0 ? Cidx + o0 : (int32_t*)(err + 4);
        V* dv = nw > 0 ? Cval + o0 : (V*)(err + 8);
        const int last = max(nw, 1) - 1;
#pragma unroll
        for (int j = 0; j < (T / 2 + 63) / 64; ++j) {
          const int m = min(t + 64 * j, last);
          const int key = K[m];
          const V val = H[m];
          di[m] = key;
          dv[m] = val;
        }
        wave_sync();
        if (!rp.advance()) break;
        continue;
      }
      clear_table<T, V>(t, K, H);
      wave_sync();
      HashSink<T, V, true> sk{K, H};
      for_row_products_pre<64, G, 4, true, V>(t, rp.a0, rp.a1, rp.d, bstart, blen, Aval, Bidx, Bval, sk, &st0);
      // the next row's first gathers (past the last row: the current row again, unused); they
      // are in flight while this row is ranked and stored
      gather_step<G, 4, true, V>(st0, t, 0, (int)min((int64_t)64, rp.na1 - rp.na0), rp.nd, Bidx, Bval);
      wave_sync();
      const int cmin = wave_reduce_dpp(sk.cmin, OpMin());
      const int cmax = wave_reduce_dpp(sk.cmax, OpMax());
      int kk[IPT], rk[IPT];
#pragma unroll
      for (int q = 0; q < IPT; ++q) kk[q] = K[slot_of<T>(t, q)];
      wave_sync();
      const int n = wave_rank_keys<T, IPT>(t, K, cmin, cmax, kk, rk);
      const int64_t o0 = rp.c0 - cbase;
      const int64_t nexp = rp.c1 - rp.c0;
      flags |= (sk.bad || n > T / 2) ? kErrTableFull : 0;
      flags |= ((int64_t)n != nexp) ? kErrCountMismatch : 0;
      // stage the ranked row in the (now free) key table: key at its rank, source slot at
      // T/2 + rank; the values stay in their hash slots
      wave_sync();
#pragma unroll
      for (int q = 0; q < IPT; ++q)
        if (rk[q] >= 0 && rk[q] < T / 2) {
          K[rk[q]] = kk[q];
          K[T / 2 + rk[q]] = slot_of<T>(t, q);
        }
      wave_sync();
      // a fixed number of unconditional, coalesced stores: lanes past the row's end repeat its
      // last element (same address, same data). Unconditional stores keep the compiler's count
      // of outstanding memory operations exact, so waiting for the next row's gathers never
      // waits for these stores. A row that came out empty (only after an inconsistency, which
      // is reported) writes to the sink behind the error word instead of the output.
      const int nw = (int)min((int64_t)min(n, T / 2), nexp);
      int32_t* di = nw > 0 ? Cidx + o0 : (int32_t*)(err + 4);
      V* dv = nw > 0 ? Cval + o0 : (V*)(err + 8);
      const int last = max(nw, 1) - 1;
#pragma unroll
      for (int j = 0; j < (T / 2 + 63) / 64; ++j) {
        const int m = min(t + 64 * j, last);
        const int key = K[m];
        const V val = H[K[T / 2 + m] & (T - 1)];
        di[m] = key;
        dv[m] = val;
      }
      wave_sync();
      if (!rp.advance()) break;
    }
    flags = wave_reduce_dpp(flags, OpOr());
    if (t == 0 && flags) atomicOr(err, flags);
    return;
  } else {
    int* BC = bc[g];
    for (int64_t slot = beg + (int64_t)blockIdx.x * GPB + g; slot < end; slot += (int64_t)gridDim.x * GPB) {
      const int row = rows[slot];
#pragma unroll
      for (int s2 = t; s2 < T; s2 += NT) {
        K[s2] = kEmpty;
        H[s2] = V(0);
      }
      gsync<NT>();
      const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
      HashSink<T, V, true> sk{K, H};
      for_row_products<NT, G, 4, true, V>(t, a0, a1, bstart, blen, Aval, Bidx, Bval, sk);
      gsync<NT>();
      const int cmin = Group<NT>::reduce(sk.cmin, OpMin(), sc[g]);
      const int cmax = Group<NT>::reduce(sk.cmax, OpMax(), sc[g]);
      int bad = sk.bad;
      const int64_t o0 = Cptr[row] - cbase;
      const int64_t nexp = Cptr[row + 1] - Cptr[row];
      const int n = block_rank_emit<V, T, NT>(t, K, H, BC, sc[g], cmin, cmax, Cidx + o0, Cval + o0, nexp);
      bad = Group<NT>::reduce(bad, OpMax(), sc[g]);
      if (t == 0) {
        if (bad) atomicOr(err, kErrTableFull);
        if ((int64_t)n != nexp) atomicOr(err, kErrCountMismatch);
      }
      gsync<NT>();
    }
  }
}

// Occupancy target of the one-wave-per-row numeric kernel (waves per SIMD); the compiler fits
// its registers to it. The workgroup-per-row shapes are limited by LDS instead. The variant with
// the register path for rows of distinct products needs more registers: 4 waves per SIMD.
#ifndef DBD_NUMERIC_WAVES_PER_EU
#define DBD_NUMERIC_WAVES_PER_EU 5
#endif
#ifndef DBD_NUMERIC_FAST_WAVES_PER_EU
#define DBD_NUMERIC_FAST_WAVES_PER_EU 4
#endif

#define DBD_NUMERIC_ARGS                                                                                     \
  const int32_t *__restrict__ rows, const int64_t *__restrict__ pbeg, const int64_t *__restrict__ pend,     \
      const int64_t *__restrict__ Aptr, const int64_t *__restrict__ bstart, const int32_t *__restrict__ blen, \
      const V *__restrict__ Aval, const int32_t *__restrict__ Bidx, const V *__restrict__ Bval,              \
      const int64_t *__restrict__ Cptr, int64_t cbase, int32_t *__restrict__ Cidx, V *__restrict__ Cval,      \
      int *__restrict__ err

template <typename V, int T, int NT, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DBD_NUMERIC_WAVES_PER_EU))) void
k_numeric_wave(DBD_NUMERIC_ARGS) {
  numeric_hash_body<V, T, NT, G>(rows, pbeg, pend, Aptr, bstart, blen, Aval, Bidx, Bval, Cptr, cbase, Cidx, Cval,
                                 err);
}

// Same rows, with the register path for rows whose products are all distinct (chosen when the
// problem has few duplicate products).
template <typename V, int T, int NT, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DBD_NUMERIC_FAST_WAVES_PER_EU))) void
k_numeric_wave_fast(DBD_NUMERIC_ARGS) {
  numeric_hash_body<V, T, NT, G, true>(rows, pbeg, pend, Aptr, bstart, blen, Aval, Bidx, Bval, Cptr, cbase, Cidx,
                                       Cval, err);
}

template <typename V, int T, int NT, int G>
__global__ __launch_bounds__(NT) void k_numeric_block(DBD_NUMERIC_ARGS) {
  numeric_hash_body<V, T, NT, G>(rows, pbeg, pend, Aptr, bstart, blen, Aval, Bidx, Bval, Cptr, cbase, Cidx, Cval,
                                 err);
}

// Long rows, any column range: the row's columns are swept in windows [lo, hi); a window starts
// at the smallest next column over the row's A entries, so empty column ranges are skipped, and
// every window's output is produced in column order (no sort).
//
// Rows of up to kSpanMax entries (with B's band index; without it, rows whose entries each have
// a subgroup of G lanes, nA <= NT/G) take windows of up to BW columns, with two walks over the
// window's products:
//   1. presence: every B column in the window sets its bit in an LDS bitmap (span walks: every
//      thread takes U elements of the flattened spans per step; cursor walks: each A entry's
//      subgroup counts its elements in the window, U loads per lane in flight);
//   2. values: the bitmap's word prefixes give every column its rank among the window's
//      columns, and A_ik * B_kj is added at that rank in an LDS value array of VC entries.
// The values are then stored as they lie (coalesced) and the columns read off the bitmap. When
// more than VC columns are present the window is cut at the VC-th column (the values walk stops
// there, the next window starts at it) and the next window is narrowed to the width that held
// VC columns; windows widen again (doubling up to BW) when they fit. Between the two cursor walks
// an active entry keeps -1 - count in its `nc` field.
//
// Rows of more entries take dense windows of VC columns in one walk (values added at column -
// lo, presence bits beside them): banded windows (below) when their entries fit the registers,
// else cursor windows, each A entry walked by a subgroup from a cursor (LongState: position,
// remaining, next column), 4 lanes per entry (8 when an entry averages 4 or more elements
// per window over the row's column span), 4 entries per subgroup in flight, a B row with more elements
// in the window than one step continuing UL steps at a time (loads issued together); an entry
// whose next column lies beyond the window costs one state read per window. Their cursors are in
// LDS (sharing the word-prefix space) up to kOldStates entries, else in the workspace `state`
// (16 bytes per A nonzero). Both paths leave the value array and the bitmap clear.
constexpr int kNumLongStates = 512;
constexpr int kSpanMax = 256;  // span-walk windows: A entries per row at most (one per thread)


template <typename V, int NT, int BW, int VC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_numeric_long(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
    const V* __restrict__ Aval, const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
    const int64_t* __restrict__ Cptr, int64_t cbase, int32_t* __restrict__ Cidx, V* __restrict__ Cval,
    LongState* __restrict__ state, int* __restrict__ err, const int32_t* __restrict__ Aidx,
    const int32_t* __restrict__ hslot, const int32_t* __restrict__ band, int64_t nb1) {
  constexpr int NWORD = BW / 32;
  constexpr int WPT = NWORD / NT;  // bitmap words per thread
  constexpr int OWORD = VC / 32;   // bitmap words of a dense window
  constexpr int NW = NT / 64;
  constexpr int U = 4;             // loads per lane in flight per walk step
  constexpr int UL = 8;            // dense windows: steps of G elements loaded together on a long B row
  constexpr int DVB = VC * (int)sizeof(V);
  constexpr int kOldStates = (NWORD * 2 + kNumLongStates * 16) / 16;
  static_assert(NWORD % NT == 0 && NT % 64 == 0 && VC < 65535 && OWORD <= NT, "shape");
  __shared__ __align__(16) unsigned char lds[DVB + NWORD * 4 + NWORD * 2 + kNumLongStates * 16];
  V* dv = reinterpret_cast<V*>(lds);
  uint32_t* bm = reinterpret_cast<uint32_t*>(lds + DVB);
  uint16_t* pre = reinterpret_cast<uint16_t*>(lds + DVB + NWORD * 4);  // exclusive prefix per word
  LongState* stl = reinterpret_cast<LongState*>(lds + DVB + NWORD * 4 + NWORD * 2);
  LongState* ostl = reinterpret_cast<LongState*>(lds + DVB + NWORD * 4);
  __shared__ int red[NW], red_cnt[NW];
  __shared__ long long redl[NW];
  __shared__ long long s_hi1, s_hi2;
  __shared__ int s_nh, s_nl;
  static_assert(2 * NT * (8 + (int)sizeof(V) + 4) <= NWORD * 2 + kNumLongStates * 16,
                "banded windows: range tables exceed the LDS they share");
  static_assert(kSpanMax <= NT && kSpanMax % 64 == 0 && kSpanMax * (16 + (int)sizeof(V)) <= kNumLongStates * 16,
                "span walks: one entry per thread, tables within the state LDS");
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int q = t; q < NWORD; q += NT) bm[q] = 0u;
  for (int q = t; q < VC; q += NT) dv[q] = V(0);
  __syncthreads();
  auto block_min = [&](int x) {  // all threads; ends with a barrier
    const int m = wave_reduce_dpp(x, OpMin());
    if (lane == 0) red[wv] = m;
    __syncthreads();
    int r = red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = min(r, red[i]);
    __syncthreads();
    return r;
  };
  int flags = 0;
  const int64_t beg = *pbeg, end = *pend;
  for (int64_t slot = beg + blockIdx.x; slot < end; slot += gridDim.x) {
    const int row = rows[slot];
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1], nA = a1 - a0;
    const int64_t c1 = Cptr[row + 1] - cbase;
    int64_t out = Cptr[row] - cbase;
    int G = long_lanes<NT>(nA);
    int NSG = NT / G, sg = t / G, lg = t & (G - 1), sgbase = lane & ~(G - 1);
    uint64_t submask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << sgbase;
    const bool spans = band != nullptr && nA <= kSpanMax;  // uniform
    if (spans || nA <= NSG) {
      // ================= two walks per window of up to BW columns =================
      // With B's band index the walks are element-parallel over the workgroup: every active
      // entry's element span in the window (banded B rows: up to the band boundary at or past the
      // window end; other B rows: at most kSpanCap elements, the window ending before the first
      // element left out) is flattened by a scan and each thread takes U elements at a time, so
      // an entry with many elements in the window costs no chain of dependent steps. The entry
      // state (B row start and length, elements consumed, next column) is in the registers of
      // thread e (rows of up to kSpanMax entries). Without the index (rows of up to NT / G
      // entries) each entry's subgroup walks it from a cursor in LDS.
      LongState* st = stl;
      int* hS = reinterpret_cast<int*>(lds + DVB + NWORD * 4 + NWORD * 2);  // span starts
      int* np = hS + kSpanMax;                                               // consumed in walk 2
      int64_t* hb0 = reinterpret_cast<int64_t*>(np + kSpanMax);             // span start positions
      V* hva = reinterpret_cast<V*>(hb0 + kSpanMax);                         // A values
      int64_t eb0 = 0;
      int eL = 0, ep = 0, enc = INT_MAX;
      V eva = V(0);
      const int32_t* ebd = nullptr;
      int lmin = INT_MAX;
      long long psum = 0;
      if (spans) {
        if (t < nA) {
          const int64_t ka = a0 + t;
          eb0 = bstart[ka];
          eL = blen[ka];
          eva = Aval[ka];
          const int sl = hslot[Aidx[ka]];
          ebd = sl >= 0 ? band + (int64_t)sl * nb1 : nullptr;
          enc = eL > 0 ? Bidx[eb0] : INT_MAX;
          lmin = enc;
          psum = eL;
        }
      } else {
        for (int64_t e = t; e < nA; e += NT) {
          LongState x;
          x.p = bstart[a0 + e];
          x.rem = blen[a0 + e];
          x.nc = x.rem > 0 ? Bidx[x.p] : INT_MAX;
          st[e] = x;
          lmin = min(lmin, x.nc);
          psum += x.rem;
        }
      }
      // flattened element f of the window belongs to entry j, the last with hS[j] <= f (f < T)
      auto span_owner = [&](int f) {
        int j = 0;
#pragma unroll
        for (int step = kSpanMax / 2; step >= 1; step >>= 1)
          if (j + step < nA && hS[j + step] <= f) j += step;
        return j;
      };
      // a window always takes the element at its first column: at most P_i windows per row
      const long long prow = Group<NT>::reduce(psum, OpAdd(), redl);
      int lo = block_min(lmin);  // also orders the state writes before the walks
      int64_t width = BW;
      long long windows = 0;
      while (lo != INT_MAX) {
        int64_t hi1 = min((int64_t)lo + width, (int64_t)INT_MAX);  // INT_MAX: exhausted entry
        int T = 0;
        if (spans) {
          // ---- the active entries' element spans; the window ends before any element a
          // capped span leaves out ----
          int len = 0, clip = INT_MAX;
          if (t < nA && (int64_t)enc < hi1) {
            int q;
            if (ebd != nullptr) {
              q = ebd[min((hi1 + (1 << kBandShift) - 1) >> kBandShift, nb1 - 1)];
            } else {
              q = min(eL, ep + kSpanCap);
              if (q < eL) clip = Bidx[eb0 + q];
            }
            len = q - ep;
          }
          if (nA <= 64) {  // uniform: wave 0 holds every entry
            if (wv == 0) {
              const int cm = wave_reduce_dpp(clip, OpMin());
              const int incl = wave_incl_scan_dpp(len);
              if (lane < nA) hS[lane] = incl - len;
              if (lane == 63) {
                s_nh = incl;
                s_hi1 = min(hi1, (int64_t)cm);
              }
            }
          } else {
            int tot;
            const int pre0 = Group<NT>::exscan(len, tot, red_cnt);
            const int cm = Group<NT>::reduce(clip, OpMin(), red);
            if (t < nA) hS[t] = pre0;
            if (t == 0) {
              s_nh = tot;
              s_hi1 = min(hi1, (int64_t)cm);
            }
          }
          if (t < nA) {
            hb0[t] = eb0 + ep;
            hva[t] = eva;
            np[t] = 0;
          }
          __syncthreads();
          T = s_nh;
          hi1 = s_hi1;
          // ---- walk 1: presence bits of [lo, hi1), element-parallel ----
          for (int f0 = 0; f0 < T; f0 += NT * U) {  // uniform
            int c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int f = f0 + u * NT + t;
              const int j = span_owner(f);
              const int64_t at = f < T ? hb0[j] + (f - hS[j]) : 0;  // lanes past T: unused read
              c[u] = Bidx[at];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (f0 + u * NT + t < T && (int64_t)c[u] < hi1) {
                const int o = c[u] - lo;
                atomicOr(&bm[o >> 5], 1u << (o & 31));
              }
            }
          }
        } else {
        // ---- walk 1: presence bits of [lo, hi1), per-entry counts ----
        for (int64_t e = sg; e < nA; e += NSG) {  // subgroup-uniform
          LongState x = st[e];
          if ((int64_t)x.nc < hi1) {
            int64_t p = x.p;
            int rem = x.rem, cnt = 0;
            for (;;) {
              int c[U];
#pragma unroll
              for (int u = 0; u < U; ++u) {
                const int k = u * G + lg;
                const int64_t at = rem > 0 ? p + min(k, rem - 1) : 0;
                const int cv = Bidx[at];
                c[u] = k < rem ? cv : INT_MAX;
              }
              int nin = 0;
#pragma unroll
              for (int u = 0; u < U; ++u) {
                const bool in = (int64_t)c[u] < hi1;
                if (in) {
                  const int o = c[u] - lo;
                  atomicOr(&bm[o >> 5], 1u << (o & 31));
                }
                nin += __popcll(__ballot(in) & submask);
              }
              cnt += nin;
              if (nin < U * G || rem <= U * G) break;  // rem shrinks every step: bounded walk
              p += U * G;
              rem -= U * G;
            }
            if (lg == 0) {
              x.nc = -1 - cnt;
              st[e] = x;
            }
          }
        }
        }
        __syncthreads();
        // ---- word prefixes, and the cut when more than VC columns are present ----
        uint32_t wd[WPT];
        int wp[WPT];
        int cnts = 0;
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
          wd[q] = bm[t * WPT + q];
          cnts += __popc(wd[q]);
        }
        const int incl = wave_incl_scan_dpp(cnts);
        if (lane == 63) red_cnt[wv] = incl;
        __syncthreads();
        int base = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
          const int ci = red_cnt[i];
          base += i < wv ? ci : 0;
          tot += ci;
        }
        base += incl - cnts;
        {
          int run = base;
#pragma unroll
          for (int q = 0; q < WPT; ++q) {
            wp[q] = run;
            pre[t * WPT + q] = (uint16_t)min(run, 65535);
            run += __popc(wd[q]);
          }
        }
        if (tot <= VC) {
          if (t == 0) s_hi2 = hi1;
        } else {
#pragma unroll
          for (int q = 0; q < WPT; ++q) {
            const int pc = __popc(wd[q]);
            if (wp[q] <= VC && VC < wp[q] + pc) {  // the VC-th set bit (from 0) is in this word
              uint32_t w = wd[q];
              for (int j = 0; j < VC - wp[q]; ++j) w &= w - 1u;
              s_hi2 = (int64_t)lo + 32 * (t * WPT + q) + (__ffs(w) - 1);
            }
          }
        }
        __syncthreads();
        const int64_t hi2 = s_hi2;
        const int nw = min(tot, VC);
        lmin = INT_MAX;
        if (spans) {
          // ---- walk 2: values of [lo, hi2) at their ranks, element-parallel; each entry's
          // elements below hi2 are a prefix of its span, whose length the last such lane of
          // every wave records ----
          for (int f0 = 0; f0 < T; f0 += NT * U) {  // uniform
            int c[U], jj[U];
            V b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int f = f0 + u * NT + t;
              jj[u] = span_owner(f);
              const int64_t at = f < T ? hb0[jj[u]] + (f - hS[jj[u]]) : 0;
              c[u] = Bidx[at];
              b[u] = Bval[at];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int f = f0 + u * NT + t;
              const int j = jj[u];
              const bool in = f < T && (int64_t)c[u] < hi2;
              if (in) {
                const int o = c[u] - lo;
                const int w = o >> 5;
                const int r = (int)pre[w] + __popc(bm[w] & ((1u << (o & 31)) - 1u));
                lds_add(&dv[r], hva[j] * b[u]);
              }
              const int jn = __shfl_down(j, 1);
              const int inn = __shfl_down((int)in, 1);
              if (in && (lane == 63 || !inn || jn != j)) atomicMax(&np[j], f - hS[j] + 1);
            }
          }
          __syncthreads();  // np complete
          if (t < nA) {
            const int took = np[t];
            if (took > 0) {
              ep += took;
              enc = ep < eL ? Bidx[eb0 + ep] : INT_MAX;
            }
            lmin = enc;
          }
        } else {
        // ---- walk 2: values of [lo, hi2) at their ranks; cursors advance to hi2 ----
        for (int64_t e = sg; e < nA; e += NSG) {  // subgroup-uniform
          LongState x = st[e];
          if (x.nc < 0) {
            const int cnt = -1 - x.nc;
            const int n = min(cnt + 1, x.rem);  // the window's elements and the one after it
            const V va = Aval[a0 + e];
            int nin2 = 0, nxt = INT_MAX;
            for (int k0 = 0; k0 < n; k0 += U * G) {
              int c[U];
              V b[U];
#pragma unroll
              for (int u = 0; u < U; ++u) {
                const int k = k0 + u * G + lg;
                const int64_t at = x.p + min(k, n - 1);
                c[u] = Bidx[at];
                b[u] = Bval[at];
              }
#pragma unroll
              for (int u = 0; u < U; ++u) {
                const int k = k0 + u * G + lg;
                const bool valid = k < n;
                const bool in = valid && (int64_t)c[u] < hi2;
                if (in) {
                  const int o = c[u] - lo;
                  const int w = o >> 5;
                  const int r = (int)pre[w] + __popc(bm[w] & ((1u << (o & 31)) - 1u));
                  lds_add(&dv[r], va * b[u]);
                }
                if (valid && !in) nxt = min(nxt, c[u]);
                nin2 += __popcll(__ballot(in) & submask);
              }
            }
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
              if (o < G) nxt = min(nxt, __shfl_xor(nxt, o));
            x.p += nin2;
            x.rem -= nin2;
            x.nc = x.rem > 0 ? nxt : INT_MAX;
            if (lg == 0) st[e] = x;
          }
          lmin = min(lmin, x.nc);
        }
        }
        __syncthreads();
        // ---- emit: values as they lie, columns from the bitmap; both left clear ----
        for (int i = t; i < nw; i += NT) {
          if (out + i < c1) Cval[out + i] = dv[i];
          else flags |= kErrWindow;
          dv[i] = V(0);
        }
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
          const int w = t * WPT + q;
          const int64_t cb = (int64_t)lo + 32 * w;
          uint32_t bits = wd[q];
          if (cb >= hi2) bits = 0u;
          else if (cb + 32 > hi2) bits &= (1u << (int)(hi2 - cb)) - 1u;
          int64_t pos = out + wp[q];
          while (bits) {
            const int bb = __ffs(bits) - 1;
            bits &= bits - 1u;
            if (pos < c1) Cidx[pos] = (int)(cb + bb);
            else flags |= kErrWindow;
            ++pos;
          }
          bm[w] = 0u;
        }
        out += nw;
        // the width that held VC columns when cut, else twice the last width
        width = tot > VC ? max((int64_t)64, hi2 - lo) : min((int64_t)BW, 2 * width);
        lo = block_min(lmin);  // also: the clears are done before the next walk 1
        if (++windows > prow) {  // uniform: no progress (reported; the row's output is not trusted)
          flags |= kErrWindow;
          break;
        }
      }
    } else {
      bool banded_done = false;  // uniform
      if (band != nullptr) {
      // ================= banded dense windows =================
      // Windows are aligned blocks of VC columns. An entry whose B row is banded knows the exact
      // range of its elements in each window from the band index; the window's elements of all
      // banded entries are gathered element-parallel over the workgroup, one per lane. The other
      // entries (B rows below the banding threshold) keep a cursor, one lane each, four elements
      // per load. Per-entry state stays in registers (up to KE banded and KE other entries per
      // thread), and each window's element loads are issued together with the next window's band
      // end and next column, so a window costs one dependent round trip to memory. Rows with more
      // entries of either kind take the cursor windows below.
      constexpr int LOGVC = __builtin_ctz(VC);
      constexpr int BPW = VC >> kBandShift;  // bands per window
      constexpr int KE = 2;
      HRec* hrec = reinterpret_cast<HRec*>(state + a0);
      LRec* lrec = reinterpret_cast<LRec*>(state + a0);
      // chunk arrays in the word-prefix and state LDS of the other paths, which both rewrite it
      // before use (the bitmap words past the dense window's must stay clear for the two walks)
      unsigned char* reg = lds + DVB + NWORD * 4;
      int64_t* hb0 = reinterpret_cast<int64_t*>(reg);                        // KE * NT
      V* hva = reinterpret_cast<V*>(reg + KE * NT * 8);                      // KE * NT
      int* hS = reinterpret_cast<int*>(reg + KE * NT * (8 + (int)sizeof(V)));  // KE * NT
      if (t == 0) s_nh = s_nl = 0;
      __syncthreads();
      int lmin = INT_MAX, lmax = -1;
      long long psum = 0;
      for (int64_t e = t; e < nA; e += NT) {
        const int64_t ka = a0 + e;
        const int L = blen[ka];
        if (L <= 0) continue;
        const int64_t b0 = bstart[ka];
        const int sl = hslot[Aidx[ka]];
        const int cf = Bidx[b0], cl = Bidx[b0 + L - 1];
        lmin = min(lmin, cf);
        lmax = max(lmax, cl);
        psum += L;
        if (sl >= 0) {
          const int h = atomicAdd(&s_nh, 1);
          hrec[h] = HRec{b0, sl, (int)e};
        } else {
          const int l = atomicAdd(&s_nl, 1);
          lrec[nA - 1 - l] = LRec{0, (int)e, cf, L};
        }
      }
      psum = Group<NT>::reduce(psum, OpAdd(), redl);
      lmax = Group<NT>::reduce(lmax, OpMax(), red_cnt);
      lmin = Group<NT>::reduce(lmin, OpMin(), red_cnt);  // the reductions also order the record writes
      const int nH = s_nh, nL = s_nl;
      if (nH <= KE * NT && nL <= KE * NT) {  // uniform
        banded_done = true;
        int w = lmin == INT_MAX ? 1 : (lmin >> LOGVC);
        const int wl = lmin == INT_MAX ? 0 : (lmax >> LOGVC);
        // banded entries k * NT + t: B row start, A value, band row, length, range [s, s1) of the
        // current window, its end one window on (s1n) and the column after the window (nxc)
        int64_t hbr[KE];
        V hv[KE];
        const int32_t* hbi[KE];
        int hL[KE], hs[KE], hs1[KE], hs1n[KE], hnx[KE];
        // other entries: B row start, A value, position, entries left, column at the position
        int64_t lb0[KE];
        V lva[KE];
        int lp[KE], lrem[KE], lnc[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) {
          const int h = k * NT + t;
          hL[k] = 0;
          hs[k] = hs1[k] = hs1n[k] = 0;
          hnx[k] = INT_MAX;
          hbr[k] = 0;
          hv[k] = V(0);
          hbi[k] = band;
          if (h < nH) {
            const HRec r = hrec[h];
            hbr[k] = r.b0;
            hv[k] = Aval[a0 + r.e];
            hbi[k] = band + (int64_t)r.slot * nb1;
            hL[k] = hbi[k][nb1 - 1];
            hs[k] = hbi[k][w * BPW];
            hs1[k] = hbi[k][(w + 1) * BPW];
          }
          lrem[k] = 0;
          lnc[k] = INT_MAX;
          lp[k] = 0;
          lb0[k] = 0;
          lva[k] = V(0);
          if (h < nL) {
            const LRec r = lrec[nA - 1 - h];
            lb0[k] = bstart[a0 + r.e];
            lva[k] = Aval[a0 + r.e];
            lrem[k] = r.rem;
            lnc[k] = r.nc;
          }
        }
        long long windows = 0;  // every window holds an element: at most psum windows
        while (w <= wl) {
          const int lo = w << LOGVC;
          const int64_t hi = (int64_t)lo + VC;
          // ---- ranges of the banded entries: exclusive scan of their counts ----
          int csum = 0;
#pragma unroll
          for (int k = 0; k < KE; ++k) csum += hs1[k] - hs[k];
          const int incl = wave_incl_scan_dpp(csum);
          if (lane == 63) red_cnt[wv] = incl;
          __syncthreads();
          int base = 0, T = 0;
#pragma unroll
          for (int i = 0; i < NW; ++i) {
            const int ci = red_cnt[i];
            base += i < wv ? ci : 0;
            T += ci;
          }
          {
            int run = base + incl - csum;  // entries k * NT + t: this thread's entries in order
#pragma unroll
            for (int k = 0; k < KE; ++k) {
              hS[t * KE + k] = run;
              hb0[t * KE + k] = hbr[k] + hs[k];
              hva[t * KE + k] = hv[k];
              run += hs1[k] - hs[k];
            }
          }
          // ---- next window's band ends and the columns after this window (loads in flight
          // with this window's gathers) ----
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            hs1n[k] = hs1[k];
            hnx[k] = INT_MAX;
            if (hs1[k] < hL[k]) {
              hs1n[k] = hbi[k][min((w + 2) * BPW, (int)nb1 - 1)];
              hnx[k] = Bidx[hbr[k] + hs1[k]];
            }
          }
          // ---- other entries: first four elements of those with one in this window ----
          int lc[KE][4];
          V lbv[KE][4];
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            const bool act = lrem[k] > 0 && (int64_t)lnc[k] < hi;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int64_t at = act ? lb0[k] + lp[k] + min(q, lrem[k] - 1) : 0;
              lc[k][q] = Bidx[at];
              lbv[k][q] = Bval[at];
            }
          }
          __syncthreads();  // the range table is complete
          // ---- banded elements: element f belongs to the table entry j with hS[j] <= f < hS[j + 1]
          // (entries in thread-major order: thread t's KE entries are t * KE .. t * KE + KE - 1) ----
          constexpr int NE = KE * NT;
          for (int f0 = 0; f0 < T; f0 += NT * U) {  // uniform
            int col[U];
            V bv[U], va[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int f = f0 + u * NT + t;
              int j = 0;
#pragma unroll
              for (int step = NE / 2; step >= 1; step >>= 1)
                if (hS[j + step] <= f) j += step;  // hS is non-decreasing; entries past nH hold T
              const int64_t at = f < T ? hb0[j] + (f - hS[j]) : 0;  // lanes past T: unused read
              col[u] = Bidx[at];
              bv[u] = Bval[at];
              va[u] = hva[j];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (f0 + u * NT + t < T) {
                const int o = col[u] - lo;
                lds_add(&dv[o], va[u] * bv[u]);
                atomicOr(&bm[o >> 5], 1u << (o & 31));
              }
            }
          }
          // ---- other entries: accumulate the loaded elements, continue a run longer than four ----
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            if (lrem[k] > 0 && (int64_t)lnc[k] < hi) {
              int took = 0;
              bool stop = false;
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                if (!stop) {
                  if (q < lrem[k] && (int64_t)lc[k][q] < hi) {
                    const int o = lc[k][q] - lo;
                    lds_add(&dv[o], lva[k] * lbv[k][q]);
                    atomicOr(&bm[o >> 5], 1u << (o & 31));
                    ++took;
                  } else {
                    stop = true;
                  }
                }
              }
              int nc = (took < 4 && took < lrem[k]) ? lc[k][took] : INT_MAX;
              lp[k] += took;
              lrem[k] -= took;
              while (!stop && lrem[k] > 0) {  // rem shrinks every round that continues
                int cc[4];
                V bb[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  const int64_t at = lb0[k] + lp[k] + min(q, lrem[k] - 1);
                  cc[q] = Bidx[at];
                  bb[q] = Bval[at];
                }
                took = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  if (!stop) {
                    if (q < lrem[k] && (int64_t)cc[q] < hi) {
                      const int o = cc[q] - lo;
                      lds_add(&dv[o], lva[k] * bb[q]);
                      atomicOr(&bm[o >> 5], 1u << (o & 31));
                      ++took;
                    } else {
                      stop = true;
                      if (q < lrem[k]) nc = cc[q];
                    }
                  }
                }
                lp[k] += took;
                lrem[k] -= took;
              }
              lnc[k] = lrem[k] > 0 ? nc : INT_MAX;
            }
          }
          // the next window: the smallest next column over all entries
          int nxt = INT_MAX;
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            nxt = min(nxt, hnx[k]);
            if (lrem[k] > 0) nxt = min(nxt, lnc[k]);
          }
          __syncthreads();  // the window's products are in LDS
          // emit in column order: one bitmap word per thread, positions by a group scan
          uint32_t wd = t < OWORD ? bm[t] : 0u;
          const int cntw = __popc(wd);
          const int inclw = wave_incl_scan_dpp(cntw);
          const int wm = wave_reduce_dpp(nxt, OpMin());
          if (lane == 63) red_cnt[wv] = inclw;
          if (lane == 0) red[wv] = wm;
          __syncthreads();
          int basew = 0, tot = 0, nlo = INT_MAX;
#pragma unroll
          for (int i = 0; i < NW; ++i) {
            const int ci = red_cnt[i];
            basew += i < wv ? ci : 0;
            tot += ci;
            nlo = min(nlo, red[i]);
          }
          if (wd) {
            int64_t pos = out + basew + inclw - cntw;
            bm[t] = 0u;
            while (wd) {
              const int bb = __ffs(wd) - 1;
              wd &= wd - 1u;
              const int o = t * 32 + bb;
              if (pos < c1) {
                Cidx[pos] = lo + o;
                Cval[pos] = dv[o];
              } else {
                flags |= kErrWindow;
              }
              dv[o] = V(0);
              ++pos;
            }
          }
          out += tot;
          // every next column lies past this window, so the loop moves on; the banded ranges
          // advance: a skipped window holds none of their elements, so the next start is this
          // end, and the next end is the prefetched one unless windows were skipped
          const int wn = nlo == INT_MAX ? wl + 1 : (nlo >> LOGVC);
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            hs[k] = hs1[k];
            hs1[k] = (wn == w + 1 || hs1[k] >= hL[k]) ? hs1n[k] : hbi[k][min((wn + 1) * BPW, (int)nb1 - 1)];
          }
          w = wn;
          __syncthreads();  // cleared before the next window's products; reduction slots free
          if (++windows > psum) {  // uniform: no progress (reported; the row's output is not trusted)
            flags |= kErrWindow;
            break;
          }
        }
      }
      }
      if (!banded_done) {
      // ================= dense windows of VC columns, one walk =================
      LongState* st = nA <= kOldStates ? ostl : state + a0;
      int lmin = INT_MAX, lmax = -1;
      long long psum = 0;
      for (int64_t e = t; e < nA; e += NT) {
        LongState x;
        x.p = bstart[a0 + e];
        x.rem = blen[a0 + e];
        x.nc = x.rem > 0 ? Bidx[x.p] : INT_MAX;
        if (x.rem > 0) lmax = max(lmax, Bidx[x.p + x.rem - 1]);
        st[e] = x;
        lmin = min(lmin, x.nc);
        psum += x.rem;
      }
      // 8 lanes per entry when an entry averages 4 or more elements per window (its B row long
      // next to the number of VC-column windows the row's column span takes), else 4
      psum = Group<NT>::reduce(psum, OpAdd(), redl);
      const int cmax = Group<NT>::reduce(lmax, OpMax(), red_cnt);
      const int cmin = Group<NT>::reduce(lmin, OpMin(), red_cnt);
      const double span = cmax >= cmin ? (double)cmax - (double)cmin + 1.0 : 1.0;
      if ((double)psum * VC >= 4.0 * (double)nA * span) {
        G = 8;
        NSG = NT / G;
        sg = t / G;
        lg = t & (G - 1);
        sgbase = lane & ~(G - 1);
        submask = ((1ull << G) - 1ull) << sgbase;
      }
      int lo = cmin;  // the reductions above also ordered the state writes before the walk
      long long windows = 0;  // at most psum windows (each takes the element at its first column)
      while (lo != INT_MAX) {
        const int64_t hi = min((int64_t)lo + VC, (int64_t)INT_MAX);  // INT_MAX: exhausted entry
        lmin = INT_MAX;
        // 4 entries per subgroup at a time: their first loads are issued together
        for (int64_t e0 = sg; e0 < nA; e0 += (int64_t)NSG * 4) {  // subgroup-uniform
          LongState x[4];
          int c[4];
          V bv[4], va[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int64_t e = e0 + (int64_t)u * NSG;
            if (e < nA) x[u] = st[e];
            else x[u] = LongState{0, INT_MAX, 0};
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            c[u] = INT_MAX;
            bv[u] = V(0);
            va[u] = V(0);
            if (x[u].nc < hi) {  // subgroup-uniform
              const int64_t at = x[u].rem > 0 ? x[u].p + min(lg, x[u].rem - 1) : 0;
              const int cv = Bidx[at];
              bv[u] = Bval[at];
              va[u] = Aval[a0 + e0 + (int64_t)u * NSG];
              c[u] = lg < x[u].rem ? cv : INT_MAX;
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (x[u].nc < hi) {
              // the first G elements were loaded above; a B row with more in the window (a long
              // B row) continues UL steps of G at a time, their loads issued together
              bool done = false;
              {
                const int cu = c[u];
                const bool in = cu != INT_MAX && (int64_t)cu < hi;  // INT_MAX: B row exhausted
                if (in) {
                  const int o = cu - lo;
                  lds_add(&dv[o], va[u] * bv[u]);
                  atomicOr(&bm[o >> 5], 1u << (o & 31));
                }
                const int nin = __popcll(__ballot(in) & submask);
                if (nin < G) {
                  x[u].nc = __shfl(cu, sgbase + nin);
                  x[u].p += nin;
                  x[u].rem -= nin;
                  done = true;
                } else {
                  x[u].p += G;
                  x[u].rem -= G;
                }
              }
              // bounded: every pass either ends (fewer than G in the window) or takes G elements
              // off rem, and with rem <= 0 every lane holds the INT_MAX sentinel, which is
              // excluded explicitly, so the pass ends
              while (!done) {  // subgroup-uniform
                int cc[UL];
                V bb[UL];
#pragma unroll
                for (int k = 0; k < UL; ++k) {
                  const int kk = k * G + lg;
                  const int64_t at = x[u].rem > 0 ? x[u].p + min(kk, x[u].rem - 1) : 0;
                  const int cv = Bidx[at];
                  bb[k] = Bval[at];
                  cc[k] = kk < x[u].rem ? cv : INT_MAX;
                }
#pragma unroll
                for (int k = 0; k < UL; ++k) {
                  if (!done) {
                    const bool in = cc[k] != INT_MAX && (int64_t)cc[k] < hi;
                    if (in) {
                      const int o = cc[k] - lo;
                      lds_add(&dv[o], va[u] * bb[k]);
                      atomicOr(&bm[o >> 5], 1u << (o & 31));
                    }
                    const int nin = __popcll(__ballot(in) & submask);
                    if (nin < G) {
                      x[u].nc = __shfl(cc[k], sgbase + nin);
                      x[u].p += nin;
                      x[u].rem -= nin;
                      done = true;
                    } else {
                      x[u].p += G;
                      x[u].rem -= G;
                    }
                  }
                }
              }
              if (lg == 0) st[e0 + (int64_t)u * NSG] = x[u];
            }
            lmin = min(lmin, x[u].nc);
          }
        }
        __syncthreads();  // the window's products are in LDS
        // emit in column order: one bitmap word per thread, positions by a group scan
        uint32_t w = t < OWORD ? bm[t] : 0u;
        const int cnt = __popc(w);
        const int incl = wave_incl_scan_dpp(cnt);
        const int wm = wave_reduce_dpp(lmin, OpMin());
        if (lane == 63) red_cnt[wv] = incl;
        if (lane == 0) red[wv] = wm;
        __syncthreads();
        int base = 0, tot = 0, nlo = INT_MAX;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
          const int ci = red_cnt[i];
          base += i < wv ? ci : 0;
          tot += ci;
          nlo = min(nlo, red[i]);
        }
        if (w) {
          int64_t pos = out + base + incl - cnt;
          bm[t] = 0u;
          while (w) {
            const int bb = __ffs(w) - 1;
            w &= w - 1u;
            const int o = t * 32 + bb;
            if (pos < c1) {
              Cidx[pos] = lo + o;
              Cval[pos] = dv[o];
            } else {
              flags |= kErrWindow;
            }
            dv[o] = V(0);
            ++pos;
          }
        }
        out += tot;
        lo = nlo;
        __syncthreads();  // cleared before the next window's products; reduction slots free
        if (++windows > psum) {  // uniform: no progress (reported; the row's output is not trusted)
          flags |= kErrWindow;
          break;
        }
      }
      }
    }
    if (out != c1) flags |= kErrWindow;
  }
  flags = wave_reduce_dpp(flags, OpOr());
  if (lane == 0 && flags) atomicOr(err, flags);
}

// ------------------------------------------------------------------------------------
// backward: dA_ik = sum_j dC_ij B_kj ; dB_kj += A_ik dC_ij  (pattern of C from forward)
// ------------------------------------------------------------------------------------
// One step of the backward walk: U entries per subgroup (as in gather_step); every product
// looks its column up in the row's dC table, dA_ik is reduced over the subgroup and
// A_ik * dC_ij is added into dB.
template <int T, int G, int U, typename V, int S>
__device__ __forceinline__ void bwd_step(int lane, int e0, int cnt, int64_t base, const ADesc<V>& d, const int* K,
                                         const V* H, const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
                                         V* __restrict__ dA, V* __restrict__ dB, int& bad) {
  const int sg = lane / G, lg = lane % G;
  Step<U, V> st;
  gather_step<G, U, true, V, S>(st, lane, e0, cnt, d, Bidx, Bval);
  V acc[U];
  static_for<U>([&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const int e = step_entry<G, S>(e0, u, sg);
    const int er = min(e0 + u, 63);
    const int lv = ebcast<G, 4 * S + (u & 3)>(d.len, er);
    const int pl = e < cnt ? lv : 0;
    const int64_t pb = ebcast64<G, 4 * S + (u & 3)>(d.b0, er);
    const V pva = ebcastv<G, 4 * S + (u & 3)>(d.va, er);
    acc[u] = V(0);
    if (lg < pl) {
      const int s = hash_find<T>(K, st.c[u]);
      if (s < 0) {
        bad = 1;  // product column absent from C's row: C does not belong to these operands
      } else {
        const V gv = H[s];
        acc[u] = gv * st.bv[u];
        if (dB) atomicAdd(&dB[pb + lg], pva * gv);
      }
    }
    for (int q = G + lg; q < pl; q += G) {
      const int64_t p = pb + q;
      const int s = hash_find<T>(K, Bidx[p]);
      if (s < 0) {
        bad = 1;
        continue;
      }
      const V gv = H[s];
      acc[u] += gv * Bval[p];
      if (dB) atomicAdd(&dB[p], pva * gv);
    }
  });
#pragma unroll
  for (int u = 0; u < U; ++u) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc[u] += __shfl_xor(acc[u], o, G);
    const int e = step_entry<G, S>(e0, u, sg);
    if (dA && lg == 0 && e < cnt) dA[base + e] = acc[u];
  }
}

// Per C row: dC_i goes into an LDS table keyed by column; then every A entry of the row walks
// its B row (descriptor chunks, U gathers in flight per subgroup as in the forward kernels),
// looks each product column up, reduces dA_ik over its subgroup and adds A_ik * dC_ij into dB.
template <typename V, int T, int NT, int G>
__global__ __launch_bounds__(wg_of<NT>()) void k_backward_hash(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int64_t* __restrict__ bstart, const int32_t* __restrict__ blen,
    const V* __restrict__ Aval, const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
    const int64_t* __restrict__ Cptr, const int32_t* __restrict__ Cidx, const V* __restrict__ dC,
    V* __restrict__ dA, V* __restrict__ dB, int* __restrict__ err) {
  constexpr int WG = wg_of<NT>();
  constexpr int GPB = WG / NT;
  constexpr int U = 4;
  static_assert(NT % 64 == 0 && NT % G == 0 && T % NT == 0, "shape");
  __shared__ int hk[GPB][T];
  __shared__ V hv[GPB][T];
  __shared__ int sc[GPB][NT > 64 ? NT / 64 : 1];
  const int t = threadIdx.x % NT;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / NT);
  const int lane = t & 63, wv = t >> 6;
  int* K = hk[g];
  V* H = hv[g];
  const int64_t beg = *pbeg, end = *pend;
  for (int64_t slot = beg + (int64_t)blockIdx.x * GPB + g; slot < end; slot += (int64_t)gridDim.x * GPB) {
    const int row = rows[slot];
#pragma unroll
    for (int s = t; s < T; s += NT) K[s] = kEmpty;
    gsync<NT>();
    const int64_t c0 = Cptr[row], c1 = Cptr[row + 1];
    int bad = 0;
    for (int64_t m = c0 + t; m < c1; m += NT) {
      const int c = Cidx[m];
      uint32_t h = hslot<T>(c);
      // keys of a C row are unique: claim the first free slot
      for (int probe = 0; probe < T; ++probe) {
        if (atomicCAS(K + h, kEmpty, c) == kEmpty) {
          H[h] = dC[m];
          break;
        }
        h = (h + 1) & (T - 1);
      }
    }
    gsync<NT>();
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
    for (int64_t base = a0 + 64 * wv; base < a1; base += NT) {  // wave-uniform
      const int cnt = (int)min((int64_t)64, a1 - base);
      const ADesc<V> d = load_desc<G, true, V>(base, a1, lane, bstart, blen, Aval);
      if constexpr (G == 16) {  // compile-time steps (DPP broadcast lanes are immediates)
        if (0 < cnt) bwd_step<T, G, U, V, 0>(lane, 0, cnt, base, d, K, H, Bidx, Bval, dA, dB, bad);
        if (16 < cnt) bwd_step<T, G, U, V, 1>(lane, 16, cnt, base, d, K, H, Bidx, Bval, dA, dB, bad);
        if (32 < cnt) bwd_step<T, G, U, V, 2>(lane, 32, cnt, base, d, K, H, Bidx, Bval, dA, dB, bad);
        if (48 < cnt) bwd_step<T, G, U, V, 3>(lane, 48, cnt, base, d, K, H, Bidx, Bval, dA, dB, bad);
      } else {
        for (int e0 = 0; e0 < cnt; e0 += U) bwd_step<T, G, U, V, 0>(lane, e0, cnt, base, d, K, H, Bidx, Bval, dA, dB, bad);
      }
    }
    bad = Group<NT>::reduce(bad, OpMax(), sc[g]);
    if (t == 0 && bad) atomicOr(err, kErrMissingKey);
    gsync<NT>();
  }
}

// Long rows: dC's row values are scattered into a dense LDS window (the C row is sorted, so
// windows follow it); B cursors advance per window; dA accumulates across windows.
template <typename V, int NT, int W, int G>
__global__ __launch_bounds__(NT) void k_backward_long(
    const int32_t* __restrict__ rows, const int64_t* __restrict__ pbeg, const int64_t* __restrict__ pend,
    const int64_t* __restrict__ Aptr, const int32_t* __restrict__ Aidx, const V* __restrict__ Aval,
    const int64_t* __restrict__ Bptr, const int32_t* __restrict__ Bidx, const V* __restrict__ Bval,
    const int64_t* __restrict__ Cptr, const int32_t* __restrict__ Cidx, const V* __restrict__ dC,
    V* __restrict__ dA, V* __restrict__ dB, int64_t* __restrict__ cursor, int* __restrict__ err) {
  __shared__ V dv[W];
  __shared__ int sc[NT / 64];
  __shared__ long long scl[NT / 64];
  const int t = threadIdx.x;
  const int64_t beg = *pbeg, end = *pend;
  // rows from the end of the range: callers put the heaviest rows last
  for (int64_t s2 = beg + blockIdx.x; s2 < end; s2 += gridDim.x) {
    const int row = rows[end - 1 - (s2 - beg)];
    const int64_t a0 = Aptr[row], a1 = Aptr[row + 1];
    const int64_t c0 = Cptr[row], c1 = Cptr[row + 1];
    for (int64_t ka = a0 + t; ka < a1; ka += NT) {
      cursor[ka] = Bptr[Aidx[ka]];
      if (dA) dA[ka] = V(0);
    }
    int64_t cc = c0;
    __syncthreads();
    int bad = 0;
    while (cc < c1) {
      const int lo = Cidx[cc];
      const int64_t hi = min((int64_t)lo + W, (int64_t)INT_MAX);  // INT_MAX: past the B row
      // C entries of this window: [cc, ce)
      // a window holds at most W distinct columns, so entries past cc + W are outside it
      long long ce_l = min((long long)c1, (long long)(cc + W));
      for (int64_t m = cc + t; m < c1 && m < cc + W; m += NT) {
        const int c = Cidx[m];
        if ((int64_t)c < hi) {
          dv[c - lo] = dC[m];
        } else {
          ce_l = min(ce_l, (long long)m);
        }
      }
      const int64_t ce = Group<NT>::reduce(ce_l, OpMin(), scl);
      for (int64_t ka = a0 + t / G; ka < a1; ka += NT / G) {
        const int k = Aidx[ka];
        const V va = Aval[ka];
        int64_t p = cursor[ka];
        const int64_t b1 = Bptr[k + 1];
        V acc = V(0);
        // one step of G, then (a long B row) UL steps at a time with their loads issued together
        int nin = G;
        for (int first = 1; nin == G && p < b1; first = 0) {  // subgroup-uniform
          constexpr int UL = 4;
          int cc[UL];
          V bb[UL];
          const int ns = first ? 1 : UL;
#pragma unroll
          for (int k = 0; k < UL; ++k) {
            const int64_t q = p + k * G + (t % G);
            const bool ok = k < ns && q < b1;
            cc[k] = ok ? Bidx[q] : INT_MAX;
            bb[k] = ok ? Bval[q] : V(0);
          }
#pragma unroll
          for (int k = 0; k < UL; ++k) {
            if (k < ns && nin == G) {
              const int64_t q = p + (t % G);
              const int c = cc[k];
              const bool in = (int64_t)c < hi;
              if (in && c >= lo) {
                const V gv = dv[c - lo];
                acc += gv * bb[k];
                if (dB) atomicAdd(&dB[q], va * gv);
              } else if (in) {
                bad = 1;  // product column absent from C's row: C does not belong to these operands
              }
              nin = __popcll(__ballot(in) & subgroup_mask<G>());
              p += nin;
            }
          }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
        if ((t % G) == 0) {
          cursor[ka] = p;
          if (dA) dA[ka] += acc;
        }
      }
      __syncthreads();
      // clear the window entries that were set
      for (int64_t m = cc + t; m < ce; m += NT) dv[Cidx[m] - lo] = V(0);
      if (ce <= cc) bad = 1;
      cc = (ce > cc) ? ce : c1;
      __syncthreads();
    }
    // every B entry reached must have been covered by a window
    for (int64_t ka = a0 + t; ka < a1; ka += NT)
      if (cursor[ka] != Bptr[Aidx[ka] + 1]) bad = 1;
    bad = Group<NT>::reduce(bad, OpMax(), sc);
    if (t == 0 && bad) atomicOr(err, kErrMissingKey);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// dispatch tables
// ------------------------------------------------------------------------------------
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

// Workgroups of kernel f (block threads) resident on one CU, from the occupancy calculator.
static int resident_blocks(const void* f, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_pair(f, block * 64 + dev);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, 0) != hipSuccess || n < 1) n = 4;
  cache[key] = n;
  return n;
}

// Grid of a row kernel: `grid` if given (persistent launches), else one group per row. The
// one-wave-per-row kernels (gpb > 1) walk many rows per wave, their row pipeline amortising
// over them: their grid is capped at what is resident at once, so no second, partly filled
// round of workgroups follows the first (all workgroups get equal row counts).
static int grid_for(int64_t nrows, int gpb, int grid, const void* f = nullptr, int block = 256) {
  if (grid > 0) return grid;
  int64_t g = (nrows + gpb - 1) / gpb;
  if (gpb > 1 && f) g = std::min<int64_t>(g, (int64_t)device_cus() * resident_blocks(f, block));
  if (g > 1048576) g = 1048576;
  if (g < 1) g = 1;
  return (int)g;
}

constexpr int kLongNT = 256;
constexpr int kSymLongW = 131072;  // bits per symbolic window (16 KB bitmap)
#ifndef DBD_NUM_LONG_BW
#define DBD_NUM_LONG_BW 65536
#endif
constexpr int kNumLongBW = DBD_NUM_LONG_BW;  // columns per numeric window (bitmap + word prefixes)
constexpr int kNumLongVC64 = 4096;  // FP64 values per window (32 KB; with 8 KB of cursors: 3 workgroups per CU)
constexpr int kNumLongVC32 = 8192;  // FP32 values per window (32 KB)
constexpr int kBwdLongW64 = 8192;
constexpr int kBwdLongW32 = 16384;
constexpr int kLongG = 16;

#define DBD_SHAPES(X, G) \
  X(64, 64, G)           \
  X(128, 64, G)          \
  X(256, 64, G)          \
  X(512, 64, G)          \
  X(1024, 128, G)        \
  X(2048, 256, G)        \
  X(4096, 512, G)        \
  X(8192, 1024, G)

int symbolic_groups_per_wg(const KernelShape& k) {
  if (k.T <= 0) return 1;  // long and hub rows: one workgroup per row
  return k.NT <= 64 ? 256 / k.NT : 1;
}

void launch_symbolic(const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                     int64_t nrows_hint, int grid, const Mat& A, const Mat& B, int64_t* nnz, int64_t* cursor,
                     int* err, hipStream_t s, const int32_t* hub_slot, int32_t* rng_cnt, int nrng) {
  const int gpb = symbolic_groups_per_wg(k);
  if (k.T == 0 || k.T == kHubShape) {  // the symbolic pass counts hub rows with the long-row windows
    const dim3 gr(grid_for(nrows_hint, gpb, grid));
    hipLaunchKernelGGL((k_symbolic_long<1024, kSymLongW, kLongG>), gr, dim3(1024), 0, s, rows, pbeg, pend, A.ptr,
                       A.bstart, A.blen, B.idx, nnz, cursor, err, hub_slot, rng_cnt, nrng, A.idx, B.hslot, B.band,
                       B.nb1);
    return;
  }
#define X(T_, NT_, G_)                                                                                         \
  if (k.T == T_ && k.NT == NT_ && k.G == G_) {                                                               \
    const dim3 gr(grid_for(nrows_hint, gpb, grid, (const void*)k_symbolic_hash<T_, NT_, G_>, wg_of<NT_>())); \
    hipLaunchKernelGGL((k_symbolic_hash<T_, NT_, G_>), gr, dim3(wg_of<NT_>()), 0, s, rows, pbeg, pend, A.ptr, \
                       A.bstart, A.blen, B.idx, nnz, err);                                                   \
    return;                                                                                                   \
  }
  DBD_SHAPES(X, 16)
  DBD_SHAPES(X, 64)
  X(16384, 1024, 16)
  X(16384, 1024, 64)
  X(1024, 64, 16)
  X(1024, 64, 64)
#undef X
}

template <typename V>
static void launch_numeric_t(const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                             int64_t nrows_hint, int grid, const Mat& A, const Mat& B, const int64_t* Cptr,
                             int64_t cbase, int32_t* Cidx, V* Cval, int64_t* cursor, int* err, hipStream_t s,
                             bool distinct) {
  const int gpb = symbolic_groups_per_wg(k);
  const V* Av = (const V*)A.val;
  const V* Bv = (const V*)B.val;
  if (k.T == 0 || k.T == kHubShape) {  // hub rows reach here only when the range kernels cannot run
    const dim3 gr(grid_for(nrows_hint, gpb, grid));
    constexpr int VC = sizeof(V) == 8 ? kNumLongVC64 : kNumLongVC32;
    hipLaunchKernelGGL((k_numeric_long<V, kLongNT, kNumLongBW, VC>), gr, dim3(kLongNT), 0, s, rows, pbeg, pend, A.ptr,
                       A.bstart, A.blen, Av, B.idx, Bv, Cptr, cbase, Cidx, Cval, (LongState*)cursor, err, A.idx,
                       B.hslot, B.band, B.nb1);
    return;
  }
#define X(T_, NT_, G_)                                                                                           \
  if (k.T == T_ && k.NT == NT_ && k.G == G_) {                                                                 \
    if constexpr (NT_ == 64) {                                                                                 \
      if (distinct) {                                                                                          \
        const dim3 gr(grid_for(nrows_hint, gpb, grid, (const void*)k_numeric_wave_fast<V, T_, NT_, G_>, 256)); \
        hipLaunchKernelGGL((k_numeric_wave_fast<V, T_, NT_, G_>), gr, dim3(256), 0, s, rows, pbeg, pend,       \
                           A.ptr, A.bstart, A.blen, Av, B.idx, Bv, Cptr, cbase, Cidx, Cval, err);             \
      } else {                                                                                                 \
        const dim3 gr(grid_for(nrows_hint, gpb, grid, (const void*)k_numeric_wave<V, T_, NT_, G_>, 256));      \
        hipLaunchKernelGGL((k_numeric_wave<V, T_, NT_, G_>), gr, dim3(256), 0, s, rows, pbeg, pend, A.ptr,     \
                           A.bstart, A.blen, Av, B.idx, Bv, Cptr, cbase, Cidx, Cval, err);                    \
      }                                                                                                        \
    } else {                                                                                                    \
      const dim3 gr(grid_for(nrows_hint, gpb, grid));                                                          \
      hipLaunchKernelGGL((k_numeric_block<V, T_, NT_, G_>), gr, dim3(NT_), 0, s, rows, pbeg, pend, A.ptr,      \
                         A.bstart, A.blen, Av, B.idx, Bv, Cptr, cbase, Cidx, Cval, err);                      \
    }                                                                                                           \
    return;                                                                                                     \
  }
  DBD_SHAPES(X, 16)
  DBD_SHAPES(X, 64)
#undef X
}

void launch_numeric(int dtype, const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                    int64_t nrows_hint, int grid, const Mat& A, const Mat& B, const int64_t* Cptr, int64_t cbase,
                    int32_t* Cidx, void* Cval, int64_t* cursor, int* err, hipStream_t s, bool distinct) {
  if (dtype == 1)
    launch_numeric_t<double>(k, rows, pbeg, pend, nrows_hint, grid, A, B, Cptr, cbase, Cidx, (double*)Cval, cursor,
                             err, s, distinct);
  else
    launch_numeric_t<float>(k, rows, pbeg, pend, nrows_hint, grid, A, B, Cptr, cbase, Cidx, (float*)Cval, cursor,
                            err, s, distinct);
}

int hub_ranges(int64_t ncols) { return (int)((ncols + kHubRange - 1) >> kHubRangeShift); }

__global__ void k_hub_flags(const int32_t* __restrict__ rows, int64_t n, const int64_t* __restrict__ Aptr,
                            int64_t* __restrict__ flag) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int r = rows[j];
    flag[j] = Aptr[r + 1] - Aptr[r] > kHubEntries ? 1 : 0;
  }
}

__global__ void k_hub_slots(const int32_t* __restrict__ rows, int64_t n, const int64_t* __restrict__ flag,
                            const int64_t* __restrict__ pos, int32_t* __restrict__ hub_slot) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    if (flag[j]) hub_slot[rows[j]] = (int32_t)pos[j];
}

__global__ void k_hub_gather_counts(const int32_t* __restrict__ rows, int64_t n, int nrng,
                                    const int32_t* __restrict__ hub_slot, const int32_t* __restrict__ rng_cnt,
                                    int64_t* __restrict__ rcnt) {
  const int64_t items = n * nrng;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < items; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = q / nrng;
    const int r = (int)(q - j * nrng);
    // a row without a slot (only if a captured pipeline's operands changed structure) has none
    const int hs = hub_slot[rows[j]];
    rcnt[q] = hs >= 0 ? rng_cnt[(int64_t)hs * nrng + r] : 0;
  }
}

void launch_hub_flags(const int32_t* rows, int64_t n, const int64_t* Aptr, int64_t* flag, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hub_flags, dim3(fill_grid(n)), dim3(256), 0, s, rows, n, Aptr, flag);
}
void launch_hub_slots(const int32_t* rows, int64_t n, const int64_t* flag, const int64_t* pos, int32_t* hub_slot,
                      hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hub_slots, dim3(fill_grid(n)), dim3(256), 0, s, rows, n, flag, pos, hub_slot);
}
void launch_hub_gather_counts(const int32_t* rows, int64_t n, int nrng, const int32_t* hub_slot,
                              const int32_t* rng_cnt, int64_t* rcnt, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hub_gather_counts, dim3(fill_grid(n * nrng)), dim3(256), 0, s, rows, n, nrng, hub_slot, rng_cnt,
                     rcnt);
}

void launch_hub_prepare(const int32_t* rows, int64_t nh, const Mat& A, const Mat& B, int nrng, int64_t* scnt,
                        int32_t* hlist, int32_t* hcount, hipStream_t s) {
  if (nh <= 0) return;
  hipLaunchKernelGGL(k_hub_prepare<256>, dim3((unsigned)nh), dim3(256), (size_t)nrng * sizeof(int), s, rows, A.ptr,
                     A.idx, A.bstart, A.blen, B.idx, B.hslot, nrng, scnt, hlist, hcount);
}

void launch_hub_scatter(int dtype, const int32_t* rows, int64_t nh, const Mat& A, const Mat& B, int nrng,
                        const int64_t* soff, int32_t* scol, void* sval, const int* err, hipStream_t s) {
  if (nh <= 0) return;
  const size_t lds = (size_t)nrng * sizeof(unsigned long long);
  if (dtype == 1)
    hipLaunchKernelGGL((k_hub_scatter<double, 256>), dim3((unsigned)nh), dim3(256), lds, s, rows, A.ptr, A.idx,
                       A.bstart, A.blen, (const double*)A.val, B.idx, (const double*)B.val, B.hslot, nrng, soff, scol,
                       (double*)sval, err);
  else
    hipLaunchKernelGGL((k_hub_scatter<float, 256>), dim3((unsigned)nh), dim3(256), lds, s, rows, A.ptr, A.idx,
                       A.bstart, A.blen, (const float*)A.val, B.idx, (const float*)B.val, B.hslot, nrng, soff, scol,
                       (float*)sval, err);
}

// Bucket capacity check of the hub path inside a captured pipeline: more short products than the
// buckets hold flags the capacity bit, on which the scatter and range kernels write nothing.
__global__ void k_hub_guard(const int64_t* __restrict__ total, int64_t cap, int* __restrict__ err) {
  if (*total > cap) atomicOr(err, kErrCapacity);
}

void launch_hub_guard(const int64_t* total, int64_t cap, int* err, hipStream_t s) {
  hipLaunchKernelGGL(k_hub_guard, dim3(1), dim3(1), 0, s, total, cap, err);
}

// One workgroup per (row, range) item; a launch holds fewer than 2^32 work-items, so rows go in
// groups of at most (2^32 / 1024 - 1) / nrng, each launch with its rows' slices of the per-row
// and per-item arrays.
template <typename V>
static void launch_hub_range_t(const int32_t* rows, int64_t nh, int nrng, const Mat& A, const Mat& B,
                               const int32_t* hlist, const int32_t* hcount, const int64_t* soff, const int32_t* scol,
                               const V* sval, const int64_t* roff, const int64_t* Cptr, int64_t cbase, int32_t* Cidx,
                               V* Cval, int* err, hipStream_t s) {
  constexpr int64_t kMaxItems = ((int64_t)1 << 32) / 1024 - 1;
  const int64_t per = std::max<int64_t>(1, kMaxItems / nrng);
  for (int64_t j0 = 0; j0 < nh; j0 += per) {
    const int64_t nj = std::min(per, nh - j0), i0 = j0 * nrng;
    hipLaunchKernelGGL((k_hub_range<V, 1024>), dim3((unsigned)(nj * nrng)), dim3(1024), 0, s, rows + j0, nrng, A.ptr,
                       A.idx, A.bstart, (const V*)A.val, B.idx, (const V*)B.val, B.hslot, B.band, B.nb1, hlist,
                       hcount + j0, soff + i0, scol, sval, roff + i0, Cptr, cbase, Cidx, Cval, err);
  }
}

void launch_hub_range(int dtype, const int32_t* rows, int64_t nh, int nrng, const Mat& A, const Mat& B,
                      const int32_t* hlist, const int32_t* hcount, const int64_t* soff, const int32_t* scol,
                      const void* sval, const int64_t* roff, const int64_t* Cptr, int64_t cbase, int32_t* Cidx,
                      void* Cval, int* err, hipStream_t s) {
  if (nh <= 0) return;
  if (dtype == 1)
    launch_hub_range_t<double>(rows, nh, nrng, A, B, hlist, hcount, soff, scol, (const double*)sval, roff, Cptr, cbase,
                               Cidx, (double*)Cval, err, s);
  else
    launch_hub_range_t<float>(rows, nh, nrng, A, B, hlist, hcount, soff, scol, (const float*)sval, roff, Cptr, cbase,
                              Cidx, (float*)Cval, err, s);
}

template <typename V>
static void launch_backward_t(const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                              int64_t nrows_hint, int grid, const Mat& A, const Mat& B, const int64_t* Cptr,
                              const int32_t* Cidx, const V* dC, V* dA, V* dB, int64_t* cursor, int* err,
                              hipStream_t s) {
  const int gpb = symbolic_groups_per_wg(k);
  const V* Av = (const V*)A.val;
  const V* Bv = (const V*)B.val;
  if (k.T == 0 || k.T == kHubShape) {  // long and hub rows: windows over the C row
    const dim3 gr(grid_for(nrows_hint, gpb, grid));
    constexpr int W = sizeof(V) == 8 ? kBwdLongW64 : kBwdLongW32;
    hipLaunchKernelGGL((k_backward_long<V, kLongNT, W, kLongG>), gr, dim3(kLongNT), 0, s, rows, pbeg, pend, A.ptr,
                       A.idx, Av, B.ptr, B.idx, Bv, Cptr, Cidx, dC, dA, dB, cursor, err);
    return;
  }
#define X(T_, NT_, G_)                                                                                             \
  if (k.T == T_ && k.NT == NT_ && k.G == G_) {                                                                   \
    const dim3 gr(grid_for(nrows_hint, gpb, grid, (const void*)k_backward_hash<V, T_, NT_, G_>, wg_of<NT_>())); \
    hipLaunchKernelGGL((k_backward_hash<V, T_, NT_, G_>), gr, dim3(wg_of<NT_>()), 0, s, rows, pbeg, pend, A.ptr, \
                       A.bstart, A.blen, Av, B.idx, Bv, Cptr, Cidx, dC, dA, dB, err);                            \
    return;                                                                                                       \
  }
  DBD_SHAPES(X, 16)
  DBD_SHAPES(X, 64)
#undef X
}

void launch_backward(int dtype, const KernelShape& k, const int32_t* rows, const int64_t* pbeg, const int64_t* pend,
                     int64_t nrows_hint, int grid, const Mat& A, const Mat& B, const int64_t* Cptr,
                     const int32_t* Cidx, const void* dC, void* dA, void* dB, int64_t* cursor, int* err,
                     hipStream_t s) {
  if (dtype == 1)
    launch_backward_t<double>(k, rows, pbeg, pend, nrows_hint, grid, A, B, Cptr, Cidx, (const double*)dC,
                              (double*)dA, (double*)dB, cursor, err, s);
  else
    launch_backward_t<float>(k, rows, pbeg, pend, nrows_hint, grid, A, B, Cptr, Cidx, (const float*)dC,
                             (float*)dA, (float*)dB, cursor, err, s);
}

bool shape_available(int phase, const KernelShape& k) {
  if (k.T == 0 || k.T == kHubShape) return true;
  if (k.G != 16 && k.G != 64) return false;
  if (k.G > k.NT) return false;
  const int T = k.T, NT = k.NT;
  const bool base = (T == 64 && NT == 64) || (T == 128 && NT == 64) || (T == 256 && NT == 64) ||
                    (T == 512 && NT == 64) || (T == 1024 && NT == 128) || (T == 2048 && NT == 256) ||
                    (T == 4096 && NT == 512) || (T == 8192 && NT == 1024);
  if (base) return true;
  return phase == 0 && ((T == 16384 && NT == 1024) || (T == 1024 && NT == 64));
}

}  // namespace dbd
