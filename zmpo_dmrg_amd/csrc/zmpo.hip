// libzmpo: MI355X (gfx950) kernels and C-ABI for the two-site sigma-vector,
// environment renormalisation and Davidson vector work of ZMPO-DMRG.
//
// The hot op is a grouped block GEMM: one launch executes every block pair
// of one or more block-sparse contraction plans.  A "task" is one output
// block; a workgroup owns one 64x64 tile of one task and loops over all of
// that task's pairs, so each output element is written exactly once, in a
// fixed order (deterministic, no atomics).  Block-internal index maps are
// 2-level (m = mhi*mlo_dim + mlo) so permuted legs are folded into the
// LDS-staged loads instead of being materialised by transpose copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "../../include/zmpo.h"

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char* what) {
  g_last_error = what;
  return code;
}

int check_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return ZMPO_OK;
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return ZMPO_ERR_HIP;
}

static_assert(sizeof(zmpo_task) == 56, "zmpo_task layout");
static_assert(sizeof(zmpo_pair) == 64, "zmpo_pair layout");
static_assert(sizeof(zmpo_unit) == 32, "zmpo_unit layout");
static_assert(sizeof(zmpo_rgroup) == 24, "zmpo_rgroup layout");
static_assert(sizeof(zmpo_seq_op) == 120, "zmpo_seq_op layout");

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Index of the calling wave in its workgroup, known to the compiler as
// wave-uniform (scalar registers, scalar branches and descriptor loads).
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}
typedef float flt4 __attribute__((ext_vector_type(4)));

// BF16 storage (the upper 16 bits of an IEEE binary32).  Values are rounded
// to nearest, ties to even, when stored; arithmetic is FP32.
struct bf16s {
  unsigned short x;
};
typedef short sh4 __attribute__((ext_vector_type(4)));

// Loads, scaling and stores of one storage type T with arithmetic in M.
template <typename T>
struct Num {
  typedef T M;
  __device__ static inline T zero() { return (T)0; }
  __device__ static inline M get(T v) { return v; }
  __device__ static inline T put(M v) { return v; }
};
template <>
struct Num<bf16s> {
  typedef float M;
  __device__ static inline bf16s zero() { return bf16s{0}; }
  __device__ static inline float get(bf16s v) { return __uint_as_float((unsigned)v.x << 16); }
  __device__ static inline bf16s put(float f) {
    // round to nearest even on the 16 dropped bits (inputs are finite)
    const unsigned u = __float_as_uint(f);
    return bf16s{(unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16)};
  }
};
// v * scale in the arithmetic type, stored back in T (exact when scale is 1)
template <typename T>
__device__ __forceinline__ T scaled(T v, typename Num<T>::M sc) {
  return Num<T>::put(Num<T>::get(v) * sc);
}
// store an accumulated value into an output element (beta: add to it)
template <typename TC, typename V>
__device__ __forceinline__ void store_acc(TC* q, V v, int beta) {
  typedef typename Num<TC>::M M;
  *q = beta ? Num<TC>::put(Num<TC>::get(*q) + (M)v) : Num<TC>::put((M)v);
}

template <typename T>
struct Acc;
// Result element r of lane l lands at row(l/16, r), column l%16 of the 16x16
// tile: FP64 interleaves rows across lane groups, FP32 keeps four consecutive
// rows per lane group.  KG = k-steps of 4 consumed by one MFMA: operand
// element (row lr, k-step s) of lane group lk holds k = 4 s + lk, so one
// BF16 16x16x16 MFMA takes four k-steps, lane group lk supplying k = lk,
// lk + 4, lk + 8, lk + 12 of both operands (the same relabelling of k on
// both sides leaves the sum unchanged).
template <>
struct Acc<double> {
  typedef dbl4 type;
  static constexpr int KG = 1;
  __device__ static inline dbl4 mfma(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Acc<float> {
  typedef flt4 type;
  static constexpr int KG = 1;
  __device__ static inline flt4 mfma(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return 4 * lk + r; }
};
template <>
struct Acc<bf16s> {
  typedef flt4 type;
  static constexpr int KG = 4;
  __device__ static inline flt4 mfma4(bf16s a0, bf16s a1, bf16s a2, bf16s a3, bf16s b0,
                                      bf16s b1, bf16s b2, bf16s b3, flt4 c) {
    const sh4 a = {(short)a0.x, (short)a1.x, (short)a2.x, (short)a3.x};
    const sh4 b = {(short)b0.x, (short)b1.x, (short)b2.x, (short)b3.x};
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int TM = 64;
constexpr int TN = 64;
constexpr int TK = 16;
constexpr int NTHREADS = 256;

// Locate the task owning global tile t (tasks sorted by tile_base).
__device__ inline int find_task(const zmpo_task* __restrict__ tasks, int ntasks, int64_t t) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_base <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// C[task] (=|+=) sum_pairs scale * A_pair(M x K) * B_pair(K x N).
// 4 waves in a 2x2 grid, each wave owns a 32x32 quadrant = 2x2 MFMA 16x16
// accumulators.  MFMA 16x16x4 operand layout: lane l holds A[l%16][l/16] and
// B[l/16][l%16]; result element r of lane l is D[Acc<T>::row(l/16, r)][l%16].
template <typename T>
__global__ __launch_bounds__(NTHREADS) void blockgemm_kernel(
    const zmpo_task* __restrict__ tasks, int ntasks, const zmpo_pair* __restrict__ pairs,
    const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C, int beta,
    int64_t ntiles) {
  typedef typename Acc<T>::type accv;
  __shared__ T As[TK][TM + 1];
  __shared__ T Bs[TK][TN + 1];
  __shared__ int64_t a_moff[TM];
  __shared__ int64_t b_noff[TN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;
  const int lrow = lane & 15;
  const int lk = lane >> 4;

  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const zmpo_task tk = tasks[find_task(tasks, ntasks, t)];
    const int64_t local = t - tk.tile_base;
    const int m0 = (int)(local / tk.tiles_n) * TM;
    const int n0 = (int)(local % tk.tiles_n) * TN;
    accv acc[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = accv{0, 0, 0, 0};

    for (int p = tk.p_begin; p < tk.p_end; ++p) {
      const zmpo_pair pr = pairs[p];
      const T scale = (T)pr.scale;
      __syncthreads();
      if (tid < TM) {
        const int m = m0 + tid;
        a_moff[tid] = (m < tk.M)
                          ? (int64_t)(m / pr.a_mlo) * pr.a_smhi + (int64_t)(m % pr.a_mlo) * pr.a_smlo
                          : (int64_t)-1;
      } else if (tid < TM + TN) {
        const int i = tid - TM;
        const int n = n0 + i;
        b_noff[i] = (n < tk.N)
                        ? (int64_t)(n / pr.b_nlo) * pr.b_snhi + (int64_t)(n % pr.b_nlo) * pr.b_snlo
                        : (int64_t)-1;
      }
      __syncthreads();
      const T* __restrict__ Ap = A + pr.a_off;
      const T* __restrict__ Bp = B + pr.b_off;
      for (int k0 = 0; k0 < pr.K; k0 += TK) {
        if (pr.a_sk == 1) {
          // K-contiguous rows: consecutive threads walk k.
          for (int e = tid; e < TM * TK; e += NTHREADS) {
            const int kk = e % TK, mm = e / TK;
            const int k = k0 + kk;
            const int64_t mo = a_moff[mm];
            T v = 0;
            if (mo >= 0 && k < pr.K) v = Ap[mo + k];
            As[kk][mm] = v * scale;
          }
        } else {
          for (int e = tid; e < TM * TK; e += NTHREADS) {
            const int mm = e % TM, kk = e / TM;
            const int k = k0 + kk;
            const int64_t mo = a_moff[mm];
            T v = 0;
            if (mo >= 0 && k < pr.K) v = Ap[mo + (int64_t)k * pr.a_sk];
            As[kk][mm] = v * scale;
          }
        }
        if (pr.b_sk == 1) {
          for (int e = tid; e < TN * TK; e += NTHREADS) {
            const int kk = e % TK, nn = e / TK;
            const int k = k0 + kk;
            const int64_t no = b_noff[nn];
            T v = 0;
            if (no >= 0 && k < pr.K) v = Bp[no + k];
            Bs[kk][nn] = v;
          }
        } else {
          for (int e = tid; e < TN * TK; e += NTHREADS) {
            const int nn = e % TN, kk = e / TN;
            const int k = k0 + kk;
            const int64_t no = b_noff[nn];
            T v = 0;
            if (no >= 0 && k < pr.K) v = Bp[no + (int64_t)k * pr.b_sk];
            Bs[kk][nn] = v;
          }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
          const T a0 = As[kk + lk][wm * 32 + lrow];
          const T a1 = As[kk + lk][wm * 32 + 16 + lrow];
          const T b0 = Bs[kk + lk][wn * 32 + lrow];
          const T b1 = Bs[kk + lk][wn * 32 + 16 + lrow];
          acc[0][0] = Acc<T>::mfma(a0, b0, acc[0][0]);
          acc[0][1] = Acc<T>::mfma(a0, b1, acc[0][1]);
          acc[1][0] = Acc<T>::mfma(a1, b0, acc[1][0]);
          acc[1][1] = Acc<T>::mfma(a1, b1, acc[1][1]);
        }
        __syncthreads();
      }
    }
    // Epilogue: one owner per output element.
    for (int rb = 0; rb < 2; ++rb) {
      for (int r = 0; r < 4; ++r) {
        const int mi = m0 + wm * 32 + rb * 16 + Acc<T>::row(lk, r);
        if (mi >= tk.M) continue;
        const int64_t moff = tk.c_off + (int64_t)(mi / tk.c_mlo) * tk.c_smhi +
                             (int64_t)(mi % tk.c_mlo) * tk.c_smlo;
        for (int cb = 0; cb < 2; ++cb) {
          const int ni = n0 + wn * 32 + cb * 16 + lrow;
          if (ni >= tk.N) continue;
          const int64_t off = moff + (int64_t)ni * tk.c_sn;
          const T v = acc[rb][cb][r];
          C[off] = beta ? C[off] + v : v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- wave tiles
// Small-block variant: one wave owns a (16*RB) x 16 tile of one task and
// walks a contiguous slice of that task's pairs, feeding operands from L1/L2
// straight into MFMA registers (LDS only stages the output tile).  RB row blocks share
// the B fragment; four accumulator chains over consecutive k-steps hide the
// MFMA dependency latency and are added in a fixed order at the end.  A unit with
// slot >= 0 writes its tile to the workspace; wave_reduce_kernel then adds the
// slots of each tile in slot order (deterministic split-K).
// Workgroups are handed to the 8 XCDs round-robin; map hardware block b to
// logical block so that each XCD runs one contiguous range of units (tasks
// that share operand blocks are adjacent, so their re-reads hit one L2).
__device__ inline int64_t xcd_logical_block(int64_t b, int64_t nb) {
  const int64_t x = b % 8, i = b / 8;
  return x * (nb / 8) + (x < nb % 8 ? x : nb % 8) + i;
}

template <typename T, int RB>
__global__ __launch_bounds__(256) void wave_gemm_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_unit* __restrict__ units, int64_t nunits,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    T* __restrict__ C, T* __restrict__ work, int beta) {
  typedef typename Acc<T>::type accv;
  __shared__ T stage[4][RB * 256];
  const int lane = threadIdx.x & 63;
  const int wv = wave_id();
  const int64_t blk = xcd_logical_block(blockIdx.x, gridDim.x);
  const int64_t u = blk * 4 + wv;
  const bool active = u < nunits;  // wave-uniform
  const zmpo_unit un = units[active ? u : 0];
  const zmpo_task tk = tasks[un.task];
  const int lr = lane & 15;
  const int lk = lane >> 4;
  // each row's 16 consecutive k values (one 128-byte line in FP64) are loaded
  // per iteration as four k-steps; NCH accumulator chains (4 for one row
  // block, 2 otherwise -- the row blocks already give independent MFMAs) keep
  // the register budget at >= 2 waves per SIMD
  constexpr int NCH = RB == 1 ? 4 : 2;
  accv acc[NCH][RB];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[c][rb] = accv{0, 0, 0, 0};
  const int p_end = active ? un.p_end : un.p_begin;
  for (int p = un.p_begin; p < p_end; ++p) {
    const zmpo_pair pr = pairs[p];
    const T scale = (T)pr.scale;
    int64_t aoff[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int m = un.m0 + rb * 16 + lr;
      aoff[rb] = m < tk.M ? pr.a_off + (int64_t)(m / pr.a_mlo) * pr.a_smhi +
                                (int64_t)(m % pr.a_mlo) * pr.a_smlo
                          : (int64_t)-1;
    }
    const int n = un.n0 + lr;
    const int64_t boff = n < tk.N ? pr.b_off + (int64_t)(n / pr.b_nlo) * pr.b_snhi +
                                        (int64_t)(n % pr.b_nlo) * pr.b_snlo
                                  : (int64_t)-1;
    const int K = pr.K;
    for (int k0 = 0; k0 < K; k0 += 16) {
      T bv[4];
      T av[4][RB];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = k0 + 4 * c + lk;
        bv[c] = (boff >= 0 && k < K) ? B[boff + (int64_t)k * pr.b_sk] : (T)0;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          av[c][rb] = (aoff[rb] >= 0 && k < K) ? A[aoff[rb] + (int64_t)k * pr.a_sk] * scale : (T)0;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          acc[c % NCH][rb] = Acc<T>::mfma(av[c][rb], bv[c], acc[c % NCH][rb]);
    }
  }
  // Stage the tile in this wave's LDS slice, then write it out row-major so
  // that consecutive lanes store consecutive addresses of each output row.
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T v = acc[0][rb][r];
#pragma unroll
      for (int c = 1; c < NCH; ++c) v += acc[c][rb][r];
      stage[wv][(rb * 16 + Acc<T>::row(lk, r)) * 16 + lr] = v;
    }
  }
  __syncthreads();
  if (!active) return;
  if (un.slot >= 0) {
    T* __restrict__ w = work + (int64_t)un.slot * (RB * 256);
    for (int e = lane; e < RB * 256; e += 64) w[e] = stage[wv][e];
    return;
  }
  const int mv = min(16 * RB, tk.M - un.m0);
  const int nv = min(16, tk.N - un.n0);
  for (int e = lane; e < mv * nv; e += 64) {
    const int row = e / nv, col = e % nv;
    const int mi = un.m0 + row, ni = un.n0 + col;
    const int64_t off = tk.c_off + (int64_t)(mi / tk.c_mlo) * tk.c_smhi +
                        (int64_t)(mi % tk.c_mlo) * tk.c_smlo + (int64_t)ni * tk.c_sn;
    const T v = stage[wv][row * 16 + col];
    C[off] = beta ? C[off] + v : v;
  }
}

template <typename T, int RB>
__global__ __launch_bounds__(256) void wave_reduce_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_rgroup* __restrict__ groups, int64_t ngroups,
    const T* __restrict__ work, T* __restrict__ C, int beta) {
  const int lane = threadIdx.x & 63;
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ngroups) return;
  const zmpo_rgroup gr = groups[g];
  const zmpo_task tk = tasks[gr.task];
  for (int e = lane; e < RB * 256; e += 64) {
    const int row = e / 16, col = e % 16;
    const int mi = gr.m0 + row, ni = gr.n0 + col;
    if (mi >= tk.M || ni >= tk.N) continue;
    T s = 0;
    for (int q = 0; q < gr.nslots; ++q) s += work[(int64_t)(gr.slot_begin + q) * (RB * 256) + e];
    const int64_t off = tk.c_off + (int64_t)(mi / tk.c_mlo) * tk.c_smhi +
                        (int64_t)(mi % tk.c_mlo) * tk.c_smlo + (int64_t)ni * tk.c_sn;
    C[off] = beta ? C[off] + s : s;
  }
}

template <typename T, int RB>
int launch_wave(const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                const void* A, const void* B, void* C, void* work, int beta, hipStream_t s) {
  if (nunits > 0) {
    const int64_t nb = (nunits + 3) / 4;
    hipLaunchKernelGGL((wave_gemm_kernel<T, RB>), dim3((unsigned)nb), dim3(256), 0, s, tasks,
                       units, nunits, pairs, (const T*)A, (const T*)B, (T*)C, (T*)work, beta);
    int rc = check_hip(hipGetLastError(), "wave_gemm launch");
    if (rc) return rc;
  }
  if (ngroups > 0) {
    const int64_t nb = (ngroups + 3) / 4;
    hipLaunchKernelGGL((wave_reduce_kernel<T, RB>), dim3((unsigned)nb), dim3(256), 0, s, tasks,
                       groups, ngroups, (const T*)work, (T*)C, beta);
    return check_hip(hipGetLastError(), "wave_reduce launch");
  }
  return ZMPO_OK;
}

// ---------------------------------------------------------------- shaped tiles
// LDS-staged variant with block-shaped tiles: one workgroup (4 waves) owns a
// TMxTN tile of one task and a slice of its pairs (the zmpo_unit schedule of
// zmpo_wavegemm).  Operand tiles are copied to LDS with consecutive threads
// on the contiguous direction of each operand (one 128-byte line per row for
// K-contiguous operands), double-buffered over k, and the MFMA 16x16 blocks
// are spread over the waves.
// Per-thread load state for the operand tiles of one pair.  Element e of a
// KTxTM (or KTxTN) tile is owned by thread e % 256; its (row, k) split
// depends on which operand direction is contiguous, so that consecutive
// threads read consecutive addresses.
template <typename T, int ROWS, int KT>
struct TileLoader {
  typedef typename Num<T>::M M;
  static constexpr int NE = ROWS * KT / 256;
  int64_t roff[NE];  // row offset (relative to base), -1 = padding row
  int rr[NE], kk[NE];
  const T* base;
  int64_t ks;
  int K;
  M scale;
  T v[NE];       // fetched values (committed to LDS later)
  M vscale;      // scale of the pair the values belong to
  unsigned vok;  // which of v are real elements
  __device__ __forceinline__ void setup(int tid, const T* b, int64_t off, int K_, int64_t sk, int mlo,
                        int64_t shi, int64_t slo, int r0, int rmax, M sc) {
    base = b + off;
    ks = sk;
    K = K_;
    scale = sc;
    // single-level row maps (mlo >= rmax: every row in the low index) need
    // no division; the test is uniform
    const bool one = mlo >= rmax;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = tid + 256 * i;
      const bool kmaj = sk == 1;
      kk[i] = kmaj ? e % KT : e / ROWS;
      rr[i] = kmaj ? e / KT : e % ROWS;
      const int r = r0 + rr[i];
      int64_t o;
      if (one)
        o = (int64_t)r * slo;
      else
        o = (int64_t)(r / mlo) * shi + (int64_t)(r % mlo) * slo;
      roff[i] = r < rmax ? o : (int64_t)-1;
    }
  }
  // Loads of one k-tile into registers, issued unconditionally (padding
  // elements read the block's first element) so that they stay in flight
  // until commit(), which zeroes padding, scales and writes the LDS tile.
  __device__ __forceinline__ void fetch(int k0) {
    vscale = scale;
    vok = 0;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int k = k0 + kk[i];
      const bool ok = roff[i] >= 0 && k < K;
      v[i] = base[ok ? roff[i] + (int64_t)k * ks : (int64_t)0];
      vok |= (unsigned)ok << i;
    }
  }
  __device__ __forceinline__ void commit(T (*dst)[ROWS + 1]) const {
#pragma unroll
    for (int i = 0; i < NE; ++i)
      dst[kk[i]][rr[i]] = ((vok >> i) & 1u) ? scaled(v[i], vscale) : Num<T>::zero();
  }
};

// One workgroup = one unit; its pairs' k-tiles form one flat, double-buffered
// pipeline (the next tile -- possibly of the next pair -- is loaded while the
// current one feeds the MFMAs), so a unit pays the memory latency about once.
template <typename T, typename TC, int TMv, int TNv>
__global__ __launch_bounds__(256) void shaped_gemm_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_unit* __restrict__ units, int64_t nunits,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    TC* __restrict__ C, TC* __restrict__ work, int beta) {
  typedef typename Acc<T>::type accv;
  typedef typename Num<T>::M M;
  constexpr int KT = 16;
  constexpr int KG = Acc<T>::KG;
  constexpr int BM = TMv / 16, BN = TNv / 16;
  constexpr int NBLK = BM * BN;
  constexpr int PER = (NBLK + 3) / 4;
  // the output staging tile reuses the operand buffers (the k loop ends with
  // a barrier): less LDS per workgroup, more workgroups per CU
  constexpr int NA = 2 * KT * (TMv + 1), NB = 2 * KT * (TNv + 1);
  constexpr int NOPB = (NA + NB) * (int)sizeof(T), NSTB = TMv * TNv * (int)sizeof(TC);
  __shared__ __align__(16) unsigned char smem[NOPB > NSTB ? NOPB : NSTB];
  T(*As)[KT][TMv + 1] = reinterpret_cast<T(*)[KT][TMv + 1]>(smem);
  T(*Bs)[KT][TNv + 1] = reinterpret_cast<T(*)[KT][TNv + 1]>(smem + NA * sizeof(T));
  TC* stage = reinterpret_cast<TC*>(smem);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t u = xcd_logical_block(blockIdx.x, gridDim.x);
  if (u >= nunits) return;  // workgroup-uniform
  const zmpo_unit un = units[u];
  const zmpo_task tk = tasks[un.task];
  accv acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = accv{0, 0, 0, 0};
  TileLoader<T, TMv, KT> la;
  TileLoader<T, TNv, KT> lb;
  // load cursor (pair, k-tile)
  int lp = un.p_begin, lkt = 0, lK = 0;
#define ZMPO_SETUP_PAIR(P)                                                                     \
  {                                                                                              \
    const zmpo_pair pr = pairs[(P)];                                                             \
    la.setup(tid, A, pr.a_off, pr.K, pr.a_sk, pr.a_mlo, pr.a_smhi, pr.a_smlo, un.m0, tk.M,      \
             (M)pr.scale);                                                                       \
    lb.setup(tid, B, pr.b_off, pr.K, pr.b_sk, pr.b_nlo, pr.b_snhi, pr.b_snlo, un.n0, tk.N, (M)1); \
    lK = pr.K;                                                                                   \
  }
  int ks_cur = 0;  // k-steps held by the buffer being computed (0 = done)
  if (lp < un.p_end) {
    ZMPO_SETUP_PAIR(lp);
    la.fetch(0);
    lb.fetch(0);
    la.commit(As[0]);
    lb.commit(Bs[0]);
    ks_cur = min(KT, lK);
    lkt = 1;
  }
  __syncthreads();
  int buf = 0;
  while (ks_cur > 0) {
    // advance the load cursor and start the next tile's loads; they are
    // committed to the other buffer after this tile's MFMAs
    int ks_nxt = 0;
    if (lkt * KT >= lK) {
      ++lp;
      lkt = 0;
      if (lp < un.p_end) ZMPO_SETUP_PAIR(lp);
    }
    if (lp < un.p_end) {
      la.fetch(lkt * KT);
      lb.fetch(lkt * KT);
      ks_nxt = min(KT, lK - lkt * KT);
      ++lkt;
    }
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4 * KG) {
      if (kk >= ks_cur) break;  // uniform
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int blkid = wv + 4 * i;
        if (blkid < NBLK) {
          const int bm = blkid % BM, bn = blkid / BM;
          const int am = bm * 16 + lr, bnn = bn * 16 + lr;
          if constexpr (KG == 1) {
            acc[i] = Acc<T>::mfma(As[buf][kk + lk][am], Bs[buf][kk + lk][bnn], acc[i]);
          } else {
            // k rows past ks_cur hold zeros in both tiles
            acc[i] = Acc<T>::mfma4(As[buf][kk + lk][am], As[buf][kk + 4 + lk][am],
                                   As[buf][kk + 8 + lk][am], As[buf][kk + 12 + lk][am],
                                   Bs[buf][kk + lk][bnn], Bs[buf][kk + 4 + lk][bnn],
                                   Bs[buf][kk + 8 + lk][bnn], Bs[buf][kk + 12 + lk][bnn], acc[i]);
          }
        }
      }
    }
    if (ks_nxt > 0) {
      la.commit(As[buf ^ 1]);
      lb.commit(Bs[buf ^ 1]);
    }
    __syncthreads();
    buf ^= 1;
    ks_cur = ks_nxt;
  }
#undef ZMPO_SETUP_PAIR
  // stage the tile, then write it out row-major (coalesced rows)
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int blkid = wv + 4 * i;
    if (blkid < NBLK) {
      const int bm = blkid % BM, bn = blkid / BM;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        stage[(bm * 16 + Acc<T>::row(lk, r)) * TNv + bn * 16 + lr] =
            Num<TC>::put((typename Num<TC>::M)acc[i][r]);
    }
  }
  __syncthreads();
  if (un.slot >= 0) {
    TC* __restrict__ w = work + (int64_t)un.slot * (TMv * TNv);
    for (int e = tid; e < TMv * TNv; e += 256) w[e] = stage[e];
    return;
  }
  const int mv = min(TMv, tk.M - un.m0);
  const int nv = min(TNv, tk.N - un.n0);
  const bool one = tk.c_mlo >= tk.M;  // single-level row map: no division
  for (int e = tid; e < TMv * TNv; e += 256) {
    const int row = e / TNv, col = e % TNv;
    if (row >= mv || col >= nv) continue;
    const int mi = un.m0 + row, ni = un.n0 + col;
    const int64_t ro = one ? (int64_t)mi * tk.c_smlo
                           : (int64_t)(mi / tk.c_mlo) * tk.c_smhi + (int64_t)(mi % tk.c_mlo) * tk.c_smlo;
    const int64_t off = tk.c_off + ro + (int64_t)ni * tk.c_sn;
    store_acc(C + off, Num<TC>::get(stage[e]), beta);
  }
}

template <typename T, int TMv, int TNv>
__global__ __launch_bounds__(256) void shaped_reduce_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_rgroup* __restrict__ groups, int64_t ngroups,
    const T* __restrict__ work, T* __restrict__ C, int beta) {
  typedef typename Num<T>::M M;
  const int64_t g = blockIdx.x;
  if (g >= ngroups) return;
  const zmpo_rgroup gr = groups[g];
  const zmpo_task tk = tasks[gr.task];
  const int mv = min(TMv, tk.M - gr.m0);
  const int nv = min(TNv, tk.N - gr.n0);
  for (int e = threadIdx.x; e < TMv * TNv; e += blockDim.x) {
    const int row = e / TNv, col = e % TNv;
    if (row >= mv || col >= nv) continue;
    M s = 0;
    for (int q = 0; q < gr.nslots; ++q)
      s += Num<T>::get(work[(int64_t)(gr.slot_begin + q) * (TMv * TNv) + e]);
    const int mi = gr.m0 + row, ni = gr.n0 + col;
    const int64_t off = tk.c_off + (int64_t)(mi / tk.c_mlo) * tk.c_smhi +
                        (int64_t)(mi % tk.c_mlo) * tk.c_smlo + (int64_t)ni * tk.c_sn;
    store_acc(C + off, s, beta);
  }
}

template <typename T, typename TC, int TMv, int TNv>
int launch_shaped(const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                  const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                  const void* A, const void* B, void* C, void* work, int beta, hipStream_t s) {
  if (nunits > 0) {
    hipLaunchKernelGGL((shaped_gemm_kernel<T, TC, TMv, TNv>), dim3((unsigned)nunits), dim3(256), 0,
                       s, tasks, units, nunits, pairs, (const T*)A, (const T*)B, (TC*)C,
                       (TC*)work, beta);
    int rc = check_hip(hipGetLastError(), "shaped_gemm launch");
    if (rc) return rc;
  }
  if (ngroups > 0) {
    hipLaunchKernelGGL((shaped_reduce_kernel<TC, TMv, TNv>), dim3((unsigned)ngroups), dim3(256), 0,
                       s, tasks, groups, ngroups, (const TC*)work, (TC*)C, beta);
    return check_hip(hipGetLastError(), "shaped_reduce launch");
  }
  return ZMPO_OK;
}

// -------------------------------------------------------------- panel tiles
// Register-blocked grouped GEMM for large blocks (at D ~ 1000 the sigma
// stage L x psi has M ~ 2500, N ~ K ~ 130 per pair; the stage x R has
// M ~ N ~ 130 and hundreds of K ~ 2500 pairs per output block).  One
// workgroup = one zmpo_unit of the shaped kernel's schedule (a TM x TN tile
// over a range of pairs, split-K slots alike); its 4 waves form a 2 x 2 grid,
// each owning (TM/2) x (TN/2) of the tile as BMW x BNW MFMA blocks, so every
// operand fragment read from LDS feeds BNW (or BMW) MFMAs.  Operand tiles
// are kept in LDS with k contiguous ([row][k], padded pitch) and a lane
// fetches its whole fragment with one or two 16-byte reads: FP64 / FP32
// lanes hold 4 consecutive k and feed four 16x16x4 MFMAs, BF16 lanes hold 8
// and feed one 16x16x32 MFMA (the same relabelling of k on both operands
// leaves every sum unchanged).  The next k-tile is loaded from global memory
// into registers before the MFMAs of the current one and written to the
// other LDS buffer after them: one barrier per k-tile.  k rows past a pair's
// K and rows / columns past the block are zero in LDS, so the MFMAs need no
// bounds tests.
typedef __bf16 bfv8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Panel;
template <>
struct Panel<double> {
  static constexpr int KT = 16, KF = 4, PITCH = 18;
  typedef dbl4 acc;
  typedef double stg;  // register type of one staged element
  struct frag {
    double v[4];
  };
  __device__ static inline void read(frag& f, const double* p) {
    const double2 x = *reinterpret_cast<const double2*>(p);
    const double2 y = *reinterpret_cast<const double2*>(p + 2);
    f.v[0] = x.x; f.v[1] = x.y; f.v[2] = y.x; f.v[3] = y.y;
  }
  __device__ static inline dbl4 mma(const frag& a, const frag& b, dbl4 c) {
#pragma unroll
    for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[s], b.v[s], c, 0, 0, 0);
    return c;
  }
  __device__ static inline int row(int g, int r) { return g + 4 * r; }
};
template <>
struct Panel<float> {
  static constexpr int KT = 16, KF = 4, PITCH = 20;
  typedef flt4 acc;
  typedef float stg;
  struct frag {
    float v[4];
  };
  __device__ static inline void read(frag& f, const float* p) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    f.v[0] = x.x; f.v[1] = x.y; f.v[2] = x.z; f.v[3] = x.w;
  }
  __device__ static inline flt4 mma(const frag& a, const frag& b, flt4 c) {
#pragma unroll
    for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[s], b.v[s], c, 0, 0, 0);
    return c;
  }
  __device__ static inline int row(int g, int r) { return 4 * g + r; }
};
template <>
struct Panel<bf16s> {
  static constexpr int KT = 32, KF = 8, PITCH = 40;
  typedef flt4 acc;
  typedef unsigned stg;  // one BF16 per 32-bit register (low half)
  struct frag {
    bfv8 v;
  };
  __device__ static inline void read(frag& f, const bf16s* p) {
    f.v = __builtin_bit_cast(bfv8, *reinterpret_cast<const uint4*>(p));
  }
  __device__ static inline flt4 mma(const frag& a, const frag& b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
  }
  __device__ static inline int row(int g, int r) { return 4 * g + r; }
};

// Staged element conversions (BF16 travels as the low half of a register).
__device__ __forceinline__ double stg_get(double v, const double*) { return v; }
__device__ __forceinline__ float stg_get(float v, const float*) { return v; }
__device__ __forceinline__ unsigned stg_load(const bf16s* p) { return p->x; }
__device__ __forceinline__ double stg_load(const double* p) { return *p; }
__device__ __forceinline__ float stg_load(const float* p) { return *p; }
__device__ __forceinline__ void stg_put(double* q, double v) { *q = v; }
__device__ __forceinline__ void stg_put(float* q, float v) { *q = v; }
__device__ __forceinline__ void stg_put(bf16s* q, unsigned v) { q->x = (unsigned short)v; }
// v * scale; mode 0: scale 1 (unchanged), 1: scale -1 (exact sign flip),
// 2: any other scale (BF16: product rounded to nearest even)
__device__ __forceinline__ double stg_scale(double v, int mode, double sc) {
  return mode == 0 ? v : mode == 1 ? -v : v * sc;
}
__device__ __forceinline__ float stg_scale(float v, int mode, float sc) {
  return mode == 0 ? v : mode == 1 ? -v : v * sc;
}
__device__ __forceinline__ unsigned stg_scale(unsigned v, int mode, float sc) {
  if (mode == 0) return v;
  if (mode == 1) return v ^ 0x8000u;
  return Num<bf16s>::put(__uint_as_float(v << 16) * sc).x;
}

// Per-thread load state of one operand tile (ROWS rows x KT k of one
// pair).  Element i of thread t: for a k-contiguous operand (sk == 1) row
// rf + i * rs at the fixed k = kb (16 / 32 consecutive threads cover a
// row's k-tile), otherwise row rf at k = kb + i * ks (consecutive threads
// cover consecutive rows).  off[i] is the element's offset at k-tile 0.
// The second form covers the tile only when ROWS divides 256: tiles with
// other row counts (ZMPO_PANEL_32x160) take k-contiguous operands only,
// which the host checks before choosing them.
template <typename T, int ROWS>
struct PanelLoader {
  static constexpr int KT = Panel<T>::KT;
  static constexpr int NE = ROWS * KT / 256;
  static_assert(NE >= 1 && NE <= 32, "panel tile too small or too large");
  typedef typename Panel<T>::stg S;
  typedef typename Num<T>::M M;
  int64_t off[NE];
  const T* base;
  int64_t sk;
  int K, kb, ks, rf, rs, mode;
  unsigned okm;
  unsigned vok;  // which fetched values are real elements
  M scale;
  __device__ __forceinline__ void setup(int tid, const T* b, int64_t boff, int K_, int64_t sk_,
                                        int mlo, int64_t shi, int64_t slo, int r0, int rmax,
                                        double sc) {
    base = b + boff;
    sk = sk_;
    K = K_;
    scale = (M)sc;
    mode = sc == 1.0 ? 0 : sc == -1.0 ? 1 : 2;
    if (sk_ == 1) {
      kb = tid % KT; ks = 0; rf = tid / KT; rs = 256 / KT;
    } else {
      kb = tid / ROWS; ks = 256 / ROWS; rf = tid % ROWS; rs = 0;
    }
    const bool one = mlo >= rmax;  // single-level row map (uniform)
    okm = 0;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int r = r0 + rf + i * rs;
      const bool ok = r < rmax;
      int64_t o = 0;
      if (ok) o = one ? (int64_t)r * slo : (int64_t)(r / mlo) * shi + (int64_t)(r % mlo) * slo;
      off[i] = o + (int64_t)(kb + i * ks) * sk_;
      okm |= (unsigned)ok << i;
    }
  }
  // Loads are issued unconditionally (padding elements read the block's
  // first element), so a tile's loads stay in flight through the MFMAs of
  // the previous tile; store() zeroes padding and applies the pair's scale
  // (the loader is not set up again between a fetch and its store).
  __device__ __forceinline__ void fetch(S (&v)[NE], int k0) {
    const int64_t ko = (int64_t)k0 * sk;
    vok = 0;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const bool ok = ((okm >> i) & 1u) && (k0 + kb + i * ks < K);
      v[i] = stg_load(base + (ok ? ko + off[i] : (int64_t)0));
      vok |= (unsigned)ok << i;
    }
  }
  __device__ __forceinline__ void store(T* lds, const S (&v)[NE]) const {
    constexpr int P = Panel<T>::PITCH;
#pragma unroll
    for (int i = 0; i < NE; ++i)
      stg_put(lds + (rf + i * rs) * P + kb + i * ks,
              ((vok >> i) & 1u) ? stg_scale(v[i], mode, scale) : (S)0);
  }
};

extern __shared__ __align__(16) unsigned char panel_lds[];

template <typename T, typename TC, int TMv, int TNv>
__global__ __launch_bounds__(256) void panel_gemm_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_unit* __restrict__ units, int64_t nunits,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    TC* __restrict__ C, TC* __restrict__ work, int beta) {
  typedef Panel<T> PT;
  typedef typename PT::acc accv;
  constexpr int KT = PT::KT, P = PT::PITCH, KF = PT::KF;
  constexpr int BMW = TMv / 32, BNW = TNv / 32;
  static_assert(BMW >= 1 && BNW >= 1, "panel tile below 32 x 32");
  T* const As0 = reinterpret_cast<T*>(panel_lds);
  T* const Bs0 = As0 + 2 * TMv * P;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = wave_id();
  const int lr = lane & 15, g = lane >> 4;
  const int wm = wv & 1, wn = wv >> 1;
  const int64_t u = xcd_logical_block(blockIdx.x, gridDim.x);
  if (u >= nunits) return;  // workgroup-uniform
  const zmpo_unit un = units[u];
  const zmpo_task tk = tasks[un.task];
  accv acc[BMW][BNW];
#pragma unroll
  for (int i = 0; i < BMW; ++i)
#pragma unroll
    for (int j = 0; j < BNW; ++j) acc[i][j] = accv{0, 0, 0, 0};
  PanelLoader<T, TMv> la;
  PanelLoader<T, TNv> lb;
  typename PT::stg sa[PanelLoader<T, TMv>::NE], sb[PanelLoader<T, TNv>::NE];
  int p = un.p_begin, k0 = 0;
#define ZMPO_PANEL_SETUP(Q)                                                                      \
  {                                                                                              \
    const zmpo_pair pr = pairs[(Q)];                                                             \
    la.setup(tid, A, pr.a_off, pr.K, pr.a_sk, pr.a_mlo, pr.a_smhi, pr.a_smlo, un.m0, tk.M,       \
             pr.scale);                                                                          \
    lb.setup(tid, B, pr.b_off, pr.K, pr.b_sk, pr.b_nlo, pr.b_snhi, pr.b_snlo, un.n0, tk.N, 1.0); \
  }
  bool have = p < un.p_end;
  if (have) {
    ZMPO_PANEL_SETUP(p);
    la.fetch(sa, 0);
    lb.fetch(sb, 0);
    la.store(As0, sa);
    lb.store(Bs0, sb);
  }
  __syncthreads();
  int cur = 0;
  while (have) {
    // advance to the next (pair, k-tile) and start its global loads
    k0 += KT;
    if (k0 >= la.K) {
      ++p;
      k0 = 0;
      if (p < un.p_end) ZMPO_PANEL_SETUP(p);
    }
    const bool nxt = p < un.p_end;
    if (nxt) {
      la.fetch(sa, k0);
      lb.fetch(sb, k0);
    }
    const T* __restrict__ As = As0 + cur * TMv * P;
    const T* __restrict__ Bs = Bs0 + cur * TNv * P;
    typename PT::frag fa[BMW], fb[BNW];
#pragma unroll
    for (int i = 0; i < BMW; ++i) PT::read(fa[i], As + (wm * (TMv / 2) + i * 16 + lr) * P + KF * g);
#pragma unroll
    for (int j = 0; j < BNW; ++j) PT::read(fb[j], Bs + (wn * (TNv / 2) + j * 16 + lr) * P + KF * g);
#pragma unroll
    for (int i = 0; i < BMW; ++i)
#pragma unroll
      for (int j = 0; j < BNW; ++j) acc[i][j] = PT::mma(fa[i], fb[j], acc[i][j]);
    if (nxt) {
      la.store(As0 + (cur ^ 1) * TMv * P, sa);
      lb.store(Bs0 + (cur ^ 1) * TNv * P, sb);
    }
    __syncthreads();
    cur ^= 1;
    have = nxt;
  }
#undef ZMPO_PANEL_SETUP
  typedef typename Num<TC>::M MC;
  if (un.slot >= 0) {
    TC* __restrict__ w = work + (int64_t)un.slot * (TMv * TNv);
#pragma unroll
    for (int i = 0; i < BMW; ++i)
#pragma unroll
      for (int j = 0; j < BNW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm * (TMv / 2) + i * 16 + PT::row(g, r);
          const int col = wn * (TNv / 2) + j * 16 + lr;
          w[row * TNv + col] = Num<TC>::put((MC)acc[i][j][r]);
        }
    return;
  }
  const int mv = min(TMv, tk.M - un.m0);
  const int nv = min(TNv, tk.N - un.n0);
  const bool one = tk.c_mlo >= tk.M;  // single-level row map: no division
  TC* __restrict__ cb = C + tk.c_off;
#pragma unroll
  for (int i = 0; i < BMW; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = wm * (TMv / 2) + i * 16 + PT::row(g, r);
      if (row >= mv) continue;
      const int mi = un.m0 + row;
      const int64_t ro = one ? (int64_t)mi * tk.c_smlo
                             : (int64_t)(mi / tk.c_mlo) * tk.c_smhi + (int64_t)(mi % tk.c_mlo) * tk.c_smlo;
#pragma unroll
      for (int j = 0; j < BNW; ++j) {
        const int col = wn * (TNv / 2) + j * 16 + lr;
        if (col < nv) store_acc(cb + ro + (int64_t)(un.n0 + col) * tk.c_sn, acc[i][j][r], beta);
      }
    }
}

template <typename T, typename TC, int TMv, int TNv>
int launch_panel(const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                 const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                 const void* A, const void* B, void* C, void* work, int beta, hipStream_t s) {
  const size_t lds = (size_t)2 * (TMv + TNv) * Panel<T>::PITCH * sizeof(T);
  if (nunits > 0) {
    if (lds > 64 * 1024) {
      // raised once per device, before the first launch that needs it
      constexpr int MAXDEV = 64;
      static int allowed[MAXDEV] = {0};
      int dev = 0;
      int rc = check_hip(hipGetDevice(&dev), "panelgemm: device");
      if (rc) return rc;
      if (dev < 0 || dev >= MAXDEV) return set_error(ZMPO_ERR_ARG, "panelgemm: device index beyond 63");
      if (!allowed[dev]) {
        rc = check_hip(hipFuncSetAttribute((const void*)panel_gemm_kernel<T, TC, TMv, TNv>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "panelgemm: LDS attribute");
        if (rc) return rc;
        allowed[dev] = 1;
      }
    }
    hipLaunchKernelGGL((panel_gemm_kernel<T, TC, TMv, TNv>), dim3((unsigned)nunits), dim3(256), lds,
                       s, tasks, units, nunits, pairs, (const T*)A, (const T*)B, (TC*)C, (TC*)work,
                       beta);
    int rc = check_hip(hipGetLastError(), "panel_gemm launch");
    if (rc) return rc;
  }
  if (ngroups > 0) {
    hipLaunchKernelGGL((shaped_reduce_kernel<TC, TMv, TNv>), dim3((unsigned)ngroups), dim3(256), 0,
                       s, tasks, groups, ngroups, (const TC*)work, (TC*)C, beta);
    return check_hip(hipGetLastError(), "panel_reduce launch");
  }
  return ZMPO_OK;
}

template <typename T, typename TC = T>
int dispatch_shaped(int shape, const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                    const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                    const void* A, const void* B, void* C, void* work, int beta, hipStream_t s) {
  switch (shape) {
    case ZMPO_TILE_64x16: return launch_shaped<T, TC, 64, 16>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_128x16: return launch_shaped<T, TC, 128, 16>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_32x32: return launch_shaped<T, TC, 32, 32>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_64x32: return launch_shaped<T, TC, 64, 32>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_16x16: return launch_shaped<T, TC, 16, 16>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_64x64: return launch_shaped<T, TC, 64, 64>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_64x64: return launch_panel<T, TC, 64, 64>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_128x64: return launch_panel<T, TC, 128, 64>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_64x128: return launch_panel<T, TC, 64, 128>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_128x128: return launch_panel<T, TC, 128, 128>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_32x64: return launch_panel<T, TC, 32, 64>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_64x32: return launch_panel<T, TC, 64, 32>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_PANEL_32x160: return launch_panel<T, TC, 32, 160>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
    default: return set_error(ZMPO_ERR_ARG, "shapedgemm: unknown tile shape");
  }
}

// ------------------------------------------------------------ streaming tiles
// A workgroup runs a contiguous range of units as one software pipeline: the
// operand tiles of item i+1 (an item = one KT-deep slice of one pair of one
// unit) are loaded into registers while item i is multiplied from LDS and
// while a finished unit is written out, so per-unit latency (descriptor,
// operand and store round trips) overlaps across the range instead of
// serialising every small block product.
constexpr int MAXUPW = 64;

// Register-staged operand loader.  Element e (< ROWS*KT) of the KTxROWS
// tile belongs to thread e % 256; for a K-contiguous operand (sk == 1)
// consecutive threads walk k, otherwise they walk rows, so that consecutive
// threads read consecutive addresses.
template <typename T, int ROWS, int KT>
struct StreamLoader {
  typedef typename Num<T>::M M;
  static constexpr int NE = (ROWS * KT + 255) / 256;
  int roff[NE];  // row offset relative to the block, -1 = padding row
  T v[NE];
  const T* base;
  int64_t ks;
  int K;
  bool kmaj;   // element layout of the pair set up last
  bool vkmaj;  // element layout of the values held in v (set by fetch)
  M scale;
  M vscale;      // scale of the pair whose values are held in v
  unsigned vok;  // which of v are real elements (the rest read the block's first)
  __device__ __forceinline__ static int kidx(bool km, int e) { return km ? e % KT : e / ROWS; }
  __device__ __forceinline__ static int ridx(bool km, int e) { return km ? e / KT : e % ROWS; }
  __device__ __forceinline__ void setup(int tid, const T* b, int64_t off, int K_, int64_t sk,
                                        int mlo, int64_t shi, int64_t slo, int r0, int rmax,
                                        M sc) {
    base = b + off;
    ks = sk;
    K = K_;
    kmaj = sk == 1;
    scale = sc;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = tid + 256 * i;
      const int r = r0 + ridx(kmaj, e);
      roff[i] = (e < ROWS * KT && r < rmax) ? (int)((r / mlo) * shi + (r % mlo) * slo) : -1;
    }
  }
  // The cursor may set up the next pair before the fetched values are
  // stored, so store() uses the layout, scale and validity recorded here.
  // Every lane issues its loads unconditionally (padding elements read the
  // block's first element), so the loads of a tile are all in flight while
  // the MFMAs of the previous one run; zeroing and scaling wait for store().
  __device__ __forceinline__ void fetch(int tid, int k0) {
    vkmaj = kmaj;
    vscale = scale;
    vok = 0;
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = tid + 256 * i;
      const int k = k0 + kidx(kmaj, e);
      const bool ok = roff[i] >= 0 && k < K;
      v[i] = base[ok ? roff[i] + (int64_t)k * ks : (int64_t)0];
      vok |= (unsigned)ok << i;
    }
  }
  __device__ __forceinline__ void store(int tid, T (*dst)[ROWS + 1]) const {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int e = tid + 256 * i;
      if (e < ROWS * KT)
        dst[kidx(vkmaj, e)][ridx(vkmaj, e)] = ((vok >> i) & 1u) ? scaled(v[i], vscale)
                                                                : Num<T>::zero();
    }
  }
};

template <typename T, typename TC, int TMv, int TNv, int KT>
__global__ __launch_bounds__(256) void stream_gemm_kernel(
    const zmpo_sunit* __restrict__ units, int64_t nunits, int upw,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    TC* __restrict__ C, TC* __restrict__ work, int beta) {
  typedef typename Acc<T>::type accv;
  typedef typename Num<T>::M M;
  constexpr int KG = Acc<T>::KG;
  static_assert(KT % (4 * KG) == 0, "k-tile must hold whole MFMA k blocks");
  constexpr int BM = TMv / 16, BN = TNv / 16;
  constexpr int NBLK = BM * BN;
  constexpr int PER = (NBLK + 3) / 4;
  constexpr int ASZ = KT * (TMv + 1) * (int)sizeof(T), STG = TMv * TNv * (int)sizeof(TC);
  constexpr int ABUF = ASZ > STG ? ASZ : STG;
  // the A tile buffer doubles as the output stage of a finished unit
  __shared__ __align__(16) unsigned char Abuf_raw[ABUF];
  __shared__ T Bs[KT][TNv + 1];
  __shared__ zmpo_sunit su[MAXUPW];
  T(*As)[TMv + 1] = reinterpret_cast<T(*)[TMv + 1]>(Abuf_raw);
  TC* Abuf = reinterpret_cast<TC*>(Abuf_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t u0 = xcd_logical_block(blockIdx.x, gridDim.x) * (int64_t)upw;
  if (u0 >= nunits) return;  // workgroup-uniform
  const int n = (int)min((int64_t)upw, nunits - u0);
  {
    static_assert(sizeof(zmpo_sunit) == 64, "zmpo_sunit layout");
    const uint4* src = reinterpret_cast<const uint4*>(units + u0);
    uint4* dst = reinterpret_cast<uint4*>(su);
    for (int e = tid; e < 4 * n; e += 256) dst[e] = src[e];
  }
  __syncthreads();

  StreamLoader<T, TMv, KT> la;
  StreamLoader<T, TNv, KT> lb;
#pragma unroll
  for (int i = 0; i < la.NE; ++i) la.v[i] = Num<T>::zero();
#pragma unroll
  for (int i = 0; i < lb.NE; ++i) lb.v[i] = Num<T>::zero();
  la.vkmaj = lb.vkmaj = true;
  la.vok = lb.vok = 0;
  la.vscale = lb.vscale = (M)1;
  accv acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = accv{0, 0, 0, 0};

  // load cursor: unit lu of the range, pair lp (< lpe), k-tile lkt; lK = 0
  // marks a unit without pairs (one empty item that writes zeros)
  int lu = 0, lp = 0, lpe = 0, lkt = 0, lK = 0;
#define ZMPO_STREAM_PAIR()                                                                    \
  {                                                                                           \
    const zmpo_pair pr = pairs[lp];                                                           \
    la.setup(tid, A, pr.a_off, pr.K, pr.a_sk, pr.a_mlo, pr.a_smhi, pr.a_smlo, su[lu].m0,      \
             su[lu].M, (M)pr.scale);                                                          \
    lb.setup(tid, B, pr.b_off, pr.K, pr.b_sk, pr.b_nlo, pr.b_snhi, pr.b_snlo, su[lu].n0,      \
             su[lu].N, (M)1);                                                                 \
    lK = pr.K;                                                                                \
  }
#define ZMPO_STREAM_UNIT()          \
  {                                 \
    lp = su[lu].p_begin;            \
    lpe = su[lu].p_end;             \
    lkt = 0;                        \
    lK = 0;                         \
    if (lp < lpe) ZMPO_STREAM_PAIR(); \
  }
  // fetch the item under the cursor into registers, report its depth and
  // whether it closes its unit, then advance the cursor
#define ZMPO_STREAM_FETCH(KS, LAST, UNIT)                                        \
  {                                                                              \
    UNIT = lu;                                                                   \
    if (lK > 0) {                                                                \
      la.fetch(tid, lkt * KT);                                                   \
      lb.fetch(tid, lkt * KT);                                                   \
      KS = min(KT, lK - lkt * KT);                                               \
    } else {                                                                     \
      KS = 0;                                                                    \
    }                                                                            \
    const bool more_k = lK > 0 && (lkt + 1) * KT < lK;                           \
    const bool more_p = lK > 0 && lp + 1 < lpe;                                  \
    LAST = !more_k && !more_p;                                                   \
    if (more_k) {                                                                \
      ++lkt;                                                                     \
    } else if (more_p) {                                                         \
      ++lp;                                                                      \
      lkt = 0;                                                                   \
      ZMPO_STREAM_PAIR();                                                        \
    } else {                                                                     \
      ++lu;                                                                      \
      if (lu < n) ZMPO_STREAM_UNIT();                                            \
    }                                                                            \
  }

  ZMPO_STREAM_UNIT();
  int cks, cu;
  bool clast;
  ZMPO_STREAM_FETCH(cks, clast, cu);
  la.store(tid, As);
  lb.store(tid, Bs);
  __syncthreads();
  while (true) {
    const bool nvalid = lu < n;  // uniform
    int nks = 0, nu = 0;
    bool nlast = false;
    if (nvalid) ZMPO_STREAM_FETCH(nks, nlast, nu);
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4 * KG) {
      if (kk >= cks) break;  // uniform
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int blkid = wv + 4 * i;
        if (blkid < NBLK) {
          const int bm = blkid % BM, bn = blkid / BM;
          const int am = bm * 16 + lr, bnn = bn * 16 + lr;
          if constexpr (KG == 1) {
            acc[i] = Acc<T>::mfma(As[kk + lk][am], Bs[kk + lk][bnn], acc[i]);
          } else {
            // k rows past cks hold zeros in both tiles
            acc[i] = Acc<T>::mfma4(As[kk + lk][am], As[kk + 4 + lk][am], As[kk + 8 + lk][am],
                                   As[kk + 12 + lk][am], Bs[kk + lk][bnn], Bs[kk + 4 + lk][bnn],
                                   Bs[kk + 8 + lk][bnn], Bs[kk + 12 + lk][bnn], acc[i]);
          }
        }
      }
    }
    if (clast) {
      __syncthreads();  // every wave is done reading the A tile
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int blkid = wv + 4 * i;
        if (blkid < NBLK) {
          const int bm = blkid % BM, bn = blkid / BM;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            Abuf[(bm * 16 + Acc<T>::row(lk, r)) * TNv + bn * 16 + lr] =
                Num<TC>::put((typename Num<TC>::M)acc[i][r]);
          acc[i] = accv{0, 0, 0, 0};
        }
      }
      __syncthreads();
      const zmpo_sunit& un = su[cu];
      if (un.slot >= 0) {
        TC* __restrict__ w = work + (int64_t)un.slot * (TMv * TNv);
        for (int e = tid; e < TMv * TNv; e += 256) w[e] = Abuf[e];
      } else {
        const int mv = min(TMv, un.M - un.m0);
        const int nv = min(TNv, un.N - un.n0);
        for (int e = tid; e < TMv * TNv; e += 256) {
          const int row = e / TNv, col = e % TNv;
          if (row < mv && col < nv) {
            const int mi = un.m0 + row, ni = un.n0 + col;
            const int64_t off = un.c_off + (int64_t)(mi / un.c_mlo) * un.c_smhi +
                                (int64_t)(mi % un.c_mlo) * un.c_smlo + (int64_t)ni * un.c_sn;
            store_acc(C + off, Num<TC>::get(Abuf[e]), beta);
          }
        }
      }
    }
    if (!nvalid) break;
    __syncthreads();  // LDS tiles (and the stage) are free
    la.store(tid, As);
    lb.store(tid, Bs);
    __syncthreads();
    cks = nks;
    clast = nlast;
    cu = nu;
  }
#undef ZMPO_STREAM_FETCH
#undef ZMPO_STREAM_UNIT
#undef ZMPO_STREAM_PAIR
}

template <typename T, typename TC, int TMv, int TNv, int KT>
int launch_stream(const zmpo_sunit* units, int64_t nunits, int upw, const zmpo_task* tasks,
                  const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                  const void* A, const void* B, void* C, void* work, int beta, hipStream_t s) {
  if (nunits > 0) {
    const int64_t nwg = (nunits + upw - 1) / upw;
    hipLaunchKernelGGL((stream_gemm_kernel<T, TC, TMv, TNv, KT>), dim3((unsigned)nwg), dim3(256),
                       0, s, units, nunits, upw, pairs, (const T*)A, (const T*)B, (TC*)C,
                       (TC*)work, beta);
    int rc = check_hip(hipGetLastError(), "stream_gemm launch");
    if (rc) return rc;
  }
  if (ngroups > 0) {
    hipLaunchKernelGGL((shaped_reduce_kernel<TC, TMv, TNv>), dim3((unsigned)ngroups), dim3(256), 0,
                       s, tasks, groups, ngroups, (const TC*)work, (TC*)C, beta);
    return check_hip(hipGetLastError(), "stream_reduce launch");
  }
  return ZMPO_OK;
}

template <typename T, typename TC, int KT>
int dispatch_stream_kt(int shape, const zmpo_sunit* units, int64_t nunits, int upw,
                       const zmpo_task* tasks, const zmpo_rgroup* groups, int64_t ngroups,
                       const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                       int beta, hipStream_t s) {
  switch (shape) {
    case ZMPO_TILE_64x16: return launch_stream<T, TC, 64, 16, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_128x16: return launch_stream<T, TC, 128, 16, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_32x32: return launch_stream<T, TC, 32, 32, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_64x32: return launch_stream<T, TC, 64, 32, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_16x16: return launch_stream<T, TC, 16, 16, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    case ZMPO_TILE_64x64: return launch_stream<T, TC, 64, 64, KT>(units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
    default: return set_error(ZMPO_ERR_ARG, "streamgemm: unknown tile shape");
  }
}

template <typename T, typename TC = T>
int dispatch_stream(int shape, int kt, const zmpo_sunit* units, int64_t nunits, int upw,
                    const zmpo_task* tasks, const zmpo_rgroup* groups, int64_t ngroups,
                    const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                    int beta, hipStream_t s) {
  if (kt == 16)
    return dispatch_stream_kt<T, TC, 16>(shape, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (kt == 32)
    return dispatch_stream_kt<T, TC, 32>(shape, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  return set_error(ZMPO_ERR_ARG, "streamgemm: kt must be 16 or 32");
}

// ------------------------------------------------------------- skinny GEMMs
// Tall blocks with few columns (sigma stage 1: M = bra x MPO-bond rows in
// the hundreds, N and K = sector dims of 5-60).  One workgroup = a strip of
// 64*RB rows x up to 64 columns of one output block; each wave owns 16*RB
// rows and works alone (no LDS, no barriers).  Operands go straight from
// global memory into the MFMA registers -- A rows K-contiguous, B rows
// N-contiguous (16 lanes = one 128-byte segment) -- all k steps of a 4*SK
// slice issued before their MFMAs; rows and columns are clamped into the
// block (those results are not stored) and A is zeroed past K, so the
// products need no other bounds tests.  Tasks may have several pairs; each
// output element is written once.
template <typename T, int RB>
__global__ __launch_bounds__(256) void skinny_gemm_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_unit* __restrict__ units, int64_t nunits,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    T* __restrict__ C, int beta) {
  typedef typename Acc<T>::type accv;
  constexpr int SK = 4;  // k steps per slice
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t u = xcd_logical_block(blockIdx.x, gridDim.x);
  if (u >= nunits) return;  // workgroup-uniform
  const zmpo_unit un = units[u];
  const zmpo_task& tk = tasks[un.task];
  const int M = tk.M, N = tk.N;
  const int rbase = un.m0 + wv * 16 * RB;
  if (rbase >= M) return;  // wave-uniform: this wave's rows are all past the block
  const int nv = min(64, N - un.n0);
  const int nt = (nv + 15) >> 4;  // 16-column MFMA tiles in use (uniform)
  accv acc[RB][4];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = accv{0, 0, 0, 0};
  for (int p = un.p_begin; p < un.p_end; ++p) {
    const zmpo_pair& pr = pairs[p];
    const int K = pr.K;
    const T sc = (T)pr.scale;
    const T* __restrict__ a = A + pr.a_off;
    const T* __restrict__ b = B + pr.b_off;
    int64_t ro[RB], co[4];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int m = min(rbase + i * 16 + lr, M - 1);
      ro[i] = pr.a_mlo >= M ? (int64_t)m * pr.a_smlo
                            : (int64_t)(m / pr.a_mlo) * pr.a_smhi + (int64_t)(m % pr.a_mlo) * pr.a_smlo;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = min(un.n0 + j * 16 + lr, N - 1);
      co[j] = pr.b_nlo >= N ? (int64_t)n * pr.b_snlo
                            : (int64_t)(n / pr.b_nlo) * pr.b_snhi + (int64_t)(n % pr.b_nlo) * pr.b_snlo;
    }
    for (int k0 = 0; k0 < K; k0 += 4 * SK) {
      T av[SK][RB], bv[SK][4];
#pragma unroll
      for (int s = 0; s < SK; ++s) {
        const int k = k0 + 4 * s + lk;
        const bool kok = k < K;
        const int64_t kc = min(k, K - 1);
#pragma unroll
        for (int i = 0; i < RB; ++i) {
          const T x = a[ro[i] + kc * pr.a_sk];
          av[s][i] = kok ? x * sc : (T)0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j < nt) bv[s][j] = b[co[j] + kc * pr.b_sk];
      }
#pragma unroll
      for (int s = 0; s < SK; ++s) {
        if (k0 + 4 * s >= K) break;  // uniform
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j >= nt) break;  // uniform
#pragma unroll
          for (int i = 0; i < RB; ++i) acc[i][j] = Acc<T>::mfma(av[s][i], bv[s][j], acc[i][j]);
        }
      }
    }
  }
  // output: lane column n0 + j*16 + lr, rows rbase + i*16 + row(lk, r)
  const bool cone = tk.c_mlo >= M;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = rbase + i * 16 + Acc<T>::row(lk, r);
      if (m >= M) continue;
      const int64_t ro = tk.c_off + (cone ? (int64_t)m * tk.c_smlo
                                          : (int64_t)(m / tk.c_mlo) * tk.c_smhi +
                                                (int64_t)(m % tk.c_mlo) * tk.c_smlo);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = j * 16 + lr;
        if (n < nv) {
          T* q = C + ro + (int64_t)(un.n0 + n) * tk.c_sn;
          const T v = acc[i][j][r];
          *q = beta ? *q + v : v;
        }
      }
    }
}

template <typename T, int RB>
int launch_skinny(const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                  const zmpo_pair* pairs, const void* A, const void* B, void* C, int beta,
                  hipStream_t s) {
  hipLaunchKernelGGL((skinny_gemm_kernel<T, RB>), dim3((unsigned)nunits), dim3(256), 0, s, tasks,
                     units, nunits, pairs, (const T*)A, (const T*)B, (T*)C, beta);
  return check_hip(hipGetLastError(), "skinny_gemm launch");
}

// ---------------------------------------------------------------- dot GEMMs
// Both operands K-contiguous (A rows and B columns are runs of K elements),
// small M and N, very long summed K (sigma stage top3 x R: sigma block rows
// = bra states, columns = right states, K = all ops x (MPO bond x right
// bond)).  One workgroup = one 32x32 output tile and one chunk of its pairs
// (the zmpo_unit schedule of zmpo_shapedgemm with tile 32x32 and split-K
// slots).  The four waves split every pair's k range (wave w takes the
// 4-deep k steps w, w+4, ...), each holding a 2x2 grid of 16x16 MFMA tiles;
// operands go straight from global memory into MFMA registers (rows and
// columns clamped into the block, A zeroed past K), and the four partial
// tiles are summed through LDS in wave order at the end.
template <typename T>
__global__ __launch_bounds__(256) void dot_gemm_kernel(
    const zmpo_task* __restrict__ tasks, const zmpo_unit* __restrict__ units, int64_t nunits,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B,
    T* __restrict__ C, T* __restrict__ work, int beta) {
  typedef typename Acc<T>::type accv;
  constexpr int TMv = 32, TNv = 32, DS = 4;  // DS: k steps loaded before their MFMAs
  __shared__ T red[3][TMv][TNv + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t u = xcd_logical_block(blockIdx.x, gridDim.x);
  if (u >= nunits) return;  // workgroup-uniform
  const zmpo_unit un = units[u];
  const zmpo_task& tk = tasks[un.task];
  const int M = tk.M, N = tk.N;
  accv acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = accv{0, 0, 0, 0};
  for (int p = un.p_begin; p < un.p_end; ++p) {
    const zmpo_pair& pr = pairs[p];
    const int K = pr.K;
    const T sc = (T)pr.scale;
    // lane rows / columns (clamped into the block; results past M, N unused)
    int64_t ar[2], bc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int m = min(un.m0 + i * 16 + lr, M - 1);
      ar[i] = pr.a_mlo >= M ? (int64_t)m * pr.a_smlo
                            : (int64_t)(m / pr.a_mlo) * pr.a_smhi + (int64_t)(m % pr.a_mlo) * pr.a_smlo;
      const int n = min(un.n0 + i * 16 + lr, N - 1);
      bc[i] = pr.b_nlo >= N ? (int64_t)n * pr.b_snlo
                            : (int64_t)(n / pr.b_nlo) * pr.b_snhi + (int64_t)(n % pr.b_nlo) * pr.b_snlo;
    }
    const T* __restrict__ a = A + pr.a_off;
    const T* __restrict__ b = B + pr.b_off;
    // this wave's k steps: 4*(wv + 4*j) .. + 3
    for (int k0 = 4 * wv; k0 < K; k0 += 16 * DS) {
      T av[DS][2], bv[DS][2];
#pragma unroll
      for (int s = 0; s < DS; ++s) {
        const int k = k0 + 16 * s + lk;
        const bool kok = k < K;
        const int kc = min(k, K - 1);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const T x = a[ar[i] + (int64_t)kc * pr.a_sk];
          av[s][i] = kok ? x * sc : (T)0;
          bv[s][i] = b[bc[i] + (int64_t)kc * pr.b_sk];
        }
      }
#pragma unroll
      for (int s = 0; s < DS; ++s) {
        if (k0 + 16 * s >= K) break;  // uniform
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = Acc<T>::mfma(av[s][i], bv[s][j], acc[i][j]);
      }
    }
  }
  // sum the four waves' partial tiles in wave order: waves 1-3 park theirs
  // in LDS, wave 0 adds them to its own and writes the tile
  if (wv > 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          red[wv - 1][i * 16 + Acc<T>::row(lk, r)][j * 16 + lr] = acc[i][j][r];
  }
  __syncthreads();
  if (wv > 0) return;
  const int mv = min(TMv, M - un.m0), nv = min(TNv, N - un.n0);
  const bool cone = tk.c_mlo >= M;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i * 16 + Acc<T>::row(lk, r);
      const int m = un.m0 + row;
      const int64_t ro = tk.c_off + (cone ? (int64_t)m * tk.c_smlo
                                          : (int64_t)(m / tk.c_mlo) * tk.c_smhi +
                                                (int64_t)(m % tk.c_mlo) * tk.c_smlo);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = j * 16 + lr;
        const T v = ((acc[i][j][r] + red[0][row][col]) + red[1][row][col]) + red[2][row][col];
        if (un.slot >= 0) {
          work[(int64_t)un.slot * (TMv * TNv) + row * TNv + col] = v;
        } else if (row < mv && col < nv) {
          T* q = C + ro + (int64_t)(un.n0 + col) * tk.c_sn;
          *q = beta ? *q + v : v;
        }
      }
    }
}

template <typename T>
int launch_dot(const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
               const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs, const void* A,
               const void* B, void* C, void* work, int beta, hipStream_t s) {
  if (nunits > 0) {
    hipLaunchKernelGGL((dot_gemm_kernel<T>), dim3((unsigned)nunits), dim3(256), 0, s, tasks, units,
                       nunits, pairs, (const T*)A, (const T*)B, (T*)C, (T*)work, beta);
    int rc = check_hip(hipGetLastError(), "dot_gemm launch");
    if (rc) return rc;
  }
  if (ngroups > 0) {
    hipLaunchKernelGGL((shaped_reduce_kernel<T, 32, 32>), dim3((unsigned)ngroups), dim3(256), 0, s,
                       tasks, groups, ngroups, (const T*)work, (T*)C, beta);
    return check_hip(hipGetLastError(), "dot_reduce launch");
  }
  return ZMPO_OK;
}

// ------------------------------------------------------------- chained GEMMs
// One workgroup = TM (= 16*RB) rows of one row group.
//  stage 0: the group's distinct first-stage A blocks (TM rows each) are
//           copied into the LDS tile atile[col][row] -- one wave per block,
//           8 independent loads per lane in flight, 16 consecutive rows per
//           128-byte segment; padding columns (up to a multiple of 4) get 0;
//  stage 1: waves take first-stage tasks round-robin; each 16x32 MFMA tile
//           accumulates over the task's pairs with A from atile and B (MPO
//           blocks, zero-padded to [ceil4(K)][ceil32(N)], L2-resident) from
//           global memory; the result goes to mid[col][row];
//  stage 2: the same with A from mid, results written to C.
// Every lane's rows are fixed for the whole unit, so the row maps (one
// division by the group's row split) are computed once per unit; padded
// operands make the inner loop free of bounds tests.  Rows past the group's
// M are zeros in atile and are not written to C.
extern __shared__ __align__(16) unsigned char chain_lds[];
// k-steps (of 4) whose operands are loaded before their MFMAs.  FP32
// operands take half the registers of FP64 ones, so their slice can be
// deeper for the same register budget (more loads in flight per wave).
#ifndef CHAIN_KS
#define CHAIN_KS 5
#endif
#ifndef CHAIN_KS_F32
#define CHAIN_KS_F32 CHAIN_KS
#endif
// BF16: whole MFMA k blocks (4 k-steps each) per slice
#ifndef CHAIN_KS_BF16
#define CHAIN_KS_BF16 8
#endif
template <typename T>
struct ChainDepth {
  static constexpr int value = CHAIN_KS;
};
template <>
struct ChainDepth<float> {
  static constexpr int value = CHAIN_KS_F32;
};
template <>
struct ChainDepth<bf16s> {
  static constexpr int value = CHAIN_KS_BF16;
  static_assert(CHAIN_KS_BF16 % 4 == 0, "BF16 chained slices hold whole MFMA k blocks");
};

template <typename T, int RB>
struct ChainSlice {
  static constexpr int KS = ChainDepth<T>::value;
  T a[KS][RB];
  T b[KS][2];
};
// BF16 operands held one per 32-bit register (arrays of 16-bit elements
// were kept in scratch memory by the compiler)
template <int RB>
struct ChainSlice<bf16s, RB> {
  static constexpr int KS = ChainDepth<bf16s>::value;
  struct R {
    unsigned v;
    __device__ R& operator=(bf16s x) {
      v = x.x;
      return *this;
    }
    __device__ operator bf16s() const { return bf16s{(unsigned short)v}; }
  };
  R a[KS][RB];
  R b[KS][2];
};

// The fields of a pair that the chained stages use (uniform per wave).
// blk: the first-stage pair's A block (zmpo_cblock index), read when the A
// operand comes straight from global memory instead of the LDS tile.
struct ChainPair {
  int64_t b_off;
  int K, c, ldb, blk;
  __device__ __forceinline__ void get(const zmpo_pair* __restrict__ pairs, int p) {
    const zmpo_pair& q = pairs[p];
    b_off = q.b_off;
    K = q.K;
    c = (int)q.a_off;
    ldb = q.b_sk;
    blk = q.pad;
  }
};

// A rows of this lane for a global-memory A block (rows split by the row
// group's rsplit; rows past M read row 0 and are never stored).
template <int RB>
struct ChainRows {
  int mb[RB], mr[RB];
};

// Operands of the k-steps [k0, k0 + 4*KS) of one pair, 32-column tile
// n0.  Every slice issues the same loads: k-steps at or past K re-read the
// last k-step that starts below K (their MFMAs are skipped, or their
// operands zeroed at use for BF16), so the number of loads in flight is
// fixed and the MFMAs of the previous slice wait only for its own operands.
// Lanes past K inside the last k-step read the zero padding of both
// operands.  Bases are wave-uniform (scalar registers); each lane adds a
// fixed 32-bit offset.
template <typename T, int RB, bool GA, bool N16>
__device__ __forceinline__ void chain_load(ChainSlice<T, RB>& sl, const ChainPair& pr, int n0,
                                           int k0, const T* __restrict__ src,
                                           const T* __restrict__ Bb, int lane_a, int lr, int lk,
                                           const T* __restrict__ Ag,
                                           const zmpo_cblock* __restrict__ blocks,
                                           const ChainRows<RB>& rows) {
  constexpr int LD = 16 * RB + 1;
  const int lane_b = lk * pr.ldb + lr;
  const T* __restrict__ b0 = Bb + pr.b_off + (int64_t)k0 * pr.ldb + n0;
  const int smax = (pr.K - 1 - k0) >> 2;  // last k-step of the slice below K (k0 < K)
  if (GA) {
    // A block in global memory: lane rows fixed, k clamped into the block
    // (columns past K meet zero B rows)
    const zmpo_cblock& bk = blocks[pr.blk];
    const T* __restrict__ ab = Ag + bk.a_off;
    int64_t ro[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) ro[i] = (int64_t)rows.mb[i] * bk.a_smhi + (int64_t)rows.mr[i] * bk.a_smlo;
#pragma unroll
    for (int s = 0; s < ChainSlice<T, RB>::KS; ++s) {
      const int sc = min(s, smax);
      const T* __restrict__ bs = b0 + 4 * sc * pr.ldb;
      sl.b[s][0] = bs[lane_b];
      if (!N16) sl.b[s][1] = bs[lane_b + 16];
      const int64_t ko = (int64_t)min(k0 + 4 * sc + lk, pr.K - 1) * bk.a_sk;
#pragma unroll
      for (int i = 0; i < RB; ++i) sl.a[s][i] = ab[ro[i] + ko];
    }
    return;
  }
  const T* __restrict__ a0 = src + (pr.c + k0) * LD;
#pragma unroll
  for (int s = 0; s < ChainSlice<T, RB>::KS; ++s) {
    const int sc = min(s, smax);
    const T* __restrict__ bs = b0 + 4 * sc * pr.ldb;
    sl.b[s][0] = bs[lane_b];
    if (!N16) sl.b[s][1] = bs[lane_b + 16];
    const T* __restrict__ as = a0 + 4 * sc * LD;
#pragma unroll
    for (int i = 0; i < RB; ++i) sl.a[s][i] = as[lane_a + i * 16];
  }
}

// One stage: tasks [t0, t1) of the group, A operand columns in src (LDS,
// stride LD), B from global.  TO_LDS: write into dst LDS columns (task c_off
// = first column; ceil4(N) columns), else into C with the task's maps, rows
// split by the unit's precomputed (mb, mr).  Each wave walks its tasks
// (t0 + wave, step 4) as one flat sequence of items -- an item is a
// 4*KS-deep k slice of one pair for one 32-column tile -- and loads
// the operands of item i+1 (possibly of the next pair or task) while the
// MFMAs of item i run; the two operand slices alternate roles, so nothing
// is copied between them.
template <typename T, int RB, bool TO_LDS, bool GA, bool N16>
struct ChainWalk {
  typedef typename Acc<T>::type accv;
  static constexpr int LD = 16 * RB + 1;
  static constexpr int KS = ChainDepth<T>::value;
  static constexpr int KD = 4 * KS;
  static constexpr int TW = N16 ? 16 : 32;  // column tile width
  static constexpr int NJ = N16 ? 1 : 2;    // 16-column MFMA blocks per tile
  const zmpo_task* __restrict__ tasks;
  const zmpo_pair* __restrict__ pairs;
  const T* __restrict__ src;
  const T* __restrict__ Bb;
  T* __restrict__ dst;
  T* __restrict__ C;
  const T* __restrict__ Ag;
  const zmpo_cblock* __restrict__ blocks;
  ChainRows<RB> rows;
  int64_t c_base;
  int t1, beta, lane_a, lr, lk;
  // cursor of the item whose operands are being loaded
  int t, n0, p, k0, tN, tpb, tpe;
  ChainPair pr;
  accv acc[RB][2];

  // load the item under the cursor into sl; returns its k-step count
  __device__ __forceinline__ int load(ChainSlice<T, RB>& sl) {
    if (p >= tpe) return 0;
    chain_load<T, RB, GA, N16>(sl, pr, n0, k0, src, Bb, lane_a, lr, lk, Ag, blocks, rows);
    return min(KS, (pr.K - k0 + 3) / 4);
  }
  // move the cursor to the next item; returns whether the previous item
  // closed its (task, column tile), and sets (ct, cn0, cN) to that tile
  __device__ __forceinline__ bool advance(int& ct, int& cn0, int& cN) {
    ct = t;
    cn0 = n0;
    cN = tN;
    if (p < tpe && k0 + KD < pr.K) {
      k0 += KD;
      return false;
    }
    k0 = 0;
    ++p;
    if (p < tpe) {
      pr.get(pairs, p);
      return false;
    }
    p = tpb;
    n0 += TW;
    if (n0 < tN) {
      if (p < tpe) pr.get(pairs, p);
      return true;
    }
    n0 = 0;
    t += 4;
    if (t < t1) {
      tN = tasks[t].N;
      tpb = tasks[t].p_begin;
      tpe = tasks[t].p_end;
      p = tpb;
      if (p < tpe) pr.get(pairs, p);
    }
    return true;
  }
  __device__ __forceinline__ void mfma(const ChainSlice<T, RB>& sl, int ks) {
    constexpr int KG = Acc<T>::KG;
#pragma unroll
    for (int s = 0; s < KS; s += KG) {
      if (s >= ks) break;  // uniform
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          if constexpr (KG == 1) {
            acc[i][j] = Acc<T>::mfma(sl.a[s][i], sl.b[s][j], acc[i][j]);
          } else {
            // k-steps at or past ks (re-read operands) contribute zero
            const T z = Num<T>::zero();
            acc[i][j] = Acc<T>::mfma4(
                sl.a[s][i], s + 1 < ks ? (T)sl.a[s + 1][i] : z, s + 2 < ks ? (T)sl.a[s + 2][i] : z,
                s + 3 < ks ? (T)sl.a[s + 3][i] : z, sl.b[s][j], s + 1 < ks ? (T)sl.b[s + 1][j] : z,
                s + 2 < ks ? (T)sl.b[s + 2][j] : z, s + 3 < ks ? (T)sl.b[s + 3][j] : z, acc[i][j]);
          }
        }
    }
  }
  __device__ __forceinline__ void flush(int ct, int cn0, int cN, const int (&mb)[RB][4],
                                        const int (&mr)[RB][4], const bool (&rok)[RB][4]) {
    const zmpo_task& tk = tasks[ct];
    if (!TO_LDS) {
      T* __restrict__ cb = C + c_base + tk.c_off;
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (rok[i][r]) {
            const int64_t ro = (int64_t)mb[i][r] * tk.c_smhi + (int64_t)mr[i][r] * tk.c_smlo;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
              const int n = cn0 + j * 16 + lr;
              if (n < cN) store_acc(cb + ro + (int64_t)n * tk.c_sn, acc[i][j][r], beta);
            }
          }
        }
    } else {
      const int co = tk.c_off;
      const int w4 = (cN + 3) & ~3;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int n = cn0 + j * 16 + lr;
        if (n >= w4) continue;
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            dst[(co + n) * LD + i * 16 + Acc<T>::row(lk, r)] =
                Num<T>::put((typename Num<T>::M)acc[i][j][r]);
      }
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) acc[i][0] = acc[i][1] = accv{0, 0, 0, 0};
  }
};

template <typename T, int RB, bool TO_LDS, bool GA, bool N16>
__device__ __forceinline__ void chain_stage(
    int t0, int t1, const zmpo_task* __restrict__ tasks, const zmpo_pair* __restrict__ pairs,
    const T* __restrict__ src, const T* __restrict__ Bb, T* __restrict__ dst,
    T* __restrict__ C, int64_t c_base, const int (&mb)[RB][4], const int (&mr)[RB][4],
    const bool (&rok)[RB][4], int beta, int wv, int lr, int lk, const T* __restrict__ Ag,
    const zmpo_cblock* __restrict__ blocks, const ChainRows<RB>& rows) {
  typedef typename Acc<T>::type accv;
  ChainWalk<T, RB, TO_LDS, GA, N16> w;
  w.Ag = Ag;
  w.blocks = blocks;
  w.rows = rows;
  w.t = t0 + wv;
  if (w.t >= t1) return;  // wave-uniform
  w.tasks = tasks;
  w.pairs = pairs;
  w.src = src;
  w.Bb = Bb;
  w.dst = dst;
  w.C = C;
  w.c_base = c_base;
  w.t1 = t1;
  w.beta = beta;
  w.lane_a = lk * (16 * RB + 1) + lr;
  w.lr = lr;
  w.lk = lk;
  w.tN = tasks[w.t].N;
  w.tpb = tasks[w.t].p_begin;
  w.tpe = tasks[w.t].p_end;
  w.n0 = 0;
  w.k0 = 0;
  w.p = w.tpb;
  w.pr.K = 0;
  if (w.p < w.tpe) w.pr.get(pairs, w.p);
#pragma unroll
  for (int i = 0; i < RB; ++i) w.acc[i][0] = w.acc[i][1] = accv{0, 0, 0, 0};
  ChainSlice<T, RB> s0, s1;
  int k_0 = w.load(s0), k_1 = 0;
  int ct, cn0, cN;
  while (true) {
    bool last = w.advance(ct, cn0, cN);
    bool have = w.t < t1;
    if (have) k_1 = w.load(s1);
    w.mfma(s0, k_0);
    if (last) w.flush(ct, cn0, cN, mb, mr, rok);
    if (!have) break;
    last = w.advance(ct, cn0, cN);
    have = w.t < t1;
    if (have) k_0 = w.load(s0);
    w.mfma(s1, k_1);
    if (last) w.flush(ct, cn0, cN, mb, mr, rok);
    if (!have) break;
  }
}

template <typename T, int RB, bool ATILE, bool N16>
__global__ __launch_bounds__(256) void chain_gemm_kernel(
    const zmpo_cunit* __restrict__ units, int64_t nunits, const zmpo_cgroup* __restrict__ groups,
    const zmpo_cblock* __restrict__ blocks, const zmpo_task* __restrict__ tasks,
    const zmpo_pair* __restrict__ pairs, const T* __restrict__ A, const T* __restrict__ B1,
    const T* __restrict__ B2, T* __restrict__ C, int beta, int kcols_max) {
  constexpr int TMv = 16 * RB, LD = TMv + 1;
  T* __restrict__ atile = reinterpret_cast<T*>(chain_lds);
  T* __restrict__ mid = atile + (ATILE ? (int64_t)kcols_max * LD : 0);
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t u = xcd_logical_block(blockIdx.x, gridDim.x);
  if (u >= nunits) return;  // workgroup-uniform
  const zmpo_cunit un = units[u];
  const zmpo_cgroup g = groups[un.group];
  const int mv = min(TMv, g.M - un.m0);

  // ---- stage 0: A blocks -> atile.  Lane rows: r0 = lane % TM.
  if (ATILE) {
    const int r0 = lane & (TMv - 1);
    const bool ok0 = r0 < mv;
    const int m = un.m0 + r0;
    const int qb = ok0 ? m / g.rsplit : 0, qr = ok0 ? m % g.rsplit : 0;
    for (int j = g.ab_begin + wv; j < g.ab_end; j += 4) {
      const zmpo_cblock bk = blocks[j];
      const T* __restrict__ a = A + un.a_base + bk.a_off +
                                ((int64_t)qb * bk.a_smhi + (int64_t)qr * bk.a_smlo);
      const int kp = (bk.K + 3) & ~3;
      const int ne = TMv * kp;
      for (int e0 = 0; e0 < ne; e0 += 64 * 8) {
        T v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = (e0 + i * 64 + lane) / TMv;
          v[i] = (ok0 && k < bk.K) ? a[(int64_t)k * bk.a_sk] : Num<T>::zero();
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int e = e0 + i * 64 + lane;
          if (e < ne) atile[(bk.col + e / TMv) * LD + r0] = v[i];
        }
      }
    }
  }
  // this lane's output rows (stage 2), split once
  int mb[RB][4], mr[RB][4];
  bool rok[RB][4];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i * 16 + Acc<T>::row(lk, r);
      rok[i][r] = row < mv;
      const int m = un.m0 + row;
      mb[i][r] = rok[i][r] ? m / g.rsplit : 0;
      mr[i][r] = rok[i][r] ? m % g.rsplit : 0;
    }
  // this lane's A rows (MFMA operand layout: row lr of each 16-row block)
  ChainRows<RB> arows;
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int row = i * 16 + lr;
    const int m = un.m0 + row;
    arows.mb[i] = row < mv ? m / g.rsplit : 0;
    arows.mr[i] = row < mv ? m % g.rsplit : 0;
  }
  if (ATILE) __syncthreads();
  // ---- stage 1: intermediate columns into LDS
  chain_stage<T, RB, true, !ATILE, N16>(g.t1_begin, g.t1_end, tasks, pairs, atile, B1 + un.b1_base,
                                   mid, nullptr, 0, mb, mr, rok, 0, wv, lr, lk, A + un.a_base,
                                   blocks, arows);
  __syncthreads();
  // ---- stage 2: output blocks
  chain_stage<T, RB, false, false, N16>(g.t2_begin, g.t2_end, tasks, pairs, mid, B2 + un.b2_base,
                                   nullptr, C, un.c_base, mb, mr, rok, beta, wv, lr, lk,
                                   nullptr, blocks, arows);
}

template <typename T, int RB, bool ATILE, bool N16>
int launch_chain_t(int kcols_max, int ncols_max, const zmpo_cunit* units, int64_t nunits,
                   const zmpo_cgroup* groups, const zmpo_cblock* blocks, const zmpo_task* tasks,
                   const zmpo_pair* pairs, const void* A, const void* B1, const void* B2, void* C,
                   int beta, hipStream_t s) {
  const size_t lds = (size_t)((ATILE ? kcols_max : 0) + ncols_max) * (16 * RB + 1) * sizeof(T);
  if (lds > 160 * 1024) return set_error(ZMPO_ERR_ARG, "chaingemm: row tile exceeds 160 KiB of LDS");
  if (lds > 64 * 1024) {
    // the LDS allowance holds for the rest of the process: raise it only
    // when a launch needs more than this device already allows, so repeated
    // launches (and launches captured into a graph) make no attribute call
    constexpr int MAXDEV = 64;
    static int allowed[MAXDEV] = {0};
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "chaingemm: device");
    if (rc) return rc;
    if (dev < 0 || dev >= MAXDEV) return set_error(ZMPO_ERR_ARG, "chaingemm: device index beyond 63");
    if ((int)lds > allowed[dev]) {
      rc = check_hip(hipFuncSetAttribute((const void*)chain_gemm_kernel<T, RB, ATILE, N16>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                     "chaingemm: LDS attribute");
      if (rc) return rc;
      allowed[dev] = (int)lds;
    }
  }
  hipLaunchKernelGGL((chain_gemm_kernel<T, RB, ATILE, N16>), dim3((unsigned)nunits), dim3(256), lds, s,
                     units, nunits, groups, blocks, tasks, pairs, (const T*)A, (const T*)B1,
                     (const T*)B2, (T*)C, beta, kcols_max);
  return check_hip(hipGetLastError(), "chain_gemm launch");
}

template <typename T, int RB, bool N16>
int launch_chain(int atile, int kcols_max, int ncols_max, const zmpo_cunit* units,
                 int64_t nunits, const zmpo_cgroup* groups, const zmpo_cblock* blocks,
                 const zmpo_task* tasks, const zmpo_pair* pairs, const void* A, const void* B1,
                 const void* B2, void* C, int beta, hipStream_t s) {
  if (atile)
    return launch_chain_t<T, RB, true, N16>(kcols_max, ncols_max, units, nunits, groups, blocks,
                                            tasks, pairs, A, B1, B2, C, beta, s);
  return launch_chain_t<T, RB, false, N16>(kcols_max, ncols_max, units, nunits, groups, blocks,
                                           tasks, pairs, A, B1, B2, C, beta, s);
}

// ---------------------------------------------------------------- vectors
// out[j] = sum_i X[j*ldx + i] * y[i], j < nx.  Deterministic two-level
// reduction: block b sums a fixed contiguous chunk, the second pass adds the
// per-block partials in block order.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void dots_partial_kernel(const T* __restrict__ X,
                                                                   int64_t ldx, int nx,
                                                                   const T* __restrict__ y,
                                                                   int64_t n, double* __restrict__ part) {
  __shared__ double red[RED_THREADS];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t beg = (int64_t)blockIdx.x * chunk;
  const int64_t end = beg + chunk < n ? beg + chunk : n;
  for (int j = 0; j < nx; ++j) {
    double s = 0.0;
    const T* __restrict__ xj = X + (int64_t)j * ldx;
    for (int64_t i = beg + threadIdx.x; i < end; i += RED_THREADS) s += (double)xj[i] * (double)y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(RED_THREADS) void dots_final_kernel(const double* __restrict__ part,
                                                                int nblk, int nx,
                                                                double* __restrict__ out) {
  __shared__ double red[RED_THREADS];
  const int j = blockIdx.x;
  if (j >= nx) return;
  double s = 0.0;
  // fixed order: thread t sums blocks t, t+256, ... then tree-reduce
  for (int b = threadIdx.x; b < nblk; b += RED_THREADS) s += part[(int64_t)j * nblk + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// out[i] = beta*out[i] + alpha * sum_j c[j] * X[j*ldx + i]   (c on the device)
template <typename T>
__global__ void lincomb_kernel(T* __restrict__ out, const T* __restrict__ X, int64_t ldx, int nx,
                               const double* __restrict__ c, double alpha, double beta,
                               int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < nx; ++j) s += c[j] * (double)X[(int64_t)j * ldx + i];
    out[i] = (T)(beta == 0.0 ? alpha * s : beta * (double)out[i] + alpha * s);
  }
}

// r[i] /= |diag[i] - e| + shift   (Davidson iop=4 preconditioner)
template <typename T>
__global__ void precond_kernel(T* __restrict__ r, const double* __restrict__ diag, double e,
                               double shift, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    r[i] = (T)((double)r[i] / (fabs(diag[i] - e) + shift));
}

// value conversion between storage types (FP64 -> BF16 rounds through FP32)
template <typename TD, typename TS>
__device__ __forceinline__ TD cvt(TS v) {
  return Num<TD>::put((typename Num<TD>::M)Num<TS>::get(v));
}

// dst[i] = src[idx[i]] (idx < 0 -> 0)
template <typename TD, typename TS>
__global__ void gather_kernel(TD* __restrict__ dst, const TS* __restrict__ src,
                              const int64_t* __restrict__ idx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx[i];
    dst[i] = j >= 0 ? cvt<TD>(src[j]) : Num<TD>::zero();
  }
}

// dst[idx[i]] = src[i]  (idx must be injective)
template <typename TD, typename TS>
__global__ void scatter_kernel(TD* __restrict__ dst, const TS* __restrict__ src,
                               const int64_t* __restrict__ idx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[idx[i]] = cvt<TD>(src[i]);
}

template <typename TD, typename TS>
__global__ void convert_kernel(TD* __restrict__ dst, const TS* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = cvt<TD>(src[i]);
}

inline unsigned grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename T>
int launch_blockgemm(const zmpo_task* tasks, int64_t ntasks, const zmpo_pair* pairs,
                     int64_t ntiles, const void* A, const void* B, void* C, int beta,
                     hipStream_t s) {
  if (ntiles <= 0 || ntasks <= 0) return ZMPO_OK;
  const int64_t cap = (int64_t)1 << 30;
  const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
  hipLaunchKernelGGL(blockgemm_kernel<T>, dim3(grid), dim3(NTHREADS), 0, s, tasks, (int)ntasks,
                     pairs, (const T*)A, (const T*)B, (T*)C, beta, ntiles);
  return check_hip(hipGetLastError(), "blockgemm launch");
}

// Reference-compatible block rule for one rank (qsym.c:24-75 semantics).
template <typename Q>
void sym_allowed_impl(int rank, int nqnum, const int* ndims, const int* status, int* allowed,
                      const Q* const* q, double tol) {
  int64_t nblk = 1;
  for (int i = 0; i < rank; ++i) nblk *= ndims[i];
  int idx[16] = {0};
  double diff[16];
  for (int64_t b = 0; b < nblk; ++b) {
    for (int iq = 0; iq < nqnum; ++iq) diff[iq] = 0.0;
    for (int r = 0; r < rank; ++r) {
      const Q* row = q[r] + (int64_t)idx[r] * nqnum;
      for (int iq = 0; iq < nqnum; ++iq)
        diff[iq] += status[r] == 0 ? (double)row[iq] : -(double)row[iq];
    }
    double tot = 0.0;
    for (int iq = 0; iq < nqnum; ++iq) tot += fabs(diff[iq]);
    allowed[b] = tot < tol;
    for (int r = rank - 1; r >= 0; --r) {  // C-order odometer
      if (++idx[r] < ndims[r]) break;
      idx[r] = 0;
    }
  }
}

}  // namespace

// ---------------------------------------------------- allocation-free launch
// Split-K workspace of zmpo_streamgemm_static: a file-scope device array,
// one copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: its device address is looked up once per
// device (hipGetSymbolAddress, that device current) and kept in
// g_scratch_ptr.  No slot needs to start at zero (every slot a launch uses
// is written in full by its unit before the reduction reads it), so nothing
// is cleared.  Calls that overlap on one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[ZMPO_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAXDEV = 64;
void* g_scratch_ptr[SCRATCH_MAXDEV] = {nullptr};
std::mutex g_scratch_mu;

int scratch_ptr(void** out) {
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "streamgemm_static: device");
  if (rc) return rc;
  if (dev < 0 || dev >= SCRATCH_MAXDEV)
    return set_error(ZMPO_ERR_ARG, "streamgemm_static: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                   "streamgemm_static: scratch address");
    if (rc) return rc;
    if (!p) return set_error(ZMPO_ERR_HIP, "streamgemm_static: scratch address is null");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ZMPO_OK;
}
}  // namespace

extern "C" {

const char* zmpo_last_error(void) { return g_last_error.c_str(); }

int zmpo_abi_version(void) { return ZMPO_ABI_VERSION; }

int zmpo_struct_sizes(int* task_size, int* pair_size) {
  if (!task_size || !pair_size) return set_error(ZMPO_ERR_ARG, "null pointer");
  *task_size = (int)sizeof(zmpo_task);
  *pair_size = (int)sizeof(zmpo_pair);
  return ZMPO_OK;
}

int zmpo_symAllowed(int rank, int nqnum, int* ndims, int* status, int* allowed, double* a0,
                    double* a1, double* a2, double* a3, double* a4, double* a5, double* a6,
                    double* a7) {
  if (rank < 2 || rank > 8) return set_error(ZMPO_ERR_ARG, "symAllowed: rank must be in [2,8]");
  if (nqnum < 1 || nqnum > 16) return set_error(ZMPO_ERR_ARG, "symAllowed: nqnum out of range");
  if (!ndims || !status || !allowed) return set_error(ZMPO_ERR_ARG, "symAllowed: null pointer");
  const double* q[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
  for (int i = 0; i < rank; ++i)
    if (!q[i]) return set_error(ZMPO_ERR_ARG, "symAllowed: null QN array");
  sym_allowed_impl<double>(rank, nqnum, ndims, status, allowed, q, 1e-10);
  return ZMPO_OK;
}

int zmpo_symAllowedInt(int rank, int nqnum, const int* ndims, const int* status, int* allowed,
                       const int32_t* const* qsyms) {
  if (rank < 1 || rank > 16) return set_error(ZMPO_ERR_ARG, "symAllowedInt: rank must be in [1,16]");
  if (nqnum < 1 || nqnum > 16) return set_error(ZMPO_ERR_ARG, "symAllowedInt: nqnum out of range");
  if (!ndims || !status || !allowed || !qsyms) return set_error(ZMPO_ERR_ARG, "symAllowedInt: null pointer");
  for (int i = 0; i < rank; ++i)
    if (!qsyms[i]) return set_error(ZMPO_ERR_ARG, "symAllowedInt: null QN array");
  sym_allowed_impl<int32_t>(rank, nqnum, ndims, status, allowed, qsyms, 0.5);
  return ZMPO_OK;
}

int zmpo_blockgemm(int dtype, const zmpo_task* tasks, int64_t ntasks, const zmpo_pair* pairs,
                   int64_t ntiles, const void* A, const void* B, void* C, int beta,
                   void* stream) {
  if (ntasks > 0 && (!tasks || !pairs || !A || !B || !C))
    return set_error(ZMPO_ERR_ARG, "blockgemm: null pointer");
  if (ntasks > 0x7fffffff) return set_error(ZMPO_ERR_ARG, "blockgemm: too many tasks");
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case ZMPO_F64: return launch_blockgemm<double>(tasks, ntasks, pairs, ntiles, A, B, C, beta, s);
    case ZMPO_F32: return launch_blockgemm<float>(tasks, ntasks, pairs, ntiles, A, B, C, beta, s);
    default: return set_error(ZMPO_ERR_ARG, "blockgemm: unsupported dtype");
  }
}

int zmpo_wavegemm(int dtype, int rb, const zmpo_task* tasks, const zmpo_unit* units,
                  int64_t nunits, const zmpo_rgroup* groups, int64_t ngroups,
                  const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                  int beta, void* stream) {
  if (nunits > 0 && (!tasks || !units || !pairs || !A || !B || !C))
    return set_error(ZMPO_ERR_ARG, "wavegemm: null pointer");
  if (ngroups > 0 && (!groups || !work)) return set_error(ZMPO_ERR_ARG, "wavegemm: null workspace");
  if ((nunits + 3) / 4 > 0x7fffffffLL || (ngroups + 3) / 4 > 0x7fffffffLL)
    return set_error(ZMPO_ERR_ARG, "wavegemm: too many units");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64) {
    switch (rb) {
      case 1: return launch_wave<double, 1>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      case 2: return launch_wave<double, 2>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      case 4: return launch_wave<double, 4>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      default: break;
    }
  } else if (dtype == ZMPO_F32) {
    switch (rb) {
      case 1: return launch_wave<float, 1>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      case 2: return launch_wave<float, 2>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      case 4: return launch_wave<float, 4>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
      default: break;
    }
  }
  return set_error(ZMPO_ERR_ARG, "wavegemm: unsupported dtype or row-block count");
}

int zmpo_shapedgemm(int dtype, int shape, const zmpo_task* tasks, const zmpo_unit* units,
                    int64_t nunits, const zmpo_rgroup* groups, int64_t ngroups,
                    const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                    int beta, void* stream) {
  if (nunits > 0 && (!tasks || !units || !pairs || !A || !B || !C))
    return set_error(ZMPO_ERR_ARG, "shapedgemm: null pointer");
  if (ngroups > 0 && (!groups || !work)) return set_error(ZMPO_ERR_ARG, "shapedgemm: null workspace");
  if (nunits > 0x7fffffffLL || ngroups > 0x7fffffffLL)
    return set_error(ZMPO_ERR_ARG, "shapedgemm: too many units");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64)
    return dispatch_shaped<double>(shape, tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_F32)
    return dispatch_shaped<float>(shape, tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_BF16)
    return dispatch_shaped<bf16s>(shape, tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_BF16_F32)
    return dispatch_shaped<bf16s, float>(shape, tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  return set_error(ZMPO_ERR_ARG, "shapedgemm: unsupported dtype");
}

int zmpo_dotgemm(int dtype, const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                 const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                 const void* A, const void* B, void* C, void* work, int beta, void* stream) {
  if (nunits > 0 && (!tasks || !units || !pairs || !A || !B || !C))
    return set_error(ZMPO_ERR_ARG, "dotgemm: null pointer");
  if (ngroups > 0 && (!groups || !work)) return set_error(ZMPO_ERR_ARG, "dotgemm: null workspace");
  if (nunits > 0x7fffffffLL || ngroups > 0x7fffffffLL)
    return set_error(ZMPO_ERR_ARG, "dotgemm: too many units");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64)
    return launch_dot<double>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_F32)
    return launch_dot<float>(tasks, units, nunits, groups, ngroups, pairs, A, B, C, work, beta, s);
  return set_error(ZMPO_ERR_ARG, "dotgemm: unsupported dtype");
}

int zmpo_skinnygemm(int dtype, int rb, const zmpo_task* tasks, const zmpo_unit* units,
                    int64_t nunits, const zmpo_pair* pairs, const void* A, const void* B, void* C,
                    int beta, void* stream) {
  if (nunits <= 0) return ZMPO_OK;
  if (!tasks || !units || !pairs || !A || !B || !C)
    return set_error(ZMPO_ERR_ARG, "skinnygemm: null pointer");
  if (nunits > 0x7fffffffLL) return set_error(ZMPO_ERR_ARG, "skinnygemm: too many units");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64) {
    if (rb == 1) return launch_skinny<double, 1>(tasks, units, nunits, pairs, A, B, C, beta, s);
    if (rb == 2) return launch_skinny<double, 2>(tasks, units, nunits, pairs, A, B, C, beta, s);
  } else if (dtype == ZMPO_F32) {
    if (rb == 1) return launch_skinny<float, 1>(tasks, units, nunits, pairs, A, B, C, beta, s);
    if (rb == 2) return launch_skinny<float, 2>(tasks, units, nunits, pairs, A, B, C, beta, s);
  }
  return set_error(ZMPO_ERR_ARG, "skinnygemm: unsupported dtype or row-block count");
}

static int chaingemm(bool n16, int dtype, int rb, int atile, int kcols_max, int ncols_max,
                     const zmpo_cunit* units, int64_t nunits, const zmpo_cgroup* groups,
                     const zmpo_cblock* blocks, const zmpo_task* tasks, const zmpo_pair* pairs,
                     const void* A, const void* B1, const void* B2, void* C, int beta,
                     void* stream) {
  if (nunits <= 0) return ZMPO_OK;
  if (!units || !groups || !blocks || !tasks || !pairs || !A || !B1 || !B2 || !C)
    return set_error(ZMPO_ERR_ARG, "chaingemm: null pointer");
  if (nunits > 0x7fffffffLL) return set_error(ZMPO_ERR_ARG, "chaingemm: too many units");
  if (ncols_max < 1 || kcols_max < 1)
    return set_error(ZMPO_ERR_ARG, "chaingemm: kcols_max and ncols_max must be >= 1");
  hipStream_t s = (hipStream_t)stream;
#define ZMPO_CHAIN(T, RB)                                                                      \
  return n16 ? launch_chain<T, RB, true>(atile, kcols_max, ncols_max, units, nunits, groups,  \
                                         blocks, tasks, pairs, A, B1, B2, C, beta, s)         \
             : launch_chain<T, RB, false>(atile, kcols_max, ncols_max, units, nunits, groups, \
                                          blocks, tasks, pairs, A, B1, B2, C, beta, s)
  if (dtype == ZMPO_F64) {
    if (rb == 1) ZMPO_CHAIN(double, 1);
    if (rb == 2) ZMPO_CHAIN(double, 2);
  } else if (dtype == ZMPO_F32) {
    if (rb == 1) ZMPO_CHAIN(float, 1);
    if (rb == 2) ZMPO_CHAIN(float, 2);
  } else if (dtype == ZMPO_BF16) {
    if (rb == 1) ZMPO_CHAIN(bf16s, 1);
    if (rb == 2) ZMPO_CHAIN(bf16s, 2);
  }
#undef ZMPO_CHAIN
  return set_error(ZMPO_ERR_ARG, "chaingemm: unsupported dtype or row-block count");
}

int zmpo_chaingemm(int dtype, int rb, int atile, int kcols_max, int ncols_max,
                   const zmpo_cunit* units, int64_t nunits, const zmpo_cgroup* groups,
                   const zmpo_cblock* blocks, const zmpo_task* tasks, const zmpo_pair* pairs,
                   const void* A, const void* B1, const void* B2, void* C, int beta,
                   void* stream) {
  return chaingemm(false, dtype, rb, atile, kcols_max, ncols_max, units, nunits, groups, blocks,
                   tasks, pairs, A, B1, B2, C, beta, stream);
}

int zmpo_chaingemm16(int dtype, int rb, int atile, int kcols_max, int ncols_max,
                     const zmpo_cunit* units, int64_t nunits, const zmpo_cgroup* groups,
                     const zmpo_cblock* blocks, const zmpo_task* tasks, const zmpo_pair* pairs,
                     const void* A, const void* B1, const void* B2, void* C, int beta,
                     void* stream) {
  return chaingemm(true, dtype, rb, atile, kcols_max, ncols_max, units, nunits, groups, blocks,
                   tasks, pairs, A, B1, B2, C, beta, stream);
}

int zmpo_streamgemm(int dtype, int shape, int kt, const zmpo_sunit* units, int64_t nunits,
                    int upw, const zmpo_task* tasks, const zmpo_rgroup* groups,
                    int64_t ngroups, const zmpo_pair* pairs, const void* A, const void* B,
                    void* C, void* work, int beta, void* stream) {
  if (nunits > 0 && (!units || !pairs || !A || !B || !C))
    return set_error(ZMPO_ERR_ARG, "streamgemm: null pointer");
  if (ngroups > 0 && (!groups || !work || !tasks))
    return set_error(ZMPO_ERR_ARG, "streamgemm: null workspace");
  if (upw < 1 || upw > MAXUPW) return set_error(ZMPO_ERR_ARG, "streamgemm: upw must be in [1, 64]");
  if (nunits < 0 || ngroups < 0 || (nunits + upw - 1) / upw > 0x7fffffffLL ||
      ngroups > 0x7fffffffLL)
    return set_error(ZMPO_ERR_ARG, "streamgemm: bad unit or group count");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64)
    return dispatch_stream<double>(shape, kt, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_F32)
    return dispatch_stream<float>(shape, kt, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_BF16)
    return dispatch_stream<bf16s>(shape, kt, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  if (dtype == ZMPO_BF16_F32)
    return dispatch_stream<bf16s, float>(shape, kt, units, nunits, upw, tasks, groups, ngroups, pairs, A, B, C, work, beta, s);
  return set_error(ZMPO_ERR_ARG, "streamgemm: unsupported dtype");
}

int64_t zmpo_scratch_bytes(void) { return (int64_t)ZMPO_SCRATCH_BYTES; }

int zmpo_streamgemm_static(int dtype, int shape, int kt, const zmpo_sunit* units,
                           int64_t nunits, int upw, const zmpo_task* tasks,
                           const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                           const void* A, const void* B, void* C, int64_t work_elems, int beta,
                           void* stream) {
  const int64_t esz = dtype == ZMPO_F64 ? 8 : (dtype == ZMPO_F32 || dtype == ZMPO_BF16_F32) ? 4
                                             : dtype == ZMPO_BF16 ? 2 : 0;
  if (!esz) return set_error(ZMPO_ERR_ARG, "streamgemm_static: unsupported dtype");
  if (work_elems < 0 || (ngroups > 0 && work_elems == 0))
    return set_error(ZMPO_ERR_ARG, "streamgemm_static: work_elems must cover the split-K slots");
  if (work_elems > (int64_t)ZMPO_SCRATCH_BYTES / esz)
    return set_error(ZMPO_ERR_ARG,
                     "streamgemm_static: split-K workspace exceeds the 64 MiB static scratch");
  void* work = nullptr;
  if (ngroups > 0) {
    int rc = scratch_ptr(&work);
    if (rc) return rc;
  }
  return zmpo_streamgemm(dtype, shape, kt, units, nunits, upw, tasks, groups, ngroups, pairs, A,
                         B, C, work, beta, stream);
}

int zmpo_dots(int dtype, const void* X, int64_t ldx, int nx, const void* y, int64_t n,
              double* out, double* work, void* stream) {
  if (nx <= 0) return ZMPO_OK;
  if (!X || !y || !out || !work) return set_error(ZMPO_ERR_ARG, "dots: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = RED_BLOCKS;
  if (dtype == ZMPO_F64)
    hipLaunchKernelGGL(dots_partial_kernel<double>, dim3(nblk), dim3(RED_THREADS), 0, s,
                       (const double*)X, ldx, nx, (const double*)y, n, work);
  else if (dtype == ZMPO_F32)
    hipLaunchKernelGGL(dots_partial_kernel<float>, dim3(nblk), dim3(RED_THREADS), 0, s,
                       (const float*)X, ldx, nx, (const float*)y, n, work);
  else
    return set_error(ZMPO_ERR_ARG, "dots: unsupported dtype");
  int rc = check_hip(hipGetLastError(), "dots partial");
  if (rc) return rc;
  hipLaunchKernelGGL(dots_final_kernel, dim3(nx), dim3(RED_THREADS), 0, s, work, nblk, nx, out);
  return check_hip(hipGetLastError(), "dots final");
}

int zmpo_dots_workspace(int nx) { return nx * RED_BLOCKS; }

int zmpo_lincomb(int dtype, void* out, const void* X, int64_t ldx, int nx, const double* coef,
                 double alpha, double beta, int64_t n, void* stream) {
  if (!out || (nx > 0 && (!X || !coef))) return set_error(ZMPO_ERR_ARG, "lincomb: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64)
    hipLaunchKernelGGL(lincomb_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)out, (const double*)X, ldx, nx, coef, alpha, beta, n);
  else if (dtype == ZMPO_F32)
    hipLaunchKernelGGL(lincomb_kernel<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)out, (const float*)X, ldx, nx, coef, alpha, beta, n);
  else
    return set_error(ZMPO_ERR_ARG, "lincomb: unsupported dtype");
  return check_hip(hipGetLastError(), "lincomb");
}

int zmpo_precond(int dtype, void* r, const double* diag, double e, double shift, int64_t n,
                 void* stream) {
  if (!r || !diag) return set_error(ZMPO_ERR_ARG, "precond: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == ZMPO_F64)
    hipLaunchKernelGGL(precond_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)r, diag, e, shift, n);
  else if (dtype == ZMPO_F32)
    hipLaunchKernelGGL(precond_kernel<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)r, diag, e, shift, n);
  else
    return set_error(ZMPO_ERR_ARG, "precond: unsupported dtype");
  return check_hip(hipGetLastError(), "precond");
}

int zmpo_gather(int dtype_dst, void* dst, int dtype_src, const void* src, const int64_t* idx,
                int64_t n, void* stream) {
  if (n <= 0) return ZMPO_OK;
  if (!dst || !src || !idx) return set_error(ZMPO_ERR_ARG, "gather: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype_dst == ZMPO_F64 && dtype_src == ZMPO_F64)
    hipLaunchKernelGGL((gather_kernel<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)dst, (const double*)src, idx, n);
  else if (dtype_dst == ZMPO_F32 && dtype_src == ZMPO_F64)
    hipLaunchKernelGGL((gather_kernel<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)dst, (const double*)src, idx, n);
  else if (dtype_dst == ZMPO_F64 && dtype_src == ZMPO_F32)
    hipLaunchKernelGGL((gather_kernel<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)dst, (const float*)src, idx, n);
  else if (dtype_dst == ZMPO_F32 && dtype_src == ZMPO_F32)
    hipLaunchKernelGGL((gather_kernel<float, float>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)dst, (const float*)src, idx, n);
  else
    return set_error(ZMPO_ERR_ARG, "gather: unsupported dtype");
  return check_hip(hipGetLastError(), "gather");
}

int zmpo_scatter(int dtype_dst, void* dst, int dtype_src, const void* src, const int64_t* idx,
                 int64_t n, void* stream) {
  if (n <= 0) return ZMPO_OK;
  if (!dst || !src || !idx) return set_error(ZMPO_ERR_ARG, "scatter: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype_dst == ZMPO_F64 && dtype_src == ZMPO_F64)
    hipLaunchKernelGGL((scatter_kernel<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)dst, (const double*)src, idx, n);
  else if (dtype_dst == ZMPO_F32 && dtype_src == ZMPO_F64)
    hipLaunchKernelGGL((scatter_kernel<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)dst, (const double*)src, idx, n);
  else if (dtype_dst == ZMPO_F64 && dtype_src == ZMPO_F32)
    hipLaunchKernelGGL((scatter_kernel<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (double*)dst, (const float*)src, idx, n);
  else if (dtype_dst == ZMPO_F32 && dtype_src == ZMPO_F32)
    hipLaunchKernelGGL((scatter_kernel<float, float>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (float*)dst, (const float*)src, idx, n);
  else
    return set_error(ZMPO_ERR_ARG, "scatter: unsupported dtype");
  return check_hip(hipGetLastError(), "scatter");
}

int zmpo_convert(int dtype_dst, void* dst, int dtype_src, const void* src, int64_t n,
                 void* stream) {
  if (n <= 0) return ZMPO_OK;
  if (!dst || !src) return set_error(ZMPO_ERR_ARG, "convert: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const unsigned g = grid_for(n, 256);
#define ZMPO_CONV(CD, TD, CS, TS)                                                             \
  if (dtype_dst == CD && dtype_src == CS) {                                                   \
    hipLaunchKernelGGL((convert_kernel<TD, TS>), dim3(g), dim3(256), 0, s, (TD*)dst,          \
                       (const TS*)src, n);                                                    \
    return check_hip(hipGetLastError(), "convert");                                           \
  }
  ZMPO_CONV(ZMPO_F32, float, ZMPO_F64, double)
  ZMPO_CONV(ZMPO_F64, double, ZMPO_F32, float)
  ZMPO_CONV(ZMPO_BF16, bf16s, ZMPO_F64, double)
  ZMPO_CONV(ZMPO_BF16, bf16s, ZMPO_F32, float)
  ZMPO_CONV(ZMPO_F64, double, ZMPO_BF16, bf16s)
  ZMPO_CONV(ZMPO_F32, float, ZMPO_BF16, bf16s)
#undef ZMPO_CONV
  return set_error(ZMPO_ERR_ARG, "convert: unsupported dtype pair");
}

// ------------------------------------------------------- schedule builder
// Host-side construction of the unit schedule of a grouped-GEMM launch (the
// work items of zmpo_wavegemm / zmpo_shapedgemm / zmpo_streamgemm): one
// unit per (output tile, chunk of the task's pairs), chunks cut where the
// summed K of the preceding pairs of the task crosses a multiple of kchunk.
// Unit order: (task, tile, chunk), or (key, chunk, task, tile) when a reuse
// key per task is given; split-K slots are numbered in (task, tile, chunk)
// order over the tasks with more than one chunk; one reduction group per
// split tile, in unit order.  tasks / pairs / reuse_key are host arrays.
int zmpo_schedule_build(int64_t nt, const zmpo_task* tasks, const zmpo_pair* pairs, int tm,
                        int tn, int64_t kchunk, const int64_t* reuse_key, zmpo_unit* units,
                        int64_t units_cap, zmpo_rgroup* groups, int64_t groups_cap,
                        int64_t* counts) {
  if (nt < 0 || tm <= 0 || tn <= 0 || kchunk <= 0 || !counts || (nt > 0 && !tasks))
    return set_error(ZMPO_ERR_ARG, "schedule_build: bad arguments");
  std::vector<int64_t> nch(nt), cfirst(nt + 1, 0), tiles(nt);
  std::vector<int32_t> cb, ce;
  cb.reserve(nt);
  ce.reserve(nt);
  for (int64_t t = 0; t < nt; ++t) {
    const zmpo_task& tk = tasks[t];
    const int64_t rows = tk.M > 0 ? ((int64_t)tk.M + tm - 1) / tm : 1;
    const int64_t cols = tk.N > 0 ? ((int64_t)tk.N + tn - 1) / tn : 1;
    tiles[t] = (rows > 0 ? rows : 1) * (cols > 0 ? cols : 1);
    cfirst[t] = (int64_t)cb.size();
    if (tk.p_end <= tk.p_begin) {
      cb.push_back(tk.p_begin);
      ce.push_back(tk.p_begin);
    } else {
      int64_t pre = 0, cur = -1;
      for (int32_t q = tk.p_begin; q < tk.p_end; ++q) {
        const int64_t cid = pre / kchunk;
        if (cid != cur) {
          cb.push_back(q);
          ce.push_back(q + 1);
          cur = cid;
        } else {
          ce.back() = q + 1;
        }
        pre += pairs[q].K;
      }
    }
    nch[t] = (int64_t)cb.size() - cfirst[t];
  }
  // split-K slot base per task ((task, tile, chunk) order)
  std::vector<int64_t> sbase(nt, -1);
  int64_t nslots = 0, nunits = 0, ngroups = 0;
  for (int64_t t = 0; t < nt; ++t) {
    nunits += tiles[t] * nch[t];
    if (nch[t] > 1) {
      sbase[t] = nslots;
      nslots += tiles[t] * nch[t];
      ngroups += tiles[t];
    }
  }
  counts[0] = nunits;
  counts[1] = ngroups;
  counts[2] = nslots;
  if (!units) return ZMPO_OK;  // count only
  if (units_cap < nunits || (ngroups > 0 && (!groups || groups_cap < ngroups)))
    return set_error(ZMPO_ERR_ARG, "schedule_build: output arrays too small");
  int64_t u = 0, gidx = 0;
  auto emit = [&](int64_t t, int64_t ti, int64_t c) {
    const int64_t cols = tiles[t] / ((int64_t)tasks[t].M > 0 ? ((int64_t)tasks[t].M + tm - 1) / tm : 1);
    zmpo_unit& un = units[u++];
    un.task = (int32_t)t;
    un.m0 = (int32_t)((ti / cols) * tm);
    un.n0 = (int32_t)((ti % cols) * tn);
    un.p_begin = cb[cfirst[t] + c];
    un.p_end = ce[cfirst[t] + c];
    un.slot = nch[t] > 1 ? (int32_t)(sbase[t] + ti * nch[t] + c) : -1;
    un.pad0 = un.pad1 = 0;
    if (nch[t] > 1 && c == 0) {
      zmpo_rgroup& g = groups[gidx++];
      g.task = (int32_t)t;
      g.m0 = un.m0;
      g.n0 = un.n0;
      g.slot_begin = un.slot;
      g.nslots = (int32_t)nch[t];
      g.pad = 0;
    }
  };
  if (!reuse_key) {
    for (int64_t t = 0; t < nt; ++t)
      for (int64_t ti = 0; ti < tiles[t]; ++ti)
        for (int64_t c = 0; c < nch[t]; ++c) emit(t, ti, c);
  } else {
    std::vector<int64_t> ord(nt);
    for (int64_t t = 0; t < nt; ++t) ord[t] = t;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int64_t a, int64_t b) { return reuse_key[a] < reuse_key[b]; });
    for (int64_t i = 0; i < nt;) {
      int64_t j = i, cmax = 0;
      while (j < nt && reuse_key[ord[j]] == reuse_key[ord[i]]) {
        cmax = nch[ord[j]] > cmax ? nch[ord[j]] : cmax;
        ++j;
      }
      for (int64_t c = 0; c < cmax; ++c)
        for (int64_t k = i; k < j; ++k) {
          const int64_t t = ord[k];
          if (c >= nch[t]) continue;
          for (int64_t ti = 0; ti < tiles[t]; ++ti) emit(t, ti, c);
        }
      i = j;
    }
  }
  if (u != nunits || gidx != ngroups) return set_error(ZMPO_ERR_ARG, "schedule_build: count mismatch");
  return ZMPO_OK;
}

// ------------------------------------------------- truncation eigensolver
// Eigenvalues (ascending) and eigenvectors of nsec symmetric sector
// matrices packed in A (sector s: n_s x n_s at offset off_s, any storage
// order since it is symmetric), by rocSOLVER's divide-and-conquer dsyevd on
// the caller's stream; eigenvector j of sector s overwrites
// A[off_s + j*n_s .. off_s + (j+1)*n_s).  W gets the eigenvalues at woff_s.
// E (>= max n_s doubles) and info (>= nsec ints) are device workspace; a
// nonzero info[s] after the stream completes means no convergence.
int zmpo_syevd_sectors(int nsec, const int64_t* n, const int64_t* off, const int64_t* woff,
                       double* A, double* W, double* E, int* info, void* stream) {
  if (nsec < 0) return set_error(ZMPO_ERR_ARG, "syevd_sectors: negative sector count");
  if (nsec == 0) return ZMPO_OK;
  if (!n || !off || !woff || !A || !W || !E || !info)
    return set_error(ZMPO_ERR_ARG, "syevd_sectors: null pointer");
  constexpr int MAXDEV = 64;
  static rocblas_handle handles[MAXDEV] = {nullptr};
  static std::mutex mu;
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "syevd_sectors: device");
  if (rc) return rc;
  if (dev < 0 || dev >= MAXDEV) return set_error(ZMPO_ERR_ARG, "syevd_sectors: device index beyond 63");
  rocblas_handle h = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success)
      return set_error(ZMPO_ERR_HIP, "syevd_sectors: rocblas_create_handle failed");
    h = handles[dev];
  }
  if (rocblas_set_stream(h, (hipStream_t)stream) != rocblas_status_success)
    return set_error(ZMPO_ERR_HIP, "syevd_sectors: rocblas_set_stream failed");
  for (int s = 0; s < nsec; ++s) {
    if (n[s] <= 0) continue;
    if (n[s] > 0x7fffffffLL) return set_error(ZMPO_ERR_ARG, "syevd_sectors: sector too large");
    const rocblas_status st =
        rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n[s],
                         A + off[s], (rocblas_int)n[s], W + woff[s], E, info + s);
    if (st != rocblas_status_success) return set_error(ZMPO_ERR_HIP, "syevd_sectors: rocsolver_dsyevd failed");
  }
  return check_hip(hipGetLastError(), "syevd_sectors");
}

// ------------------------------------------------------------ device store
int zmpo_alloc(void** out, int64_t bytes) {
  if (!out || bytes < 0) return set_error(ZMPO_ERR_ARG, "alloc: bad arguments");
  *out = nullptr;
  if (bytes == 0) return ZMPO_OK;
  return check_hip(hipMalloc(out, (size_t)bytes), "alloc");
}

int zmpo_free(void* p) {
  if (!p) return ZMPO_OK;
  return check_hip(hipFree(p), "free");
}

int zmpo_memcpy(void* dst, const void* src, int64_t bytes, int kind, void* stream) {
  if (bytes <= 0) return ZMPO_OK;
  if (!dst || !src) return set_error(ZMPO_ERR_ARG, "memcpy: null pointer");
  hipMemcpyKind k;
  switch (kind) {
    case ZMPO_H2D: k = hipMemcpyHostToDevice; break;
    case ZMPO_D2H: k = hipMemcpyDeviceToHost; break;
    case ZMPO_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return set_error(ZMPO_ERR_ARG, "memcpy: unknown direction");
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = check_hip(hipMemcpyAsync(dst, src, (size_t)bytes, k, s), "memcpy");
  if (rc || kind == ZMPO_D2D) return rc;
  // host buffers are pageable: the copy is complete when the call returns
  return check_hip(hipStreamSynchronize(s), "memcpy sync");
}

// ---------------------------------------------------------- launch sequences
// A recorded list of grouped-GEMM and chained launches over numbered operand
// slots: a whole sigma (or renormalisation, or diagonal) is one call that
// binds the slots to device buffers and issues every launch in order on one
// stream.  Tables are device pointers owned by the caller.
struct zmpo_seq {
  std::vector<zmpo_seq_op> ops;
};

}  // extern "C"

namespace {
int seq_check_slot(int x, int nslots) { return x < nslots; }
}  // namespace

extern "C" {

int zmpo_seq_create(zmpo_seq** out) {
  if (!out) return set_error(ZMPO_ERR_ARG, "seq_create: null pointer");
  *out = new (std::nothrow) zmpo_seq();
  return *out ? ZMPO_OK : set_error(ZMPO_ERR_ARG, "seq_create: out of host memory");
}

int zmpo_seq_destroy(zmpo_seq* s) {
  delete s;
  return ZMPO_OK;
}

int zmpo_seq_add(zmpo_seq* s, const zmpo_seq_op* op) {
  if (!s || !op) return set_error(ZMPO_ERR_ARG, "seq_add: null pointer");
  if (op->kind < ZMPO_SEQ_SHAPED || op->kind > ZMPO_SEQ_GATHER)
    return set_error(ZMPO_ERR_ARG, "seq_add: unknown launch kind");
  if (op->kind == ZMPO_SEQ_GATHER) {
    // FP64 gather C[i] = A[idx[i]], idx = op.tasks, n = op.nunits
    if (op->slot_a < 0 || op->slot_c < 0 || (op->nunits > 0 && !op->tasks))
      return set_error(ZMPO_ERR_ARG, "seq_add: gather needs slots a, c and an index array");
    zmpo_seq_op g = *op;
    g.slot_b = g.slot_a;  // no second operand
    g.slot_b2 = -1;
    g.slot_work = -1;
    s->ops.push_back(g);
    return ZMPO_OK;
  }
  if (op->slot_a < 0 || op->slot_b < 0 || op->slot_c < 0 ||
      ((op->kind == ZMPO_SEQ_CHAIN || op->kind == ZMPO_SEQ_CHAIN16) && op->slot_b2 < 0))
    return set_error(ZMPO_ERR_ARG, "seq_add: missing operand slot");
  s->ops.push_back(*op);
  return ZMPO_OK;
}

int64_t zmpo_seq_length(const zmpo_seq* s) { return s ? (int64_t)s->ops.size() : -1; }

int zmpo_seq_run(const zmpo_seq* s, int dtype, void* const* slots, int nslots, void* stream) {
  if (!s || (nslots > 0 && !slots)) return set_error(ZMPO_ERR_ARG, "seq_run: null pointer");
  for (const zmpo_seq_op& op : s->ops) {
    const int hi = op.slot_a > op.slot_b ? op.slot_a : op.slot_b;
    const int hi2 = op.slot_c > op.slot_b2 ? op.slot_c : op.slot_b2;
    if (!seq_check_slot(hi > hi2 ? hi : hi2, nslots) || !seq_check_slot(op.slot_work, nslots))
      return set_error(ZMPO_ERR_ARG, "seq_run: operand slot beyond nslots");
    void* A = slots[op.slot_a];
    void* B = slots[op.slot_b];
    void* C = slots[op.slot_c];
    void* W = op.slot_work >= 0 ? slots[op.slot_work] : nullptr;
    const int dt = (op.out_f32 && dtype == ZMPO_BF16) ? ZMPO_BF16_F32 : dtype;
    int rc = ZMPO_OK;
    switch (op.kind) {
      case ZMPO_SEQ_SHAPED:
        rc = zmpo_shapedgemm(dt, op.shape, (const zmpo_task*)op.tasks, (const zmpo_unit*)op.units,
                             op.nunits, (const zmpo_rgroup*)op.groups, op.ngroups,
                             (const zmpo_pair*)op.pairs, A, B, C, W, op.beta, stream);
        break;
      case ZMPO_SEQ_STREAM:
        rc = zmpo_streamgemm(dt, op.shape, op.kt, (const zmpo_sunit*)op.units, op.nunits, op.upw,
                             (const zmpo_task*)op.tasks, (const zmpo_rgroup*)op.groups, op.ngroups,
                             (const zmpo_pair*)op.pairs, A, B, C, W, op.beta, stream);
        break;
      case ZMPO_SEQ_GATHER:
        rc = zmpo_gather(ZMPO_F64, C, ZMPO_F64, A, (const int64_t*)op.tasks, op.nunits, stream);
        break;
      default:
        rc = (op.kind == ZMPO_SEQ_CHAIN ? zmpo_chaingemm : zmpo_chaingemm16)(
            dt, op.rb, op.atile, op.kmax, op.nmax, (const zmpo_cunit*)op.units, op.nunits,
            (const zmpo_cgroup*)op.groups, (const zmpo_cblock*)op.blocks,
            (const zmpo_task*)op.tasks, (const zmpo_pair*)op.pairs, A, B, slots[op.slot_b2], C,
            op.beta, stream);
    }
    if (rc) return rc;
  }
  return ZMPO_OK;
}

// Named entry points over recorded sequences (the operand order of the
// reference's calls).
int zmpo_sigma_two_site(const zmpo_seq* plan, int dtype, const void* L, const void* R,
                        const void* Wi_padded, const void* Wj_padded, const void* psi,
                        void* sigma, void* top1, void* top3, void* work, void* stream) {
  if (!plan || !L || !R || !Wi_padded || !Wj_padded || !psi || !sigma || !top1 || !top3)
    return set_error(ZMPO_ERR_ARG, "sigma_two_site: null pointer");
  void* slots[9] = {(void*)L,   (void*)R, (void*)Wi_padded, (void*)Wj_padded, (void*)psi,
                    sigma,      top1,     top3,             work};
  return zmpo_seq_run(plan, dtype, slots, 9, stream);
}

int zmpo_renorm(const zmpo_seq* plan, int dtype, const void* site, const void* env_in,
                const void* W, void* t1, void* t2, void* env_out, void* work, void* stream) {
  if (!plan || !site || !env_in || !W || !t1 || !t2 || !env_out)
    return set_error(ZMPO_ERR_ARG, "renorm: null pointer");
  void* slots[7] = {(void*)site, (void*)env_in, (void*)W, t1, t2, env_out, work};
  return zmpo_seq_run(plan, dtype, slots, 7, stream);
}

int zmpo_hdiag(const zmpo_seq* plan, const double* L, const double* R, const double* mpo_factors,
               double* ldiag, double* rdiag, double* x, double* diag, void* work, void* stream) {
  if (!plan || !L || !R || !mpo_factors || !ldiag || !rdiag || !x || !diag)
    return set_error(ZMPO_ERR_ARG, "hdiag: null pointer");
  void* slots[8] = {(void*)L, (void*)R, (void*)mpo_factors, ldiag, rdiag, x, diag, work};
  return zmpo_seq_run(plan, ZMPO_F64, slots, 8, stream);
}

int zmpo_davidson_residual(int64_t n, const double* W, const double* V, int64_t ld, int nb,
                           const double* coef, double theta, const double* diag, double shift,
                           double* r, void* stream) {
  if (n < 0 || nb < 0) return set_error(ZMPO_ERR_ARG, "davidson_residual: bad sizes");
  if (!r || (nb > 0 && (!W || !V || !coef)))
    return set_error(ZMPO_ERR_ARG, "davidson_residual: null pointer");
  int rc = zmpo_lincomb(ZMPO_F64, r, W, ld, nb, coef, 1.0, 0.0, n, stream);
  if (rc) return rc;
  rc = zmpo_lincomb(ZMPO_F64, r, V, ld, nb, coef, -theta, 1.0, n, stream);
  if (rc || !diag) return rc;
  return zmpo_precond(ZMPO_F64, r, diag, theta, shift, n, stream);
}

}  // extern "C"
