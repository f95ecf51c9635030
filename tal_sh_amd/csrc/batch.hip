// Persistent batched-contraction kernel (gfx950).
//
// Thousands of small contractions of uneven shape run in one launch.  Each
// contraction is described on the device by a BatchTask: operand pointers,
// the matrix sizes M (uncontracted L dims), N (uncontracted R dims), K
// (contracted dims), a tile shape chosen by the host autotuner, and offset
// tables that map a row / column / contracted index to an element offset in
// each tensor (lM, lK, rK, rN, dM, dN).  With those tables the operands are
// gathered straight from their dlf storage into LDS tiles, so no operand is
// transposed: D(m,n) (+)= alpha * sum_k L[lM[m]+lK[k]] * R[rK[k]+rN[n]] is
// written to D[dM[m]+dN[n]].  The reference runs such tasks one TTGT schedule
// (up to 4 transposes + a GEMM launch) at a time (tensor_algebra_gpu_nvidia.cu:
// 5847-7188).
//
// Work distribution: every (task, tile) pair is a work item; a fixed grid of
// workgroups pulls items from a global counter until the counter passes the
// total, which every workgroup reaches, so the grid always drains.  The
// counter is cleared by a one-thread kernel in the same stream (also inside
// a captured graph).  Each D element is produced by exactly one work item, so
// results are deterministic and independent of the batch composition.
#include "batch.hpp"

#include <hip/hip_runtime.h>

namespace tal {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename S>
struct BMfma;
template <>
struct BMfma<double> {
  using acc_t = f64x4;
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct BMfma<float> {
  using acc_t = f32x4;
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int kThreads = 256;
constexpr int kBK = 16;
constexpr int kMaxTile = 64;

__global__ void batch_reset_kernel(unsigned long long *counter) { *counter = 0ull; }

// One work item = one tm x tn tile of one task's D.  Element i of a thread
// in the A tile is (row am(i), step column ak(i)): along m (AM = 0: one row
// tid % tm, k = tid / tm + i * 256 / tm) when an uncontracted index is the
// contiguous one of L, along k (AM = 1: k = tid % 16, rows tid / 16 + 16 i)
// when a contracted index is, so consecutive lanes read consecutive
// addresses either way; likewise for B with n (BM).  At the start of a tile
// the workgroup copies the tile's row and column offsets and, when K <= kKC
// (KL = 1), the task's contracted-index offsets from the task tables into LDS
// with one load per thread and entry, all in flight at once; the K loop and
// the epilogue then read offsets from LDS, so a data gather waits for one
// global load, not two dependent ones.  The next K step is gathered into registers while the
// current one is multiplied from LDS (double-buffered, one barrier per step).
constexpr int kKC = 512;  // tasks with K <= kKC keep their contracted-index offsets in LDS
template <typename S>
struct TileSmem {
  S A[2][kMaxTile][kBK + 1];
  S B[2][kMaxTile][kBK + 1];
  int lm[kMaxTile], rn[kMaxTile], dm[kMaxTile], dn[kMaxTile];
  int lk[kKC], rk[kKC];
};

template <typename S, int AM, int BM, int KL>
__device__ __forceinline__ void tile_body(const BatchTask &T, long long local, const int *__restrict__ tabs,
                                          TileSmem<S> &sm, int tid) {
  using MF = BMfma<S>;
  using acc_t = typename MF::acc_t;
  constexpr int NR = kMaxTile * kBK / kThreads;  // elements per thread and operand (tile 64 x 16)
  const int lane = tid & 63, wave = tid >> 6;
  const int tm = T.tm, tn = T.tn, K = T.K, M = T.M, N = T.N;
  const int m0 = (int)(local % T.tiles_m) * tm, n0 = (int)(local / T.tiles_m) * tn;
  const int *lM = tabs + T.tab, *lK = lM + M, *rK = lK + K, *rN = rK + K, *dM = rN + N, *dN = dM + M;
  const S *L = (const S *)T.L, *R = (const S *)T.R;
  const int na = tm * kBK / kThreads, nb = tn * kBK / kThreads;
  const int tml = __builtin_ctz(tm), tnl = __builtin_ctz(tn);  // tm, tn are powers of two
  // row / column offsets of the tile (-1 past the edge); the previous tile's
  // epilogue read these arrays before the barrier that ended its item
  if (tid < kMaxTile) {
    const int m = m0 + tid, n = n0 + tid;
    sm.lm[tid] = (tid < tm && m < M) ? lM[m] : -1;
    sm.dm[tid] = (tid < tm && m < M) ? dM[m] : 0;
    sm.rn[tid] = (tid < tn && n < N) ? rN[n] : -1;
    sm.dn[tid] = (tid < tn && n < N) ? dN[n] : 0;
  }
  auto am = [&](int i) { return AM ? (tid >> 4) + kBK * i : tid & (tm - 1); };
  auto ak = [&](int i) { return AM ? tid & (kBK - 1) : (tid >> tml) + i * (kThreads >> tml); };
  auto bn = [&](int i) { return BM ? (tid >> 4) + kBK * i : tid & (tn - 1); };
  auto bk = [&](int i) { return BM ? tid & (kBK - 1) : (tid >> tnl) + i * (kThreads >> tnl); };
  const int subm = tm >> 4, nsub = subm * (tn >> 4);
  acc_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = acc_t{0, 0, 0, 0};
  // two register sets: step s is gathered into set s & 1 two steps ahead
  S ra0[NR], rb0[NR], ra1[NR], rb1[NR];
  int aoff[NR], boff[NR];
  const int *lk = KL ? sm.lk : lK, *rk = KL ? sm.rk : rK;
#define TAL_GATHER(RA, RB, k0)                                             \
  _Pragma("unroll") for (int i = 0; i < NR; ++i) {                         \
    const int ka = (k0) + ak(i), kb = (k0) + bk(i);                        \
    RA[i] = (aoff[i] >= 0 && ka < K) ? L[aoff[i] + lk[ka]] : (S)0;         \
    RB[i] = (boff[i] >= 0 && kb < K) ? R[rk[kb] + boff[i]] : (S)0;         \
  }
#define TAL_STAGE(RA, RB, buf)                                             \
  _Pragma("unroll") for (int i = 0; i < NR; ++i) {                         \
    if (i < na) sm.A[buf][am(i)][ak(i)] = RA[i];                           \
    if (i < nb) sm.B[buf][bn(i)][bk(i)] = RB[i];                           \
  }
  constexpr int UK = sizeof(S) == 8 ? 1 : 2;  // fewer live fragments for FP64 (register budget)
#define TAL_COMPUTE(buf)                                                                                   \
  _Pragma("unroll UK") for (int ks = 0; ks < kBK / 4; ++ks) {                                              \
    const int kk = ks * 4 + (lane >> 4);                                                                   \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                                        \
      const int sidx = wave + 4 * q;                                                                       \
      if (sidx < nsub) {                                                                                   \
        const int si = sidx % subm, sj = sidx / subm;                                                      \
        acc[q] = MF::op(sm.A[buf][si * 16 + (lane & 15)][kk], sm.B[buf][sj * 16 + (lane & 15)][kk], acc[q]); \
      }                                                                                                    \
    }                                                                                                      \
  }
  if (KL)  // K <= kKC: the contracted-index offsets of the whole task in LDS
    for (int k = tid; k < K; k += kThreads) {
      sm.lk[k] = lK[k];
      sm.rk[k] = rK[k];
    }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    aoff[i] = i < na ? sm.lm[am(i)] : -1;
    boff[i] = i < nb ? sm.rn[bn(i)] : -1;
  }
  const int nk = (K + kBK - 1) / kBK;
  TAL_GATHER(ra0, rb0, 0)
  TAL_STAGE(ra0, rb0, 0)
  __syncthreads();
  if (nk > 1) {
    TAL_GATHER(ra1, rb1, kBK)
  }
  if (nk > 2) {
    TAL_GATHER(ra0, rb0, 2 * kBK)
  }
  for (int kt = 0; kt < nk; kt += 2) {
    // even step kt from buffer 0; step kt+1 (set 1) goes to buffer 1, set 1 refills with kt+3
    TAL_COMPUTE(0)
    if (kt + 1 < nk) {
      TAL_STAGE(ra1, rb1, 1)
    }
    __syncthreads();
    if (kt + 3 < nk) {
      TAL_GATHER(ra1, rb1, (kt + 3) * kBK)
    }
    if (kt + 1 >= nk) break;
    // odd step kt+1 from buffer 1; step kt+2 (set 0) goes to buffer 0, set 0 refills with kt+4
    TAL_COMPUTE(1)
    if (kt + 2 < nk) {
      TAL_STAGE(ra0, rb0, 0)
    }
    __syncthreads();
    if (kt + 4 < nk) {
      TAL_GATHER(ra0, rb0, (kt + 4) * kBK)
    }
  }
#undef TAL_COMPUTE
#undef TAL_GATHER
#undef TAL_STAGE
  S *D = (S *)T.D;
  const S alpha = (S)T.alpha;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int sidx = wave + 4 * q;
    if (sidx >= nsub) continue;
    const int si = sidx % subm, sj = sidx / subm;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mi = si * 16 + MF::row(lane, r), ni = sj * 16 + (lane & 15);
      if (m0 + mi < M && n0 + ni < N) {
        S *d = D + sm.dm[mi] + sm.dn[ni];
        const S v = alpha * acc[q][r];
        *d = T.beta_one ? *d + v : v;
      }
    }
  }
}

// The index of the next work item is fetched (one atomic by thread 0) while
// the current tile is computed: the atomic's result stays in a register until
// the tile is done (the tile's barriers wait for LDS traffic only) and is then
// published through a two-slot LDS buffer, so neither the atomic's round trip
// nor a task search sits between tiles.
template <typename S>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void batch_contract_kernel(
    const BatchTask *__restrict__ tasks, const long long *__restrict__ prefix, const int *__restrict__ item_task,
    const int *__restrict__ tabs, unsigned long long *counter, long long total) {
  __shared__ TileSmem<S> sm;
  __shared__ long long s_item[2];
  const int tid = threadIdx.x;
  if (tid == 0) s_item[0] = (long long)atomicAdd(counter, 1ull);
  __syncthreads();
  long long item = s_item[0];
  for (int it = 1; item < total; ++it) {
    unsigned long long next = 0;
    if (tid == 0) next = atomicAdd(counter, 1ull);
    const int t = item_task[item];
    const BatchTask &T = tasks[t];
    const long long local = item - prefix[t];
    switch (T.amode * 2 + T.bmode + (T.K <= kKC ? 4 : 0)) {  // uniform over the block
      case 0: tile_body<S, 0, 0, 0>(T, local, tabs, sm, tid); break;
      case 1: tile_body<S, 0, 1, 0>(T, local, tabs, sm, tid); break;
      case 2: tile_body<S, 1, 0, 0>(T, local, tabs, sm, tid); break;
      case 3: tile_body<S, 1, 1, 0>(T, local, tabs, sm, tid); break;
      case 4: tile_body<S, 0, 0, 1>(T, local, tabs, sm, tid); break;
      case 5: tile_body<S, 0, 1, 1>(T, local, tabs, sm, tid); break;
      case 6: tile_body<S, 1, 0, 1>(T, local, tabs, sm, tid); break;
      default: tile_body<S, 1, 1, 1>(T, local, tabs, sm, tid); break;
    }
    // slot it & 1 was last read before the barriers of the previous tile
    if (tid == 0) s_item[it & 1] = (long long)next;
    __syncthreads();  // the next index is visible; every thread leaves the loop together
    item = s_item[it & 1];
  }
}

}  // namespace

int batch_launch(int dk, const BatchTask *tasks, const long long *prefix, const int *item_task, const int *tabs,
                 unsigned long long *counter, long long total, int grid, hipStream_t s) {
  if (!tasks || !prefix || !item_task || !tabs || !counter || total < 0 || grid <= 0) return -1;
  batch_reset_kernel<<<1, 1, 0, s>>>(counter);
  if (total == 0) return hipGetLastError() == hipSuccess ? 0 : -2;
  switch (dk) {
    case 8:
      batch_contract_kernel<double><<<grid, kThreads, 0, s>>>(tasks, prefix, item_task, tabs, counter, total);
      break;
    case 4:
      batch_contract_kernel<float><<<grid, kThreads, 0, s>>>(tasks, prefix, item_task, tabs, counter, total);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace tal
