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

// One work item = one tm x tn tile of one task's D.  Thread tid always loads
// row tid % tm of the A tile and column tid % tn of the B tile (tm, tn divide
// the 256 threads), so the row/column offsets come from the tables once per
// item and only the contracted-index table is read per element.  The next
// K step is gathered into registers while the current one is multiplied
// from LDS (double-buffered, one barrier per step).
template <typename S>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void batch_contract_kernel(const BatchTask *__restrict__ tasks,
                                                                  const long long *__restrict__ prefix, int ntasks,
                                                                  const int *__restrict__ tabs,
                                                                  unsigned long long *counter, long long total) {
  using MF = BMfma<S>;
  using acc_t = typename MF::acc_t;
  constexpr int NR = kMaxTile * kBK / kThreads;  // registers per operand (tile 64 x 16)
  __shared__ S As[2][kMaxTile][kBK + 1];
  __shared__ S Bs[2][kMaxTile][kBK + 1];
  __shared__ long long s_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  while (true) {
    if (tid == 0) s_item = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long item = s_item;
    __syncthreads();
    if (item >= total) break;
    // task owning this item: last t with prefix[t] <= item
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const BatchTask &T = tasks[lo];
    const long long local = item - prefix[lo];
    const int tm = T.tm, tn = T.tn, K = T.K;
    const int m0 = (int)(local % T.tiles_m) * tm, n0 = (int)(local / T.tiles_m) * tn;
    const int *lM = tabs + T.tab, *lK = lM + T.M, *rK = lK + T.K, *rN = rK + T.K, *dM = rN + T.N, *dN = dM + T.M;
    const S *L = (const S *)T.L, *R = (const S *)T.R;
    // this thread's A row / B column and its k positions within a K step
    const int am = tid % tm, ak = tid / tm, astep = kThreads / tm, na = tm * kBK / kThreads;
    const int bn = tid % tn, bk = tid / tn, bstep = kThreads / tn, nb = tn * kBK / kThreads;
    const bool arow_ok = m0 + am < T.M, bcol_ok = n0 + bn < T.N;
    const int64_t arow = arow_ok ? lM[m0 + am] : 0, bcol = bcol_ok ? rN[n0 + bn] : 0;
    S ra[NR], rb[NR];
    auto gather = [&](int k0) {
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const int ka = k0 + ak + i * astep, kb = k0 + bk + i * bstep;
        ra[i] = (i < na && arow_ok && ka < K) ? L[arow + lK[ka]] : (S)0;
        rb[i] = (i < nb && bcol_ok && kb < K) ? R[rK[kb] + bcol] : (S)0;
      }
    };
    auto stage = [&](int buf) {
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        if (i < na) As[buf][am][ak + i * astep] = ra[i];
        if (i < nb) Bs[buf][bn][bk + i * bstep] = rb[i];
      }
    };
    const int subm = tm >> 4, nsub = subm * (tn >> 4);
    acc_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = acc_t{0, 0, 0, 0};
    gather(0);
    stage(0);
    __syncthreads();
    const int nk = (K + kBK - 1) / kBK;
    for (int kt = 0; kt < nk; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nk) gather((kt + 1) * kBK);
#pragma unroll
      for (int ks = 0; ks < kBK / 4; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int sidx = wave + 4 * q;
          if (sidx < nsub) {
            const int si = sidx % subm, sj = sidx / subm;
            acc[q] = MF::op(As[buf][si * 16 + (lane & 15)][kk], Bs[buf][sj * 16 + (lane & 15)][kk], acc[q]);
          }
        }
      }
      if (kt + 1 < nk) stage(buf ^ 1);
      __syncthreads();
    }
    S *D = (S *)T.D;
    const S alpha = (S)T.alpha;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int sidx = wave + 4 * q;
      if (sidx >= nsub) continue;
      const int si = sidx % subm, sj = sidx / subm;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + si * 16 + MF::row(lane, r), n = n0 + sj * 16 + (lane & 15);
        if (m < T.M && n < T.N) {
          S *d = D + dM[m] + dN[n];
          const S v = alpha * acc[q][r];
          *d = T.beta_one ? *d + v : v;
        }
      }
    }
  }
}

}  // namespace

int batch_launch(int dk, const BatchTask *tasks, const long long *prefix, int ntasks, const int *tabs,
                 unsigned long long *counter, long long total, int grid, hipStream_t s) {
  if (!tasks || !prefix || !tabs || !counter || ntasks <= 0 || total < 0 || grid <= 0) return -1;
  batch_reset_kernel<<<1, 1, 0, s>>>(counter);
  if (total == 0) return hipGetLastError() == hipSuccess ? 0 : -2;
  switch (dk) {
    case 8:
      batch_contract_kernel<double><<<grid, kThreads, 0, s>>>(tasks, prefix, ntasks, tabs, counter, total);
      break;
    case 4:
      batch_contract_kernel<float><<<grid, kThreads, 0, s>>>(tasks, prefix, ntasks, tabs, counter, total);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace tal
