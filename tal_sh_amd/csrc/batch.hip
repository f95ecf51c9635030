// Persistent batched-contraction kernel (gfx950).
//
// Thousands of small contractions of uneven shape run in one launch.  Each
// wave is an independent worker: the host planner gives it a fixed list of
// work items (wave tiles of one contraction, self-contained 256-byte
// descriptors, batch_item.hpp) and the wave computes them one after the
// other with no barrier, no LDS and no atomic.  Operands are read straight
// from their dlf storage into the 16x16x4 MFMA operand registers: a row,
// column or contracted index becomes a byte offset by mixed-radix digits
// (multiply-high division), and buffer loads with the operand's size as
// their range return 0 for every index past an edge, so no operand is
// transposed and no load is masked.  Contracted indices are loaded kDepth
// steps of 4 ahead of the MFMAs that consume them.  The reference runs such
// tasks one TTGT schedule (up to 4 transposes + a GEMM launch) at a time
// (tensor_algebra_gpu_nvidia.cu:5847-7188).
//
// Each D element is produced by exactly one item, in a fixed order of the
// contracted indices, so results are deterministic and independent of the
// batch composition and of the schedule.
#include "batch.hpp"

#include <hip/hip_runtime.h>

namespace tal {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Raw buffer resource (gfx9 format word) over `bytes` bytes at p: the range
// check drops every access at or past `bytes`.
__device__ inline __amdgpu_buffer_rsrc_t rsrc(uint64_t p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

template <typename S>
struct Wave;
template <>
struct Wave<double> {
  using acc_t = f64x4;
  static constexpr int kDepth = kBatchDepthF64;
  __device__ static double load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  }
  __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 0);
  }
  __device__ static acc_t mfma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Wave<float> {
  using acc_t = f32x4;
  static constexpr int kDepth = kBatchDepthF32;
  __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  }
  __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, 0, 0);
  }
  __device__ static acc_t mfma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int kWaves = 4;  // waves per workgroup (independent workers)

// `nx` is the wave's next descriptor (this one for its last item): one
// scalar load per 64-byte line of it is issued once this item's own
// descriptor fields are in registers, and consumed after the item, so the
// next item starts from the scalar cache instead of memory.
template <typename S, int BI, int BJ>
__device__ __forceinline__ void item_body(const BatchItem &it, int lane, const BatchItem &nx) {
  using W = Wave<S>;
  using acc_t = typename W::acc_t;
  constexpr int P = W::kDepth;
  const auto rl = rsrc(it.L, it.lbytes), rr = rsrc(it.R, it.rbytes), rd = rsrc(it.D, it.dbytes);
  uint32_t arow[BI], bcol[BJ], dcol[BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i) arow[i] = a_row_offset(it, i, lane);
#pragma unroll
  for (int j = 0; j < BJ; ++j) b_col_offsets(it, j, lane, &bcol[j], &dcol[j]);
  acc_t acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  S a[P][BI], b[P][BJ];
  // loads of K step c into register set s; a step past K loads zeros (kOutside)
  auto load = [&](int s, int c) {
    uint32_t kl, kr;
    k_offsets(it, c, lane, &kl, &kr);
#pragma unroll
    for (int i = 0; i < BI; ++i) a[s][i] = W::load(rl, arow[i] + kl);
#pragma unroll
    for (int j = 0; j < BJ; ++j) b[s][j] = W::load(rr, kr + bcol[j]);
  };
  const int nsteps = (it.K + 3) / 4;
#pragma unroll
  for (int s = 0; s < P; ++s) load(s, s);
  const int pf0 = nx.M, pf1 = (int)nx.gm.e0, pf2 = nx.gn.s[0][0], pf3 = nx.gk.s[0][0];
  for (int c0 = 0; c0 < nsteps; c0 += P) {
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (c0 + s < nsteps) {  // uniform over the wave
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < BJ; ++j) acc[i][j] = W::mfma(a[s][i], b[s][j], acc[i][j]);
      }
      load(s, c0 + s + P);
    }
  }
  const S alpha = (S)it.alpha;
  const bool accumulate = item_accumulates(it);
  uint32_t drow[BI][4];  // D offsets of the accumulator rows, computed before the stores
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) drow[i][r] = d_row_offset(it, i, acc_row((int)sizeof(S), lane, r));
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        const uint32_t off = drow[i][r] + dcol[j];
        S v = alpha * acc[i][j][r];
        if (accumulate) v = W::load(rd, off) + v;
        W::store(rd, off, v);
      }
    }
  asm volatile("" ::"s"(pf0), "s"(pf1), "s"(pf2), "s"(pf3));
}

// Workgroup b's wave v is worker w = b * kWaves + v and runs items
// [start[w], start[w + 1]) of `items`; every worker's list is finite, so the
// grid drains.
template <typename S>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4))) void batch_contract_kernel(const BatchItem *__restrict__ items,
                                                                    const int *__restrict__ start, int nworkers) {
  const int lane = threadIdx.x & 63;
  const int w = (int)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  if (w >= nworkers) return;
  const int p1 = start[w + 1];
  for (int p = start[w]; p < p1; ++p) {
    const BatchItem &it = items[p], &nx = items[p + 1 < p1 ? p + 1 : p];
    switch (it.shape & 15) {  // (BI - 1) | (BJ - 1) << 2, uniform over the wave
      case 0: item_body<S, 1, 1>(it, lane, nx); break;
      case 1: item_body<S, 2, 1>(it, lane, nx); break;
      case 4: item_body<S, 1, 2>(it, lane, nx); break;
      case 5: item_body<S, 2, 2>(it, lane, nx); break;
      case 3: item_body<S, 4, 1>(it, lane, nx); break;
      case 12: item_body<S, 1, 4>(it, lane, nx); break;
      default: break;  // not produced by the planner
    }
  }
}

}  // namespace

int batch_workers_per_block() { return kWaves; }

int batch_blocks_per_cu(int dk, int *blocks) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (dk == 8)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_contract_kernel<double>, kWaves * 64, 0);
  else if (dk == 4)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_contract_kernel<float>, kWaves * 64, 0);
  if (e != hipSuccess || nb <= 0) return -1;
  *blocks = nb;
  return 0;
}

int batch_launch(int dk, const BatchItem *items, const int *start, int nblocks, hipStream_t s) {
  if (!items || !start || nblocks <= 0) return -1;
  switch (dk) {
    case 8: batch_contract_kernel<double><<<nblocks, kWaves * 64, 0, s>>>(items, start, nblocks * kWaves); break;
    case 4: batch_contract_kernel<float><<<nblocks, kWaves * 64, 0, s>>>(items, start, nblocks * kWaves); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace tal
