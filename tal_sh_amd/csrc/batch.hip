// Persistent batched-contraction kernel (gfx950).
//
// Thousands of small contractions of uneven shape run in one launch.  Each
// wave is an independent worker (no barrier): it pulls work items (wave tiles
// of one contraction, self-contained 256-byte descriptors, batch_item.hpp)
// from its XCD's queue with one atomic per item, issued as the previous item
// starts, and turns to the other queues once its own is empty, so the load
// balances itself whatever each item costs.  Operands are read straight from their
// dlf storage: a row, column or contracted index becomes a byte offset by
// mixed-radix digits (multiply-high division), and buffer loads with the
// operand's size as their range return 0 for every index past an edge, so no
// operand is transposed and no load is masked.  Each operand tile of
// kBatchKC contracted indices is loaded along the operand's contiguous index
// (coalesced), written to the wave's own LDS region and read back as MFMA
// fragments; the loads of the next chunk(s) are in flight while the MFMAs of
// the current one run.  The reference runs such tasks one TTGT schedule (up
// to 4 transposes + a GEMM launch) at a time (tensor_algebra_gpu_nvidia.cu:
// 5847-7188).
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
constexpr int kTile = 32;  // largest tile side (rows or columns)
constexpr int kPad = kBatchKC + 1;  // LDS row length: conflict-free fragment reads

// The wave's LDS region: the staged A (rows x contracted) and B (columns x
// contracted) tiles of the current chunk.
template <typename S>
struct WaveLds {
  S a[kTile][kPad];
  S b[kTile][kPad];
};

// One item: tile of BI x BJ blocks; AK / BK: L / R staged along k (else along
// its rows / columns).  `next` is lane 0's pending fetch of the wave's next
// queue position (an atomic issued just before this item); once the first
// chunk has arrived, so has that result: one scalar load per 64-byte line of
// the descriptor at queue position base + next (if inside [.., end)) is then
// issued and consumed after the item, so the next item starts from the
// scalar cache instead of memory.
template <typename S, int BI, int BJ, bool AK, bool BK>
__device__ __forceinline__ void item_body(const BatchItem &it, int lane, const BatchItem *__restrict__ items,
                                          unsigned next, int base, int end, WaveLds<S> &lds) {
  using W = Wave<S>;
  using acc_t = typename W::acc_t;
  constexpr int TA = 16 * BI, TB = 16 * BJ;                       // tile rows, tile columns
  constexpr int NA = TA * kBatchKC / 64, NB = TB * kBatchKC / 64;  // loads per chunk and lane
  constexpr int PD = W::kDepth;
  const auto rl = rsrc(it.L, it.lbytes), rr = rsrc(it.R, it.rbytes), rd = rsrc(it.D, it.dbytes);
  // Along k each load has its own row (offsets fixed per item) and the lane
  // one contracted index; along the rows the lane keeps one row and each load
  // takes the contracted index of lane c (< kBatchKC) by a lane shuffle.
  uint32_t aoff[AK ? NA : 1], boff[BK ? NB : 1];
#pragma unroll
  for (int j = 0; j < (AK ? NA : 1); ++j) {
    int r, c;
    stage_coords(AK, TA, lane, j, &r, &c);
    aoff[j] = row_offset(it, it.m0 + r, 0);
  }
#pragma unroll
  for (int j = 0; j < (BK ? NB : 1); ++j) {
    int r, c;
    stage_coords(BK, TB, lane, j, &r, &c);
    boff[j] = col_offset(it, it.n0 + r, 0);
  }
  S ra[PD][NA], rb[PD][NB];
  auto load = [&](int set, int chunk) {
    uint32_t kl, kr;  // offsets of contracted index chunk * kBatchKC + lane % kBatchKC
    k_offsets(it, chunk * kBatchKC + (lane % kBatchKC), &kl, &kr);
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, c;
      stage_coords(AK, TA, lane, j, &r, &c);
      const uint32_t o = AK ? aoff[j] + kl : aoff[0] + (uint32_t)__shfl((int)kl, c, 64);
      ra[set][j] = W::load(rl, o);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int r, c;
      stage_coords(BK, TB, lane, j, &r, &c);
      const uint32_t o = BK ? boff[j] + kr : boff[0] + (uint32_t)__shfl((int)kr, c, 64);
      rb[set][j] = W::load(rr, o);
    }
  };
  auto stage = [&](int set) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, c;
      stage_coords(AK, TA, lane, j, &r, &c);
      lds.a[r][c] = ra[set][j];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int r, c;
      stage_coords(BK, TB, lane, j, &r, &c);
      lds.b[r][c] = rb[set][j];
    }
  };
  acc_t acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  auto compute = [&]() {
#pragma unroll
    for (int s = 0; s < kBatchKC / 4; ++s) {
      S fa[BI], fb[BJ];
#pragma unroll
      for (int i = 0; i < BI; ++i) fa[i] = lds.a[frag_row(i, lane)][frag_k(s, lane)];
#pragma unroll
      for (int j = 0; j < BJ; ++j) fb[j] = lds.b[frag_row(j, lane)][frag_k(s, lane)];
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) acc[i][j] = W::mfma(fa[i], fb[j], acc[i][j]);
    }
  };
  const int nch = (it.K + kBatchKC - 1) / kBatchKC;
#pragma unroll
  for (int q = 0; q < PD; ++q)
    if (q < nch) load(q, q);
  int pf0 = 0, pf1 = 0, pf2 = 0, pf3 = 0;
  // chunk c is staged from register set c % PD, whose loads then move on to
  // chunk c + PD; LDS accesses of one wave complete in order, so staging chunk
  // c cannot overtake the fragment reads of chunk c - 1
  for (int c0 = 0; c0 < nch; c0 += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      if (c0 + q < nch) {  // uniform over the wave
        stage(q);
        if (c0 + q + PD < nch) load(q, c0 + q + PD);
        if (c0 + q == 0) {  // the next item's descriptor, into the scalar cache
          const int pn = base + __builtin_amdgcn_readfirstlane((int)next);
          const BatchItem &nx = items[pn < end ? pn : base];
          pf0 = nx.M;
          pf1 = (int)nx.gm.e0;
          pf2 = nx.gn.s[0][0];
          pf3 = nx.gk.s[0][0];
        }
        compute();
      }
    }
  }
  const S alpha = (S)it.alpha;
  const bool accumulate = item_accumulates(it);
  uint32_t dcol[BJ];
#pragma unroll
  for (int j = 0; j < BJ; ++j) dcol[j] = col_offset(it, it.n0 + frag_row(j, lane), 1);
#pragma unroll
  for (int i = 0; i < BI; ++i) {
    uint32_t drow[4];  // D offsets of this block row's accumulator rows
#pragma unroll
    for (int r = 0; r < 4; ++r) drow[r] = row_offset(it, it.m0 + 16 * i + acc_row((int)sizeof(S), lane, r), 1);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        const uint32_t off = drow[r] + dcol[j];
        S v = alpha * acc[i][j][r];
        if (accumulate) v = W::load(rd, off) + v;
        W::store(rd, off, v);
      }
    }
  }
  asm volatile("" ::"s"(pf0), "s"(pf1), "s"(pf2), "s"(pf3));
}

// Queue q holds items [qstart[q], qstart[q + 1]); counters[kCounterStride * q]
// counts the positions taken from it.  Workgroup b starts on queue b % nq
// and, once a fetch comes back past the end of a queue, moves on to the next
// one; after a fetch past the end of all nq queues the wave exits, so every
// wave exits and the grid drains.
constexpr int kCounterStride = 64;  // one 256-byte line per counter

__global__ void batch_reset_kernel(unsigned *counters, int nq) {
  if ((int)threadIdx.x < nq) counters[kCounterStride * threadIdx.x] = 0u;
}

template <typename S>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4))) void batch_contract_kernel(
    const BatchItem *__restrict__ items, const int *__restrict__ qstart, unsigned *counters, int nq) {
  __shared__ WaveLds<S> lds[kWaves];
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  int q = (int)blockIdx.x % nq, visited = 1;
  unsigned got = 0;
  if (lane == 0) got = atomicAdd(&counters[kCounterStride * q], 1u);
  int cur = __builtin_amdgcn_readfirstlane((int)got);
  while (true) {
    const int base = qstart[q], end = qstart[q + 1];
    if (base + cur >= end) {
      if (visited == nq) break;
      q = q + 1 == nq ? 0 : q + 1;
      ++visited;
      if (lane == 0) got = atomicAdd(&counters[kCounterStride * q], 1u);
      cur = __builtin_amdgcn_readfirstlane((int)got);
      continue;
    }
    unsigned next = 0;
    if (lane == 0) next = atomicAdd(&counters[kCounterStride * q], 1u);
    const BatchItem &it = items[base + cur];
    // (BI - 1) | (BJ - 1) << 2 | along-k bits of L and R, uniform over the wave
#define TAL_ITEM(BI, BJ, AK, BK)                                        \
  case (BI - 1) | (BJ - 1) << 2 | AK << 5 | BK << 6:                    \
    item_body<S, BI, BJ, AK, BK>(it, lane, items, next, base, end, lds[v]); \
    break;
#define TAL_ITEMS(AK, BK) TAL_ITEM(1, 1, AK, BK) TAL_ITEM(2, 1, AK, BK) TAL_ITEM(1, 2, AK, BK) TAL_ITEM(2, 2, AK, BK)
    switch (it.shape & 0x6f) {
      TAL_ITEMS(0, 0)
      TAL_ITEMS(1, 0)
      TAL_ITEMS(0, 1)
      TAL_ITEMS(1, 1)
      default: break;  // not produced by the planner
    }
#undef TAL_ITEMS
#undef TAL_ITEM
    cur = __builtin_amdgcn_readfirstlane((int)next);
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

int batch_launch(int dk, const BatchItem *items, const int *qstart, unsigned *counters, int nq, int nblocks,
                 hipStream_t s) {
  if (!items || !qstart || !counters || nq <= 0 || nq > kWaves * 64 || nblocks <= 0) return -1;
  batch_reset_kernel<<<1, kWaves * 64, 0, s>>>(counters, nq);
  switch (dk) {
    case 8: batch_contract_kernel<double><<<nblocks, kWaves * 64, 0, s>>>(items, qstart, counters, nq); break;
    case 4: batch_contract_kernel<float><<<nblocks, kWaves * 64, 0, s>>>(items, qstart, counters, nq); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t batch_counter_bytes(int nq) { return (size_t)nq * kCounterStride * sizeof(unsigned); }

}  // namespace tal
