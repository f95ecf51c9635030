// Persistent batched-contraction kernel (gfx950).
//
// Thousands of small contractions of uneven shape run in one launch.  Each
// workgroup pulls work items (workgroup tiles of one contraction,
// self-contained 256-byte descriptors, batch_item.hpp) from its XCD's queue
// with one atomic per item, issued as the previous item starts, and turns to
// the other queues once its own is empty, so the load balances itself
// whatever each item costs.  Operands are read straight from their dlf
// storage: a row, column or contracted index becomes a byte offset by
// mixed-radix digits (multiply-high division), and buffer loads with the
// operand's size as their range return 0 for every index past an edge, so no
// operand is transposed and no load is masked.  Each operand tile of
// kBatchKC contracted indices is loaded along the operand's contiguous index
// (coalesced) into a double-buffered LDS tile that the four waves read as
// MFMA fragments, one barrier per chunk; the loads of the next chunks are in
// flight while the MFMAs of the current one run.  The reference runs such
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

constexpr int kWaves = 4;  // waves per workgroup, 2 x 2 quadrants of the tile
constexpr int kThreads = kWaves * 64;
constexpr int kTile = 64;           // largest tile side (rows or columns)
constexpr int kPad = kBatchKC + 1;  // LDS row length: conflict-free fragment reads

// Double-buffered staged tiles: A (rows x contracted), B (columns x contracted).
template <typename S>
struct TileLds {
  S a[2][kTile][kPad];
  S b[2][kTile][kPad];
};

// One item: tile of 32 BI x 32 BJ; AK / BK: L / R staged along k (else along
// its rows / columns).  `next` is thread 0's pending fetch of the
// workgroup's next queue position (an atomic issued just before this item).
// Once the first chunk has arrived, so has that result, and wave 0 loads one
// word of each 64-byte line of that descriptor (queue position base + next,
// if inside [.., end)) into the scalar cache, so the next item's descriptor
// does not come from memory.
template <typename S, int BI, int BJ, bool AK, bool BK>
__device__ __forceinline__ void item_body(const BatchItem &it, int tid, const BatchItem *__restrict__ items,
                                          unsigned next, int base, int end, TileLds<S> &lds, int *pos_slot) {
  using W = Wave<S>;
  using acc_t = typename W::acc_t;
  constexpr int TA = 32 * BI, TB = 32 * BJ;                             // tile rows, tile columns
  constexpr int NA = TA * kBatchKC / kThreads, NB = TB * kBatchKC / kThreads;  // loads per chunk and thread
  constexpr int PD = W::kDepth;
  const int lane = tid & 63, wave = tid >> 6;
  const auto rl = rsrc(it.L, it.lbytes), rr = rsrc(it.R, it.rbytes), rd = rsrc(it.D, it.dbytes);
  // Along k each load has its own row (offsets fixed per item) and the
  // thread one contracted index; along the rows the thread keeps one row and
  // each load takes the contracted index computed by lane c (< kBatchKC) of
  // the same wave, by a lane shuffle.
  uint32_t aoff[AK ? NA : 1], boff[BK ? NB : 1];
#pragma unroll
  for (int j = 0; j < (AK ? NA : 1); ++j) {
    int r, c;
    stage_coords(AK, TA, kThreads, tid, j, &r, &c);
    aoff[j] = row_offset(it, it.m0 + r, 0);
  }
#pragma unroll
  for (int j = 0; j < (BK ? NB : 1); ++j) {
    int r, c;
    stage_coords(BK, TB, kThreads, tid, j, &r, &c);
    boff[j] = col_offset(it, it.n0 + r, 0);
  }
  S ra[PD][NA], rb[PD][NB];
  auto load = [&](int set, int chunk) {
    uint32_t kl, kr;  // offsets of contracted index chunk * kBatchKC + lane % kBatchKC
    k_offsets(it, chunk * kBatchKC + (lane % kBatchKC), &kl, &kr);
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, c;
      stage_coords(AK, TA, kThreads, tid, j, &r, &c);
      const uint32_t o = AK ? aoff[j] + kl : aoff[0] + (uint32_t)__shfl((int)kl, c, 64);
      ra[set][j] = W::load(rl, o);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int r, c;
      stage_coords(BK, TB, kThreads, tid, j, &r, &c);
      const uint32_t o = BK ? boff[j] + kr : boff[0] + (uint32_t)__shfl((int)kr, c, 64);
      rb[set][j] = W::load(rr, o);
    }
  };
  auto stage = [&](int buf, int set) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, c;
      stage_coords(AK, TA, kThreads, tid, j, &r, &c);
      lds.a[buf][r][c] = ra[set][j];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int r, c;
      stage_coords(BK, TB, kThreads, tid, j, &r, &c);
      lds.b[buf][r][c] = rb[set][j];
    }
  };
  acc_t acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  const int qr = quad_row(BI, wave), qc = quad_col(BJ, wave);
  // a quadrant wholly past the tile's last row or column has nothing to compute
  const bool busy = it.m0 + qr < it.M && it.n0 + qc < it.N;
  auto compute = [&](int buf, int nsteps) {  // the chunk's first nsteps MFMA steps (the rest are past K)
    if (!busy) return;
#pragma unroll
    for (int s = 0; s < kBatchKC / 4; ++s) {
      if (s >= nsteps) break;  // uniform over the wave
      S fa[BI], fb[BJ];
#pragma unroll
      for (int i = 0; i < BI; ++i) fa[i] = lds.a[buf][qr + frag_row(i, lane)][frag_k(s, lane)];
#pragma unroll
      for (int j = 0; j < BJ; ++j) fb[j] = lds.b[buf][qc + frag_row(j, lane)][frag_k(s, lane)];
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
  // Chunk c is staged from register set c % PD into LDS buffer c % 2 (whose
  // previous chunk, c - 2, every wave finished reading before the barrier
  // that ended chunk c - 1), the set then loads chunk c + PD, and after the
  // barrier the waves multiply chunk c.
  for (int c0 = 0; c0 < nch; c0 += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int c = c0 + q;
      if (c < nch) {  // uniform over the workgroup
        stage(c & 1, q);
        if (c + PD < nch) load(q, c + PD);
        if (c == 0 && wave == 0) {  // the next item's descriptor, into the scalar cache
          const int pn = base + __builtin_amdgcn_readfirstlane((int)next);
          const BatchItem &nx = items[pn < end ? pn : base];
          pf0 = nx.M;
          pf1 = (int)nx.gm.e0;
          pf2 = nx.gn.s[0][0];
          pf3 = nx.gk.s[0][0];
        }
        __syncthreads();
        compute(c & 1, (it.K - c * kBatchKC + 3) / 4);
      }
    }
  }
  // Thread 0 publishes the next queue position now, while no store is
  // outstanding: waiting for the fetch later would also wait for the stores.
  if (tid == 0) *pos_slot = (int)next;
  const S alpha = (S)it.alpha;
  const bool accumulate = item_accumulates(it);
  uint32_t dcol[BJ];
#pragma unroll
  for (int j = 0; j < BJ; ++j) dcol[j] = col_offset(it, it.n0 + qc + frag_row(j, lane), 1);
#pragma unroll
  for (int i = 0; i < BI; ++i) {
    uint32_t drow[4];  // D offsets of this block row's accumulator rows
#pragma unroll
    for (int r = 0; r < 4; ++r)
      drow[r] = row_offset(it, it.m0 + qr + 16 * i + acc_row((int)sizeof(S), lane, r), 1);
    S dv[4][BJ];  // an accumulating item reads all of the block row's D first
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int j = 0; j < BJ; ++j) dv[r][j] = accumulate ? W::load(rd, drow[r] + dcol[j]) : (S)0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int j = 0; j < BJ; ++j) W::store(rd, drow[r] + dcol[j], dv[r][j] + alpha * acc[i][j][r]);
  }
  asm volatile("" ::"s"(pf0), "s"(pf1), "s"(pf2), "s"(pf3));
}

// Queue q holds items [qstart[q], qstart[q + 1]); counters[kCounterStride * q]
// counts the positions taken from it.  Workgroup b starts on queue b % nq
// and, once a fetch comes back past the end of a queue, moves on to the next
// one; after a fetch past the end of all nq queues it exits, so every
// workgroup exits and the grid drains.  Thread 0 fetches; the position is
// published through a two-slot LDS word behind a barrier.
constexpr int kCounterStride = 64;  // one 256-byte line per counter

__global__ void batch_reset_kernel(unsigned *counters, int nq) {
  if ((int)threadIdx.x < nq) counters[kCounterStride * threadIdx.x] = 0u;
}

template <typename S>
__global__ __launch_bounds__(kThreads) void batch_contract_kernel(const BatchItem *__restrict__ items,
                                                                 const int *__restrict__ qstart,
                                                                 unsigned *counters, int nq) {
  __shared__ TileLds<S> lds;
  __shared__ int s_pos[2];
  const int tid = threadIdx.x;
  int q = (int)blockIdx.x % nq, visited = 1, slot = 0;
  if (tid == 0) s_pos[0] = (int)atomicAdd(&counters[kCounterStride * q], 1u);
  __syncthreads();
  int cur = s_pos[0];
  while (true) {
    const int base = qstart[q], end = qstart[q + 1];
    slot ^= 1;  // the other slot was last read before the barrier that published `cur`
    if (base + cur >= end) {
      if (visited == nq) break;
      q = q + 1 == nq ? 0 : q + 1;
      ++visited;
      if (tid == 0) s_pos[slot] = (int)atomicAdd(&counters[kCounterStride * q], 1u);
      __syncthreads();
      cur = s_pos[slot];
      continue;
    }
    unsigned next = 0;
    if (tid == 0) next = atomicAdd(&counters[kCounterStride * q], 1u);
    const BatchItem &it = items[base + cur];
    // (BI - 1) | (BJ - 1) << 2 | along-k bits of L and R, uniform over the workgroup
#define TAL_ITEM(BI, BJ, AK, BK)                                         \
  case (BI - 1) | (BJ - 1) << 2 | AK << 5 | BK << 6:                     \
    item_body<S, BI, BJ, AK, BK>(it, tid, items, next, base, end, lds, &s_pos[slot]); \
    break;
#define TAL_ITEMS(AK, BK) TAL_ITEM(1, 1, AK, BK) TAL_ITEM(2, 1, AK, BK) TAL_ITEM(1, 2, AK, BK) TAL_ITEM(2, 2, AK, BK)
    switch (it.shape & 0x6f) {
      TAL_ITEMS(0, 0)
      TAL_ITEMS(1, 0)
      TAL_ITEMS(0, 1)
      TAL_ITEMS(1, 1)
      default:  // not produced by the planner: skipped
        if (tid == 0) s_pos[slot] = (int)next;
        break;
    }
#undef TAL_ITEMS
#undef TAL_ITEM
    // every wave has finished the item's LDS reads before the next item stages
    // into LDS, and the next position (published by the item) is visible
    __syncthreads();
    cur = s_pos[slot];
  }
}

}  // namespace

int batch_blocks_per_cu(int dk, int *blocks) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (dk == 8)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_contract_kernel<double>, kThreads, 0);
  else if (dk == 4)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_contract_kernel<float>, kThreads, 0);
  if (e != hipSuccess || nb <= 0) return -1;
  *blocks = nb;
  return 0;
}

int batch_launch(int dk, const BatchItem *items, const int *qstart, unsigned *counters, int nq, int nblocks,
                 hipStream_t s) {
  if (!items || !qstart || !counters || nq <= 0 || nq > kThreads || nblocks <= 0) return -1;
  batch_reset_kernel<<<1, kThreads, 0, s>>>(counters, nq);
  switch (dk) {
    case 8: batch_contract_kernel<double><<<nblocks, kThreads, 0, s>>>(items, qstart, counters, nq); break;
    case 4: batch_contract_kernel<float><<<nblocks, kThreads, 0, s>>>(items, qstart, counters, nq); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t batch_counter_bytes(int nq) { return (size_t)nq * kCounterStride * sizeof(unsigned); }

}  // namespace tal
