// Tensor permutation (copy / permuted add), init, scale and squared 2-norm
// kernels for gfx950.
//
// Replaces the reference's transpose kernel gpu_tensor_block_copy_dlf__
// (tensor_algebra_gpu_nvidia.cu:1125-1457), its never-launched permuted-add
// twin gpu_tensor_block_add_dlf__ (:785-1122), gpu_array_add__ (:442-480),
// gpu_array_init__ (:337-345), gpu_array_scale__ (:410-438) and the
// never-launched gpu_array_norm2__ (:240-334).
//
// Permutation design: the host fuses input dimensions that stay adjacent in
// the output and drops unit extents, so every permutation reduces to a tiled
// 2-D problem over (a) the input's fastest dimension and (b) the dimension
// that becomes the output's fastest one, with all remaining dimensions
// enumerated by the block index.  If (a) is also the output's fastest
// dimension the copy streams rows directly (both sides coalesced); otherwise
// a T x T tile is staged through LDS with a +1 column pad so both the read
// and the write side are full contiguous runs.  The elementwise add uses a
// separately rounded multiply and add, matching the reference CPU path bit
// for bit; copies are pure data movement.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

// Elementwise arithmetic must round every multiply and add separately, as the
// reference CPU path does: no fused multiply-add contraction in this file.
#pragma clang fp contract(off)

namespace tal {

namespace {

constexpr int kMaxRank = 56;
constexpr int kThreads = 256;

// ---------------------------------------------------------------- element ops
template <typename T>
struct Elem;

template <>
struct Elem<float> {
  __device__ static float conj(float x, int) { return x; }
  __device__ static float axpy(float d, float l, float ar, float) { return __fadd_rn(d, __fmul_rn(l, ar)); }
  __device__ static float scale(float d, float ar, float) { return __fmul_rn(d, ar); }
  __device__ static double sq(float x) { return (double)__fmul_rn(x, x); }
};
template <>
struct Elem<double> {
  __device__ static double conj(double x, int) { return x; }
  __device__ static double axpy(double d, double l, double ar, double) { return __dadd_rn(d, __dmul_rn(l, ar)); }
  __device__ static double scale(double d, double ar, double) { return __dmul_rn(d, ar); }
  __device__ static double sq(double x) { return x * x; }
};
template <>
struct Elem<float2> {
  __device__ static float2 conj(float2 x, int c) { return c ? make_float2(x.x, -x.y) : x; }
  __device__ static float2 mul(float2 x, float cr, float ci) {
    float t1 = __fmul_rn(x.x, cr), t2 = __fmul_rn(x.y, ci), t3 = __fmul_rn(x.x, ci), t4 = __fmul_rn(x.y, cr);
    return make_float2(__fsub_rn(t1, t2), __fadd_rn(t3, t4));
  }
  __device__ static float2 axpy(float2 d, float2 l, float cr, float ci) {
    float2 p = mul(l, cr, ci);
    return make_float2(__fadd_rn(d.x, p.x), __fadd_rn(d.y, p.y));
  }
  __device__ static float2 scale(float2 d, float cr, float ci) { return mul(d, cr, ci); }
  __device__ static double sq(float2 x) { return (double)__fmul_rn(x.x, x.x) + (double)__fmul_rn(x.y, x.y); }
};
template <>
struct Elem<double2> {
  __device__ static double2 conj(double2 x, int c) { return c ? make_double2(x.x, -x.y) : x; }
  __device__ static double2 mul(double2 x, double cr, double ci) {
    double t1 = __dmul_rn(x.x, cr), t2 = __dmul_rn(x.y, ci), t3 = __dmul_rn(x.x, ci), t4 = __dmul_rn(x.y, cr);
    return make_double2(__dsub_rn(t1, t2), __dadd_rn(t3, t4));
  }
  __device__ static double2 axpy(double2 d, double2 l, double cr, double ci) {
    double2 p = mul(l, cr, ci);
    return make_double2(__dadd_rn(d.x, p.x), __dadd_rn(d.y, p.y));
  }
  __device__ static double2 scale(double2 d, double cr, double ci) { return mul(d, cr, ci); }
  __device__ static double sq(double2 x) { return x.x * x.x + x.y * x.y; }
};

template <typename T>
struct Real {
  using type = T;
};
template <>
struct Real<float2> {
  using type = float;
};
template <>
struct Real<double2> {
  using type = double;
};

// ------------------------------------------------------------ permute args
// After fusion a permutation is a 3-D tile problem over dims a (input
// fastest, stride 1), b and c, with every other dim enumerated by the block
// index.  Tile mode: b is the output's fastest dim (output stride 1); row
// mode: a is fastest on both sides and b is the dim with the next smallest
// output stride.  c is an outer dim carried through the tile so that one
// block moves kElems elements with all of its loads issued before any store.
// NO: capacity for outer dims (a launch of two permutations packs both
// argument sets into the kernel arguments, so it uses a short form).
template <int NO>
struct PermArgsT {
  int nouter;
  int64_t ea, eb, ec;          // tile-dim extents
  int64_t isb, isc;            // input strides of b and c (a has stride 1)
  int64_t osa, osb, osc;       // output strides of a, b, c
  int64_t ntile_a, ntile_b, ntile_c;
  int tc;                      // tile size along c (<= TCMAX of the kernel)
  int tb_log;                  // row mode: log2 of the tile size along b
  int64_t oe[NO];              // outer extents
  int64_t ois[NO];             // outer input strides
  int64_t oos[NO];             // outer output strides
};
using PermArgs = PermArgsT<kMaxRank>;
constexpr int kPairOuter = 6;

template <class PA>
__device__ inline void tile_origin(const PA &p, int64_t blk, int TA, int TB, int64_t &a0, int64_t &b0,
                                   int64_t &c0, int64_t &ib, int64_t &ob) {
  int64_t t = blk % p.ntile_a;
  blk /= p.ntile_a;
  a0 = t * TA;
  t = blk % p.ntile_b;
  blk /= p.ntile_b;
  b0 = t * TB;
  t = blk % p.ntile_c;
  blk /= p.ntile_c;
  c0 = t * p.tc;
  ib = a0 + b0 * p.isb + c0 * p.isc;
  ob = a0 * p.osa + b0 * p.osb + c0 * p.osc;
  for (int k = 0; k < p.nouter; ++k) {
    const int64_t e = p.oe[k];
    const int64_t q = blk / e;
    const int64_t r = blk - q * e;
    blk = q;
    ib += r * p.ois[k];
    ob += r * p.oos[k];
  }
}

// Elements moved by one block.  Small blocks win on MI355X: with 128-byte
// runs (extent-16 R8 dims) 1024-element blocks beat 4096-element ones by up
// to 25 % on the slowest permutations of the 16^6 sweep (more blocks resident
// per CU, shorter tail), and a flat copy peaks with one 16-byte vector per
// thread.  4-byte elements move 2048 per block (8 KB, as for 8-byte ones):
// 5.6 % faster on the 96^4 R4 operand transposes than 1024.
template <typename T>
constexpr int elems_per_block() {
  return sizeof(T) == 4 ? 2048 : 1024;
}

// V consecutive elements moved as one 8- or 16-byte access.
template <typename T, int V>
struct Pack {
  T v[V];
};

template <typename T, int V>
__device__ inline Pack<T, V> ldv(const T *p) {
  Pack<T, V> r;
  if constexpr (sizeof(T) * V == 16) {
    const uint4 u = *reinterpret_cast<const uint4 *>(p);
    __builtin_memcpy(&r, &u, 16);
  } else if constexpr (sizeof(T) * V == 8) {
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    __builtin_memcpy(&r, &u, 8);
  } else {
#pragma unroll
    for (int t = 0; t < V; ++t) r.v[t] = p[t];
  }
  return r;
}

template <typename T, int V>
__device__ inline void stv(T *p, const Pack<T, V> &r) {
  if constexpr (sizeof(T) * V == 16) {
    uint4 u;
    __builtin_memcpy(&u, &r, 16);
    *reinterpret_cast<uint4 *>(p) = u;
  } else if constexpr (sizeof(T) * V == 8) {
    uint2 u;
    __builtin_memcpy(&u, &r, 8);
    *reinterpret_cast<uint2 *>(p) = u;
  } else {
#pragma unroll
    for (int t = 0; t < V; ++t) p[t] = r.v[t];
  }
}

template <typename T, int V, int OP>
__device__ inline void emit(T *o, Pack<T, V> x, typename Real<T>::type ar, typename Real<T>::type ai, int conj) {
#pragma unroll
  for (int t = 0; t < V; ++t) x.v[t] = Elem<T>::conj(x.v[t], conj);
  if (OP == PERM_ADD) {
    const Pack<T, V> d = ldv<T, V>(o);
#pragma unroll
    for (int t = 0; t < V; ++t) x.v[t] = Elem<T>::axpy(d.v[t], x.v[t], ar, ai);
  }
  stv<T, V>(o, x);
}

// Row mode: dim a is contiguous on both sides; no staging needed.  Accesses
// are V elements wide along a (the host guarantees V | extent of a and
// 16-byte aligned bases when V > 1).  The tile size along b is a power of two
// chosen at launch (p.tb_log) to fit the extent of b, so a small b does not
// leave most of the block idle; c takes the rest of the block.
template <typename T, int TA, int V, int OP, class PA>
__device__ inline void perm_rows_body(const T *__restrict__ src, T *__restrict__ dst, const PA &p, int64_t blk,
                                      typename Real<T>::type ar, typename Real<T>::type ai, int conj) {
  constexpr int EPB = elems_per_block<T>(), N = EPB / kThreads / V, AV = TA / V;
  const int tbl = p.tb_log, tb = 1 << tbl;
  int64_t a0, b0, c0, ib, ob;
  tile_origin(p, blk, TA, tb, a0, b0, c0, ib, ob);
  const int na = (int)min<int64_t>(TA, p.ea - a0), nb = (int)min<int64_t>(tb, p.eb - b0);
  const int nc = (int)min<int64_t>(p.tc, p.ec - c0);
  Pack<T, V> v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int u = threadIdx.x + q * kThreads;
    const int ia = (u % AV) * V, jb = (u / AV) & (tb - 1), kc = (u / AV) >> tbl;
    if (ia < na && jb < nb && kc < nc) v[q] = ldv<T, V>(src + ib + ia + jb * p.isb + kc * p.isc);
  }
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int u = threadIdx.x + q * kThreads;
    const int ia = (u % AV) * V, jb = (u / AV) & (tb - 1), kc = (u / AV) >> tbl;
    if (ia < na && jb < nb && kc < nc) emit<T, V, OP>(dst + ob + ia + jb * p.osb + kc * p.osc, v[q], ar, ai, conj);
  }
}
template <typename T, int TA, int V, int OP>
__global__ __launch_bounds__(kThreads) void perm_rows_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             PermArgs p, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  perm_rows_body<T, TA, V, OP>(src, dst, p, blockIdx.x, ar, ai, conj);
}

// Tile mode: stage a (a x b x c) tile through LDS (+1 pad on the a axis) so
// the reads run along a and the writes along b, V elements per access on
// both sides (the host guarantees V | extents of a and b when V > 1).
template <typename T, int TA, int TB, int V, int OP, class PA>
__device__ inline void perm_tile_body(const T *__restrict__ src, T *__restrict__ dst, const PA &p, int64_t blk,
                                      typename Real<T>::type ar, typename Real<T>::type ai, int conj) {
  constexpr int EPB = elems_per_block<T>(), N = EPB / kThreads / V, TCMAX = EPB / (TA * TB);
  constexpr int AV = TA / V, BV = TB / V;
  __shared__ T tile[TCMAX][TB][TA + 1];
  int64_t a0, b0, c0, ib, ob;
  tile_origin(p, blk, TA, TB, a0, b0, c0, ib, ob);
  const int na = (int)min<int64_t>(TA, p.ea - a0), nb = (int)min<int64_t>(TB, p.eb - b0);
  const int nc = (int)min<int64_t>(p.tc, p.ec - c0);
  Pack<T, V> v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int u = threadIdx.x + q * kThreads;
    const int ia = (u % AV) * V, jb = (u / AV) % TB, kc = u / (AV * TB);
    if (ia < na && jb < nb && kc < nc) v[q] = ldv<T, V>(src + ib + ia + jb * p.isb + kc * p.isc);
  }
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int u = threadIdx.x + q * kThreads;
    const int ia = (u % AV) * V, jb = (u / AV) % TB, kc = u / (AV * TB);
    if (ia < na && jb < nb && kc < nc)
#pragma unroll
      for (int t = 0; t < V; ++t) tile[kc][jb][ia + t] = v[q].v[t];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int u = threadIdx.x + q * kThreads;
    const int jb = (u % BV) * V, ia = (u / BV) % TA, kc = u / (BV * TA);
    if (ia < na && jb < nb && kc < nc) {
      Pack<T, V> x;
#pragma unroll
      for (int t = 0; t < V; ++t) x.v[t] = tile[kc][jb + t][ia];
      emit<T, V, OP>(dst + ob + jb + ia * p.osa + kc * p.osc, x, ar, ai, conj);
    }
  }
}
template <typename T, int TA, int TB, int V, int OP>
__global__ __launch_bounds__(kThreads) void perm_tile_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             PermArgs p, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  perm_tile_body<T, TA, TB, V, OP>(src, dst, p, blockIdx.x, ar, ai, conj);
}

// Two permutation copies of one kind and tile shape in one launch (the A and
// B operand transposes of a contraction): blocks [0, nblk0) move the first,
// the rest the second.  One launch instead of two on a short contraction's
// critical path.
template <typename T>
struct PermPair {
  const T *src[2];
  T *dst[2];
  PermArgsT<kPairOuter> p[2];
  int64_t nblk0;
};
template <typename T, int TA, int TB, int V, bool ROWS>
__global__ __launch_bounds__(kThreads) void perm_pair_kernel(PermPair<T> q) {
  using R = typename Real<T>::type;
  const int64_t b = blockIdx.x;
  if constexpr (ROWS) {
    if (b < q.nblk0) perm_rows_body<T, TA, V, PERM_COPY>(q.src[0], q.dst[0], q.p[0], b, (R)1, (R)0, 0);
    else perm_rows_body<T, TA, V, PERM_COPY>(q.src[1], q.dst[1], q.p[1], b - q.nblk0, (R)1, (R)0, 0);
  } else {
    if (b < q.nblk0) perm_tile_body<T, TA, TB, V, PERM_COPY>(q.src[0], q.dst[0], q.p[0], b, (R)1, (R)0, 0);
    else perm_tile_body<T, TA, TB, V, PERM_COPY>(q.src[1], q.dst[1], q.p[1], b - q.nblk0, (R)1, (R)0, 0);
  }
}

// Contiguous (identity after fusion) copy / add: V-wide accesses, 4 in
// flight per thread; the tail (n % (V*4*threads)) is handled element-wise.
template <typename T, int V, int OP>
__global__ __launch_bounds__(kThreads) void perm_flat_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             int64_t n, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  constexpr int U = 1;
  const int64_t nvec = n / V;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (U - 1) * stride < nvec; i += U * stride) {
    Pack<T, V> v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = ldv<T, V>(src + (i + k * stride) * V);
#pragma unroll
    for (int k = 0; k < U; ++k) emit<T, V, OP>(dst + (i + k * stride) * V, v[k], ar, ai, conj);
  }
  for (; i < nvec; i += stride) emit<T, V, OP>(dst + i * V, ldv<T, V>(src + i * V), ar, ai, conj);
  for (int64_t e = nvec * V + (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += stride)
    emit<T, 1, OP>(dst + e, ldv<T, 1>(src + e), ar, ai, conj);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void init_kernel(T *__restrict__ d, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (; i < n; i += stride) d[i] = v;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scale_kernel(T *__restrict__ d, int64_t n, typename Real<T>::type ar,
                                                         typename Real<T>::type ai) {
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (; i < n; i += stride) d[i] = Elem<T>::scale(d[i], ar, ai);
}

// ------------------------------------------------------------------ norm2
// Stage-1 blocks (2 per CU): 512 read a 134 MB tensor in 23.6 us against
// 24.9 us with 1,024, 25.0 with 2,048 and 25.4 with 8 loads in flight.
constexpr int kNormBlocks = 512;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline double block_sum(double v) {
  __shared__ double ws[kThreads / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) t += ws[i];
  return t;
}

// Stage 1: grid-stride over V-wide packs with four independent accumulators
// per thread; the grid size depends only on n, so the summation order (and
// the result) is reproducible run to run.
template <typename T, int V>
__global__ __launch_bounds__(kThreads) void norm2_stage1(const T *__restrict__ x, int64_t n,
                                                         double *__restrict__ partials) {
  const int64_t nvec = n / V, stride = (int64_t)gridDim.x * kThreads;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + 3 * stride < nvec; i += 4 * stride) {
    Pack<T, V> v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ldv<T, V>(x + (i + k * stride) * V);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int t = 0; t < V; ++t) acc[k] += Elem<T>::sq(v[k].v[t]);
  }
  for (; i < nvec; i += stride) {
    const Pack<T, V> v = ldv<T, V>(x + i * V);
#pragma unroll
    for (int t = 0; t < V; ++t) acc[0] += Elem<T>::sq(v.v[t]);
  }
  if (blockIdx.x == 0)
    for (int64_t e = nvec * V + threadIdx.x; e < n; e += kThreads) acc[0] += Elem<T>::sq(x[e]);
  const double t = block_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Stage 2: one block sums the partials in a fixed order and stores the result
// into the rank-0 output body with the output's data kind.
template <typename T>
__global__ __launch_bounds__(kThreads) void norm2_stage2(const double *__restrict__ partials, int np, T *out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kThreads) acc += partials[i];
  double t = block_sum(acc);
  if (threadIdx.x == 0) {
    T v;
    if constexpr (std::is_floating_point<T>::value) {
      v = (T)t;
    } else {
      v.x = (typename Real<T>::type)t;
      v.y = 0;
    }
    *out = v;
  }
}

int grid_for(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 22));
}

template <typename T, int TA, int V>
void launch_rows(const T *src, T *dst, const PermArgs &p, int op, typename Real<T>::type ar,
                 typename Real<T>::type ai, int conj, int64_t nblk, hipStream_t s) {
  if (op == PERM_COPY)
    launch_clocked(perm_rows_kernel<T, TA, V, PERM_COPY>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p, ar, ai,
                                                                                          conj);
  else
    launch_clocked(perm_rows_kernel<T, TA, V, PERM_ADD>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p, ar, ai,
                                                                                         conj);
}

template <typename T, int TA, int TB, int V>
void launch_tile(const T *src, T *dst, const PermArgs &p, int op, typename Real<T>::type ar,
                 typename Real<T>::type ai, int conj, int64_t nblk, hipStream_t s) {
  if (op == PERM_COPY)
    launch_clocked(perm_tile_kernel<T, TA, TB, V, PERM_COPY>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p, ar, ai,
                                                                                              conj);
  else
    launch_clocked(perm_tile_kernel<T, TA, TB, V, PERM_ADD>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p, ar, ai,
                                                                                             conj);
}

// Widest access (elements) for type T: 16 bytes.
template <typename T>
constexpr int vec_width() {
  return sizeof(T) >= 16 ? 1 : 16 / (int)sizeof(T);
}

// A permutation after fusion: mode 0 = nothing to move, 1 = flat copy
// (identity), 2 = row problem, 3 = tile problem (tile ta x tb).
struct PermPlan {
  int mode = 0;
  int ta = 0, tb = 0;
  bool vec = false;  // 16-byte accesses
  int64_t nblk = 0, vol = 0;
  PermArgs p;
};

// Fuses the permutation and picks its kernel and launch shape.  Variant 0
// takes as third tile dimension the remaining one with the smallest input
// stride and enumerates the outer dimensions in input order; variant 1 takes
// the remaining one with the smallest output stride instead, and variant 2
// enumerates the outer dimensions in output-stride order.  Every variant moves
// each element once with the same arithmetic (bit-identical results); which
// is fastest depends on the permutation.
constexpr int kPermVariants = 3;
template <typename T>
int plan_perm(const T *src, T *dst, int rank, const int *dims, const int *o2n, PermPlan *pl, int variant = 0) {
  int64_t vol = 1;
  for (int k = 0; k < rank; ++k) vol *= dims[k];
  pl->mode = 0;
  if (vol == 0) return 0;
  int64_t oext[kMaxRank], ostride_of_out[kMaxRank];
  for (int k = 0; k < rank; ++k) oext[o2n[k]] = dims[k];
  int64_t acc = 1;
  for (int j = 0; j < rank; ++j) {
    ostride_of_out[j] = acc;
    acc *= oext[j];
  }
  // non-unit dims in input order with their output slots
  std::vector<int64_t> ext;
  std::vector<int> opos;
  for (int k = 0; k < rank; ++k)
    if (dims[k] > 1) {
      ext.push_back(dims[k]);
      opos.push_back(o2n[k]);
    }
  std::vector<int> orank(opos.size());
  {
    std::vector<int> sorted = opos;
    std::sort(sorted.begin(), sorted.end());
    for (size_t i = 0; i < opos.size(); ++i)
      orank[i] = (int)(std::lower_bound(sorted.begin(), sorted.end(), opos[i]) - sorted.begin());
  }
  // fuse runs of input-adjacent dims whose output slots are also adjacent
  std::vector<int64_t> fe, fis, fos;
  {
    int64_t istr = 1;
    size_t i = 0;
    while (i < ext.size()) {
      int64_t e = ext[i];
      const int64_t os = ostride_of_out[opos[i]];
      size_t j = i + 1;
      while (j < ext.size() && orank[j] == orank[j - 1] + 1) {
        e *= ext[j];
        ++j;
      }
      fe.push_back(e);
      fis.push_back(istr);
      fos.push_back(os);
      istr *= e;
      i = j;
    }
  }
  const int r = (int)fe.size();
  constexpr int VW = vec_width<T>();
  const bool aligned = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  if (r <= 1) {
    pl->mode = 1;
    pl->vol = vol;
    pl->vec = aligned;
    return 0;
  }
  int b = -1;
  for (int k = 0; k < r; ++k)
    if (fos[k] == 1) b = k;
  if (b < 0) return -3;
  const bool rows = (b == 0);
  if (rows) {  // second tile dim: the remaining dim with the smallest output stride
    int c = -1;
    for (int k = 1; k < r; ++k)
      if (c < 0 || fos[k] < fos[c]) c = k;
    b = c;
  }
  // third tile dim: the remaining dim with the smallest input stride (variant
  // 1: output stride)
  int c = -1;
  for (int k = 1; k < r; ++k)
    if (k != b && (c < 0 || (variant == 1 ? fos[k] < fos[c] : fis[k] < fis[c]))) c = k;
  PermArgs &p = pl->p;
  std::memset(&p, 0, sizeof(p));
  p.ea = fe[0];
  p.osa = fos[0];
  p.eb = fe[b];
  p.isb = fis[b];
  p.osb = fos[b];
  p.ec = c >= 0 ? fe[c] : 1;
  p.isc = c >= 0 ? fis[c] : 0;
  p.osc = c >= 0 ? fos[c] : 0;
  p.nouter = 0;
  int64_t nouter_vol = 1;
  std::vector<int> outer;
  for (int k = 1; k < r; ++k)
    if (k != b && k != c) outer.push_back(k);
  if (variant == 2) std::stable_sort(outer.begin(), outer.end(), [&](int x, int y) { return fos[x] < fos[y]; });
  for (int k : outer) {
    p.oe[p.nouter] = fe[k];
    p.ois[p.nouter] = fis[k];
    p.oos[p.nouter] = fos[k];
    p.nouter++;
    nouter_vol *= fe[k];
  }
  constexpr int EPB = elems_per_block<T>();
  int ta, tb;
  if (rows) {
    ta = p.ea >= 64 ? 64 : p.ea >= 32 ? 32 : p.ea >= 16 ? 16 : 8;
    // b: the smallest power of two covering its extent, within the block
    p.tb_log = 0;
    while ((1 << p.tb_log) < p.eb && (ta << (p.tb_log + 1)) <= EPB) ++p.tb_log;
    tb = 1 << p.tb_log;
  } else {
    const int64_t m = std::min(p.ea, p.eb);
    if (m >= 24) { ta = 32; tb = 32; }
    else if (m >= 12) { ta = 16; tb = 16; }
    else { ta = 8; tb = 8; }
  }
  p.tc = (int)std::min<int64_t>(p.ec, EPB / (ta * tb));
  p.ntile_a = (p.ea + ta - 1) / ta;
  p.ntile_b = (p.eb + tb - 1) / tb;
  p.ntile_c = (p.ec + p.tc - 1) / p.tc;
  const int64_t nblk = p.ntile_a * p.ntile_b * p.ntile_c * nouter_vol;
  if (nblk <= 0 || nblk > 0x7fffffffLL) return -4;
  // vector accesses need V | extent on each contiguous side (so every other
  // stride is a multiple of V) and 16-byte aligned bodies
  const bool vec = aligned && (p.ea % VW == 0) && (rows || p.eb % VW == 0) && VW > 1;
  pl->mode = rows ? 2 : 3;
  pl->ta = ta;
  pl->tb = tb;
  pl->vec = vec;
  pl->nblk = nblk;
  pl->vol = vol;
  return 0;
}

template <typename T>
int launch_plan(const T *src, T *dst, const PermPlan &pl, int op, double ar_, double ai_, int conj, hipStream_t s) {
  using R = typename Real<T>::type;
  const R ar = (R)ar_, ai = (R)ai_;
  constexpr int VW = vec_width<T>();
  const PermArgs &p = pl.p;
  const bool vec = pl.vec;
  const int64_t nblk = pl.nblk, vol = pl.vol;
  if (pl.mode == 0) return 0;
  if (pl.mode == 1) {
    const int g = grid_for(vol / VW);
    if (vec) {
      if (op == PERM_COPY)
        launch_clocked(perm_flat_kernel<T, VW, PERM_COPY>, dim3(g), dim3(kThreads), s, src, dst, vol, ar, ai, conj);
      else
        launch_clocked(perm_flat_kernel<T, VW, PERM_ADD>, dim3(g), dim3(kThreads), s, src, dst, vol, ar, ai, conj);
    } else {
      if (op == PERM_COPY)
        launch_clocked(perm_flat_kernel<T, 1, PERM_COPY>, dim3(g), dim3(kThreads), s, src, dst, vol, ar, ai, conj);
      else
        launch_clocked(perm_flat_kernel<T, 1, PERM_ADD>, dim3(g), dim3(kThreads), s, src, dst, vol, ar, ai, conj);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  const int ta = pl.ta;
  if (pl.mode == 2) {
#define TAL_ROWS(A)                                                            \
  (vec ? launch_rows<T, A, VW>(src, dst, p, op, ar, ai, conj, nblk, s)         \
       : launch_rows<T, A, 1>(src, dst, p, op, ar, ai, conj, nblk, s))
    if (ta == 64) TAL_ROWS(64);
    else if (ta == 32) TAL_ROWS(32);
    else if (ta == 16) TAL_ROWS(16);
    else TAL_ROWS(8);
#undef TAL_ROWS
  } else {
#define TAL_TILE(A, B)                                                         \
  (vec ? launch_tile<T, A, B, VW>(src, dst, p, op, ar, ai, conj, nblk, s)      \
       : launch_tile<T, A, B, 1>(src, dst, p, op, ar, ai, conj, nblk, s))
    if (ta == 32) TAL_TILE(32, 32);
    else if (ta == 16) TAL_TILE(16, 16);
    else TAL_TILE(8, 8);
#undef TAL_TILE
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Per-shape choice among the plan variants.  After a first untimed launch
// (which warms the shape's kernel and data), the next launches of a
// permutation shape on a device (extents, permutation, element size, access
// width) run one distinct variant each, timed by its kernel's dispatch
// timestamps (event timing around a launch adds ~20 us of noise); a later
// launch, once all of them have completed, keeps the fastest for the shape.
// Nothing waits on the events, and a launch made while the choice is open
// (or under stream capture) takes variant 0.  Only permutations of at least
// kPermTuneBytes are tuned (below that a launch is mostly latency), for at
// most kPermTuneMax shapes per process; cleared at shutdown.  On the 34
// permutations of a 16^6 R8 tensor the variants differ by 0-10 % (one by
// 40 %), and the slowest permutation after the choice runs 1 % faster.
constexpr size_t kPermTuneMax = 256;
constexpr size_t kPermTuneBytes = (size_t)8 << 20;
struct PermTune {
  int next = -1, choice = -1;
  bool distinct[kPermVariants] = {};
  hipEvent_t ev[kPermVariants][2] = {};
};
struct PermTuner {
  std::mutex mu;
  std::map<std::vector<int64_t>, PermTune> shapes;
};
PermTuner &perm_tuner() {
  static PermTuner t;
  return t;
}

void release_tune(PermTune &t) {
  for (auto &e : t.ev)
    for (auto &x : e) {
      if (x) (void)hipEventDestroy(x);
      x = nullptr;
    }
}

bool same_plan(const PermPlan &a, const PermPlan &b) {
  return a.mode == b.mode && a.ta == b.ta && a.tb == b.tb && a.vec == b.vec && a.nblk == b.nblk &&
         std::memcmp(&a.p, &b.p, sizeof(PermArgs)) == 0;
}

// The variant a launch of this shape runs; *ev gets the event pair to record
// around it while the variant is being timed (else nullptr).
template <typename T>
int pick_variant(const PermPlan &base, const T *src, T *dst, int rank, const int *dims, const int *o2n,
                 hipStream_t s, hipEvent_t **ev) {
  *ev = nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::vector<int64_t> key = {dev, (int64_t)sizeof(T), base.vec ? 1 : 0, rank};
  key.insert(key.end(), dims, dims + rank);
  key.insert(key.end(), o2n, o2n + rank);
  PermTuner &tu = perm_tuner();
  std::lock_guard<std::mutex> lk(tu.mu);
  auto it = tu.shapes.find(key);
  if (it == tu.shapes.end()) {
    if (tu.shapes.size() >= kPermTuneMax) return 0;
    PermTune t;
    t.distinct[0] = true;
    for (int v = 1; v < kPermVariants; ++v) {
      PermPlan q;
      t.distinct[v] = plan_perm(src, dst, rank, dims, o2n, &q, v) == 0 && !same_plan(q, base);
    }
    it = tu.shapes.emplace(std::move(key), t).first;
  }
  PermTune &t = it->second;
  if (t.choice >= 0) return t.choice;
  if (t.next < 0) {  // the shape's first launch
    t.next = 0;
    return 0;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return 0;
  while (t.next < kPermVariants && !t.distinct[t.next]) ++t.next;
  if (t.next < kPermVariants) {
    const int v = t.next++;
    if ((t.ev[v][0] || hipEventCreate(&t.ev[v][0]) == hipSuccess) &&
        (t.ev[v][1] || hipEventCreate(&t.ev[v][1]) == hipSuccess)) {
      *ev = t.ev[v];
    } else {
      t.distinct[v] = false;
    }
    return v;
  }
  for (int v = 0; v < kPermVariants; ++v)
    if (t.distinct[v] && t.ev[v][1] && hipEventQuery(t.ev[v][1]) == hipErrorNotReady) return 0;
  float best = 0.0f;
  int bv = 0;
  bool any = false;
  for (int v = 0; v < kPermVariants; ++v) {
    float ms = 0.0f;
    if (t.distinct[v] && t.ev[v][0] && t.ev[v][1] && hipEventElapsedTime(&ms, t.ev[v][0], t.ev[v][1]) == hipSuccess &&
        (!any || ms < best)) {
      best = ms;
      bv = v;
      any = true;
    }
  }
  t.choice = bv;
  release_tune(t);
  return bv;
}

template <typename T>
int launch_perm_typed(const T *src, T *dst, int rank, const int *dims, const int *o2n, int op, double ar,
                      double ai, int conj, hipStream_t s) {
  PermPlan pl;
  if (int e = plan_perm(src, dst, rank, dims, o2n, &pl)) return e;
  if (pl.mode < 2 || (size_t)pl.vol * sizeof(T) < kPermTuneBytes)
    return launch_plan(src, dst, pl, op, ar, ai, conj, s);
  hipEvent_t *ev = nullptr;
  const int v = pick_variant(pl, src, dst, rank, dims, o2n, s, &ev);
  if (v != 0 && plan_perm(src, dst, rank, dims, o2n, &pl, v)) return -3;
  if (!ev) return launch_plan(src, dst, pl, op, ar, ai, conj, s);
  // a timed variant: its kernel's own dispatch timestamps go to the shape's
  // events; the calling operation's clock, if on, gets event records around it
  KernelClock &kc = kernel_clock();
  const KernelClock op_clock = kc;
  if (op_clock.on && !op_clock.started) (void)hipEventRecord(op_clock.t0, s);
  kc.on = true;
  kc.started = false;
  kc.t0 = ev[0];
  kc.t1 = ev[1];
  const int e = launch_plan(src, dst, pl, op, ar, ai, conj, s);
  kc = op_clock;
  if (op_clock.on) {
    (void)hipEventRecord(op_clock.t1, s);
    kc.started = true;
  }
  return e;
}

template <typename T, int TA, int TB, int V, bool ROWS>
void launch_pair(const PermPair<T> &q, int64_t nblk, hipStream_t s) {
  launch_clocked(perm_pair_kernel<T, TA, TB, V, ROWS>, dim3((unsigned)nblk), dim3(kThreads), s, q);
}

// Two permutation copies: one launch when both take the same kernel (row or
// tile problems of one tile shape and access width with few outer dims),
// else one launch each.
template <typename T>
int launch_perm_pair_typed(const T *src0, T *dst0, int rank0, const int *dims0, const int *o2n0, const T *src1,
                           T *dst1, int rank1, const int *dims1, const int *o2n1, hipStream_t s) {
  PermPlan a, b;
  if (int e = plan_perm(src0, dst0, rank0, dims0, o2n0, &a)) return e;
  if (int e = plan_perm(src1, dst1, rank1, dims1, o2n1, &b)) return e;
  const bool pair = a.mode >= 2 && a.mode == b.mode && a.ta == b.ta && a.tb == b.tb && a.vec == b.vec &&
                    a.p.nouter <= kPairOuter && b.p.nouter <= kPairOuter && a.nblk + b.nblk <= 0x7fffffffLL;
  if (!pair) {
    if (int e = launch_plan(src0, dst0, a, PERM_COPY, 1.0, 0.0, 0, s)) return e;
    return launch_plan(src1, dst1, b, PERM_COPY, 1.0, 0.0, 0, s);
  }
  PermPair<T> q;
  std::memset(&q, 0, sizeof(q));
  q.src[0] = src0;
  q.dst[0] = dst0;
  q.src[1] = src1;
  q.dst[1] = dst1;
  q.nblk0 = a.nblk;
  const PermPlan *pl[2] = {&a, &b};
  for (int j = 0; j < 2; ++j) {
    const PermArgs &x = pl[j]->p;
    auto &y = q.p[j];
    y.nouter = x.nouter;
    y.ea = x.ea, y.eb = x.eb, y.ec = x.ec;
    y.isb = x.isb, y.isc = x.isc;
    y.osa = x.osa, y.osb = x.osb, y.osc = x.osc;
    y.ntile_a = x.ntile_a, y.ntile_b = x.ntile_b, y.ntile_c = x.ntile_c;
    y.tc = x.tc;
    y.tb_log = x.tb_log;
    for (int k = 0; k < x.nouter; ++k) {
      y.oe[k] = x.oe[k];
      y.ois[k] = x.ois[k];
      y.oos[k] = x.oos[k];
    }
  }
  constexpr int VW = vec_width<T>();
  const int64_t n = a.nblk + b.nblk;
  if (a.mode == 2) {
#define TAL_PAIR_ROWS(A) \
  (a.vec ? launch_pair<T, A, 1, VW, true>(q, n, s) : launch_pair<T, A, 1, 1, true>(q, n, s))
    if (a.ta == 64) TAL_PAIR_ROWS(64);
    else if (a.ta == 32) TAL_PAIR_ROWS(32);
    else if (a.ta == 16) TAL_PAIR_ROWS(16);
    else TAL_PAIR_ROWS(8);
#undef TAL_PAIR_ROWS
  } else {
#define TAL_PAIR_TILE(A) \
  (a.vec ? launch_pair<T, A, A, VW, false>(q, n, s) : launch_pair<T, A, A, 1, false>(q, n, s))
    if (a.ta == 32) TAL_PAIR_TILE(32);
    else if (a.ta == 16) TAL_PAIR_TILE(16);
    else TAL_PAIR_TILE(8);
#undef TAL_PAIR_TILE
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename T>
T make_val(double vr, double vi) {
  if constexpr (std::is_floating_point<T>::value) {
    return (T)vr;
  } else {
    T v;
    v.x = (typename Real<T>::type)vr;
    v.y = (typename Real<T>::type)vi;
    return v;
  }
}

// ------------------------------------------------------------ slice / insert
// A block of a dlf tensor ("slice", extents S) is copied out of or into a
// larger dlf tensor (extents B) at offsets O.  Both keep the dimension order,
// so after the host drops unit extents and fuses dimensions the slice spans
// fully, the problem is a set of rows: a run of L elements contiguous on both
// sides (L = 1 when the first remaining dimension is strided in the big
// tensor), enumerated by the remaining row dimensions.  A block covers
// kSliceRows x `ry` rows x (rx * V) run elements: each thread moves V
// elements of kSliceRows rows (all loads issued before the stores), and
// decomposes each row index with 32-bit arithmetic when the row count allows
// (the last row digit needs no division).
constexpr int kSliceRows = 4;
struct SliceArgs {
  int nrd;           // row dimensions
  int rx_log;        // log2 of threads along the run
  int64_t L;         // run length (elements)
  int64_t nrows, nchunk;
  int64_t src0, dst0;
  int64_t re[kMaxRank], rss[kMaxRank], rds[kMaxRank];  // row extents, source / destination strides
};

template <typename R>
__device__ inline void slice_row(const SliceArgs &p, R r, int64_t *so, int64_t *dof) {
  for (int k = 0; k + 1 < p.nrd; ++k) {
    const R q = r / (R)p.re[k];
    const R i = r - q * (R)p.re[k];
    r = q;
    *so += (int64_t)i * p.rss[k];
    *dof += (int64_t)i * p.rds[k];
  }
  if (p.nrd > 0) {
    *so += (int64_t)r * p.rss[p.nrd - 1];
    *dof += (int64_t)r * p.rds[p.nrd - 1];
  }
}

template <typename T, int V, int OP, bool NARROW>
__global__ __launch_bounds__(kThreads) void slice_kernel(const T *__restrict__ src, T *__restrict__ dst, SliceArgs p) {
  const int rx = 1 << p.rx_log, ry = kThreads >> p.rx_log;
  const int tx = threadIdx.x & (rx - 1), ty = threadIdx.x >> p.rx_log;
  const int64_t chunk = (int64_t)blockIdx.x % p.nchunk, blk = (int64_t)blockIdx.x / p.nchunk;
  const int64_t e = (chunk * rx + tx) * V;
  if (e >= p.L) return;
  int64_t so[kSliceRows], dof[kSliceRows];
  bool ok[kSliceRows];
#pragma unroll
  for (int i = 0; i < kSliceRows; ++i) {
    const int64_t r = (blk * kSliceRows + i) * ry + ty;
    ok[i] = r < p.nrows;
    so[i] = p.src0 + e;
    dof[i] = p.dst0 + e;
    if (ok[i]) {
      if (NARROW)
        slice_row<uint32_t>(p, (uint32_t)r, &so[i], &dof[i]);
      else
        slice_row<int64_t>(p, r, &so[i], &dof[i]);
    }
  }
  Pack<T, V> v[kSliceRows];
#pragma unroll
  for (int i = 0; i < kSliceRows; ++i)
    if (ok[i]) v[i] = ldv<T, V>(src + so[i]);
#pragma unroll
  for (int i = 0; i < kSliceRows; ++i)
    if (ok[i]) emit<T, V, OP>(dst + dof[i], v[i], (typename Real<T>::type)1, (typename Real<T>::type)0, 0);
}

template <typename T>
int launch_slice_typed(int mode, const T *src, T *dst, int rank, const int *sdims, const int *bdims,
                       const int *offs, int accumulative, hipStream_t s) {
  // strides of both tensors; dims with slice extent 1 only move the base
  int64_t sst[kMaxRank], bst[kMaxRank], vol = 1, bvol = 1, base = 0;
  for (int k = 0; k < rank; ++k) {
    if (sdims[k] < 1 || bdims[k] < 1 || offs[k] < 0 || (int64_t)offs[k] + sdims[k] > bdims[k]) return -1;
    sst[k] = vol;
    bst[k] = bvol;
    vol *= sdims[k];
    bvol *= bdims[k];
    base += (int64_t)offs[k] * bst[k];
  }
  int64_t ext[kMaxRank], xs[kMaxRank], xb[kMaxRank];  // fused dims: extent, slice stride, big stride
  int n = 0;
  for (int k = 0; k < rank; ++k) {
    if (sdims[k] == 1) continue;
    if (n > 0 && xs[n - 1] * ext[n - 1] == sst[k] && xb[n - 1] * ext[n - 1] == bst[k]) {
      ext[n - 1] *= sdims[k];
      continue;
    }
    ext[n] = sdims[k];
    xs[n] = sst[k];
    xb[n] = bst[k];
    ++n;
  }
  SliceArgs p;
  int first = 0;
  p.L = 1;
  if (n > 0 && xs[0] == 1 && xb[0] == 1) {
    p.L = ext[0];
    first = 1;
  }
  p.nrd = n - first;
  p.nrows = 1;
  for (int k = first; k < n; ++k) {
    p.re[k - first] = ext[k];
    p.rss[k - first] = mode == 0 ? xb[k] : xs[k];
    p.rds[k - first] = mode == 0 ? xs[k] : xb[k];
    p.nrows *= ext[k];
  }
  p.src0 = mode == 0 ? base : 0;
  p.dst0 = mode == 0 ? 0 : base;
  // vector width: every run start and both bases must be whole vectors
  constexpr int VW = vec_width<T>();
  bool vec = VW > 1 && p.L % VW == 0 && p.src0 % VW == 0 && p.dst0 % VW == 0 && (uintptr_t)src % 16 == 0 &&
             (uintptr_t)dst % 16 == 0;
  for (int k = 0; k < p.nrd && vec; ++k) vec = p.rss[k] % VW == 0 && p.rds[k] % VW == 0;
  const int V = vec ? VW : 1;
  const int64_t nvec = p.L / V;
  int rx_log = 0;
  while (rx_log < 8 && ((int64_t)1 << rx_log) < nvec) ++rx_log;
  p.rx_log = rx_log;
  const int64_t rx = (int64_t)1 << rx_log, ry = kThreads / rx;
  p.nchunk = (nvec + rx - 1) / rx;
  const int64_t nblk = p.nchunk * ((p.nrows + ry * kSliceRows - 1) / (ry * kSliceRows));
  if (nblk > 0x7fffffffLL) return -4;
  const bool narrow = p.nrows <= 0xffffffffLL;  // row indices and extents fit 32 bits
#define TAL_SLICE_LAUNCH(VV, NW)                                                                                   \
  (accumulative                                                                                                  \
       ? launch_clocked(slice_kernel<T, VV, PERM_ADD, NW>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p) \
       : launch_clocked(slice_kernel<T, VV, PERM_COPY, NW>, dim3((unsigned)nblk), dim3(kThreads), s, src, dst, p))
  if (V > 1) {
    if (narrow)
      TAL_SLICE_LAUNCH(VW, true);
    else
      TAL_SLICE_LAUNCH(VW, false);
  } else {
    if (narrow)
      TAL_SLICE_LAUNCH(1, true);
    else
      TAL_SLICE_LAUNCH(1, false);
  }
#undef TAL_SLICE_LAUNCH
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace

static bool valid_perm(int rank, const int *dims, const int *o2n) {
  if (rank < 0 || rank > kMaxRank || (rank > 0 && (!dims || !o2n))) return false;
  bool seen[kMaxRank] = {false};
  for (int k = 0; k < rank; ++k) {
    if (dims[k] < 0 || o2n[k] < 0 || o2n[k] >= rank || seen[o2n[k]]) return false;
    seen[o2n[k]] = true;
  }
  return true;
}

void perm_tune_clear() {
  PermTuner &tu = perm_tuner();
  std::lock_guard<std::mutex> lk(tu.mu);
  for (auto &x : tu.shapes) release_tune(x.second);
  tu.shapes.clear();
}

int launch_permute(int dk, const void *src, void *dst, int rank, const int *dims, const int *o2n, int op,
                   double ar, double ai, int conj, hipStream_t s) {
  if (!src || !dst || !valid_perm(rank, dims, o2n)) return -1;
  if (op != PERM_COPY && op != PERM_ADD) return -1;
  switch (dk) {
    case 4: return launch_perm_typed<float>((const float *)src, (float *)dst, rank, dims, o2n, op, ar, ai, 0, s);
    case 8: return launch_perm_typed<double>((const double *)src, (double *)dst, rank, dims, o2n, op, ar, ai, 0, s);
    case 14:
      return launch_perm_typed<float2>((const float2 *)src, (float2 *)dst, rank, dims, o2n, op, ar, ai, conj, s);
    case 18:
      return launch_perm_typed<double2>((const double2 *)src, (double2 *)dst, rank, dims, o2n, op, ar, ai, conj, s);
    default: return -1;
  }
}


int launch_permute_pair(int dk, const void *src0, void *dst0, int rank0, const int *dims0, const int *o2n0,
                        const void *src1, void *dst1, int rank1, const int *dims1, const int *o2n1, hipStream_t s) {
  if (!src0 || !dst0 || !src1 || !dst1 || !valid_perm(rank0, dims0, o2n0) || !valid_perm(rank1, dims1, o2n1))
    return -1;
#define TAL_PAIR_ARGS(T)                                                                                    \
  (const T *)src0, (T *)dst0, rank0, dims0, o2n0, (const T *)src1, (T *)dst1, rank1, dims1, o2n1, s
  switch (dk) {
    case 4: return launch_perm_pair_typed<float>(TAL_PAIR_ARGS(float));
    case 8: return launch_perm_pair_typed<double>(TAL_PAIR_ARGS(double));
    case 14: return launch_perm_pair_typed<float2>(TAL_PAIR_ARGS(float2));
    case 18: return launch_perm_pair_typed<double2>(TAL_PAIR_ARGS(double2));
    default: return -1;
  }
#undef TAL_PAIR_ARGS
}

int launch_slice(int dk, int mode, const void *src, void *dst, int rank, const int *sdims, const int *bdims,
                 const int *offsets, int accumulative, hipStream_t s) {
  if (rank < 0 || rank > kMaxRank || (rank > 0 && (!sdims || !bdims || !offsets)) || !src || !dst ||
      (mode != 0 && mode != 1))
    return -1;
  for (int k = 0; k < rank; ++k)
    if (sdims[k] < 1 || offsets[k] < 0 || bdims[k] < 1 || (int64_t)offsets[k] + sdims[k] > bdims[k]) return -1;
  switch (dk) {
    case 4: return launch_slice_typed(mode, (const float *)src, (float *)dst, rank, sdims, bdims, offsets, accumulative, s);
    case 8: return launch_slice_typed(mode, (const double *)src, (double *)dst, rank, sdims, bdims, offsets, accumulative, s);
    case 14: return launch_slice_typed(mode, (const float2 *)src, (float2 *)dst, rank, sdims, bdims, offsets, accumulative, s);
    case 18: return launch_slice_typed(mode, (const double2 *)src, (double2 *)dst, rank, sdims, bdims, offsets, accumulative, s);
    default: return -1;
  }
}

int launch_init(int dk, void *d, size_t vol, double vr, double vi, hipStream_t s) {
  if (!d) return -1;
  if (vol == 0) return 0;
  const int g = grid_for((int64_t)vol);
  switch (dk) {
    case 4: launch_clocked(init_kernel<float>, dim3(g), dim3(kThreads), s, (float *)d, vol, make_val<float>(vr, vi)); break;
    case 8: launch_clocked(init_kernel<double>, dim3(g), dim3(kThreads), s, (double *)d, vol, make_val<double>(vr, vi)); break;
    case 14: launch_clocked(init_kernel<float2>, dim3(g), dim3(kThreads), s, (float2 *)d, vol, make_val<float2>(vr, vi)); break;
    case 18: launch_clocked(init_kernel<double2>, dim3(g), dim3(kThreads), s, (double2 *)d, vol, make_val<double2>(vr, vi)); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_scale(int dk, void *d, size_t vol, double ar, double ai, hipStream_t s) {
  if (!d) return -1;
  if (vol == 0) return 0;
  const int g = grid_for((int64_t)vol);
  switch (dk) {
    case 4: launch_clocked(scale_kernel<float>, dim3(g), dim3(kThreads), s, (float *)d, vol, (float)ar, (float)ai); break;
    case 8: launch_clocked(scale_kernel<double>, dim3(g), dim3(kThreads), s, (double *)d, vol, ar, ai); break;
    case 14: launch_clocked(scale_kernel<float2>, dim3(g), dim3(kThreads), s, (float2 *)d, vol, (float)ar, (float)ai); break;
    case 18: launch_clocked(scale_kernel<double2>, dim3(g), dim3(kThreads), s, (double2 *)d, vol, ar, ai); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t norm2_scratch_elems() { return kNormBlocks; }

// Fixed scratch of the allocation-free norm2 (stage-1 partial sums, one double
// per block; stage 1 writes every partial that stage 2 reads, so nothing needs
// clearing between calls).  Each GPU holds its own copy of the array for the
// life of the process; host code reaches it only through the address that
// hipGetSymbolAddress returns on that GPU.
static __device__ __align__(256) double scratch[kNormBlocks];
static double *g_scratch[kNoAllocMaxDevices] = {};

int noalloc_scratch_setup(int hip_dev) {
  if (hip_dev < 0 || hip_dev >= kNoAllocMaxDevices) return -1;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
  g_scratch[hip_dev] = (double *)p;
  return 0;
}

size_t norm2_noalloc_scratch_bytes() { return sizeof(scratch); }

int launch_norm2_noalloc(int hip_dev, int dk, const void *src, size_t vol, void *out, hipStream_t s) {
  if (hip_dev < 0 || hip_dev >= kNoAllocMaxDevices || !g_scratch[hip_dev]) return -3;
  // the stage-1 grid never exceeds kNormBlocks partials, the array's length
  static_assert(sizeof(scratch) == kNormBlocks * sizeof(double), "one partial per stage-1 block");
  return launch_norm2(dk, src, vol, g_scratch[hip_dev], out, s);
}

KernelClock &kernel_clock() {
  static thread_local KernelClock c;
  return c;
}

template <typename T>
static int launch_norm2_typed(const T *x, size_t vol, double *partials, void *out, hipStream_t s) {
  const int64_t n = (int64_t)vol;
  constexpr int VW = vec_width<T>();
  const bool aligned = ((uintptr_t)x % 16) == 0;
  const int64_t packs = aligned ? n / VW : n;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kNormBlocks, (packs + 4 * kThreads - 1) / (4 * kThreads)));
  if (aligned)
    launch_clocked(norm2_stage1<T, VW>, dim3(nb), dim3(kThreads), s, x, n, partials);
  else
    launch_clocked(norm2_stage1<T, 1>, dim3(nb), dim3(kThreads), s, x, n, partials);
  launch_clocked(norm2_stage2<T>, dim3(1), dim3(kThreads), s, partials, nb, (T *)out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_norm2(int dk, const void *src, size_t vol, double *partials, void *out, hipStream_t s) {
  if (!src || !partials || !out) return -1;
  switch (dk) {
    case 4: return launch_norm2_typed<float>((const float *)src, vol, partials, out, s);
    case 8: return launch_norm2_typed<double>((const double *)src, vol, partials, out, s);
    case 14: return launch_norm2_typed<float2>((const float2 *)src, vol, partials, out, s);
    case 18: return launch_norm2_typed<double2>((const double2 *)src, vol, partials, out, s);
    default: return -1;
  }
}

}  // namespace tal
