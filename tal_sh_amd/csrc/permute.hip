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
struct PermArgs {
  int nouter;                 // number of outer (block-enumerated) dims
  int64_t ea, eb;             // extents of tile dims a and b
  int64_t isb, osa;           // input stride of b, output stride of a
  int64_t ntile_a, ntile_b;   // tile counts along a and b
  int64_t oe[kMaxRank];       // outer extents
  int64_t ois[kMaxRank];      // outer input strides
  int64_t oos[kMaxRank];      // outer output strides
};

__device__ inline void outer_offsets(const PermArgs &p, int64_t blk, int64_t &ta, int64_t &tb, int64_t &ibase,
                                     int64_t &obase) {
  ta = blk % p.ntile_a;
  blk /= p.ntile_a;
  tb = blk % p.ntile_b;
  blk /= p.ntile_b;
  ibase = 0;
  obase = 0;
  for (int k = 0; k < p.nouter; ++k) {
    int64_t e = p.oe[k];
    int64_t q = blk / e;
    int64_t r = blk - q * e;
    blk = q;
    ibase += r * p.ois[k];
    obase += r * p.oos[k];
  }
}

// Output fastest dim == input fastest dim (a): stream rows of a, rows indexed by b.
template <typename T, int TA, int TB, int OP>
__global__ __launch_bounds__(kThreads) void perm_rows_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             PermArgs p, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  int64_t ta, tb, ib, ob;
  outer_offsets(p, blockIdx.x, ta, tb, ib, ob);
  const int64_t a0 = ta * TA, b0 = tb * TB;
  const int na = (int)min<int64_t>(TA, p.ea - a0), nb = (int)min<int64_t>(TB, p.eb - b0);
  ib += a0 + b0 * p.isb;
  ob += a0 + b0 * p.osa;  // here osa holds the output stride of b
#pragma unroll
  for (int e = threadIdx.x; e < TA * TB; e += kThreads) {
    const int ia = e % TA, jb = e / TA;
    if (ia < na && jb < nb) {
      T v = Elem<T>::conj(src[ib + ia + jb * p.isb], conj);
      T *o = dst + ob + ia + jb * p.osa;
      if (OP == PERM_COPY)
        *o = v;
      else
        *o = Elem<T>::axpy(*o, v, ar, ai);
    }
  }
}

// General case: tile (a = input fastest) x (b = output fastest) through LDS.
template <typename T, int TA, int TB, int OP>
__global__ __launch_bounds__(kThreads) void perm_tile_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             PermArgs p, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  __shared__ T tile[TB][TA + 1];
  int64_t ta, tb, ib, ob;
  outer_offsets(p, blockIdx.x, ta, tb, ib, ob);
  const int64_t a0 = ta * TA, b0 = tb * TB;
  const int na = (int)min<int64_t>(TA, p.ea - a0), nb = (int)min<int64_t>(TB, p.eb - b0);
  ib += a0 + b0 * p.isb;
  ob += b0 + a0 * p.osa;
#pragma unroll
  for (int e = threadIdx.x; e < TA * TB; e += kThreads) {
    const int ia = e % TA, jb = e / TA;
    if (ia < na && jb < nb) tile[jb][ia] = src[ib + ia + jb * p.isb];
  }
  __syncthreads();
#pragma unroll
  for (int e = threadIdx.x; e < TA * TB; e += kThreads) {
    const int jb = e % TB, ia = e / TB;
    if (ia < na && jb < nb) {
      T v = Elem<T>::conj(tile[jb][ia], conj);
      T *o = dst + ob + jb + ia * p.osa;
      if (OP == PERM_COPY)
        *o = v;
      else
        *o = Elem<T>::axpy(*o, v, ar, ai);
    }
  }
}

// Contiguous (identity after fusion) copy / add.
template <typename T, int OP>
__global__ __launch_bounds__(kThreads) void perm_flat_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             int64_t n, typename Real<T>::type ar,
                                                             typename Real<T>::type ai, int conj) {
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (; i < n; i += stride) {
    T v = Elem<T>::conj(src[i], conj);
    if (OP == PERM_COPY)
      dst[i] = v;
    else
      dst[i] = Elem<T>::axpy(dst[i], v, ar, ai);
  }
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
constexpr int kNormBlocks = 1024;

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

// Stage 1: block b sums the fixed contiguous chunk [b*chunk, (b+1)*chunk).
template <typename T>
__global__ __launch_bounds__(kThreads) void norm2_stage1(const T *__restrict__ x, int64_t n, int64_t chunk,
                                                         double *__restrict__ partials) {
  const int64_t beg = (int64_t)blockIdx.x * chunk;
  const int64_t end = min<int64_t>(n, beg + chunk);
  double acc = 0.0;
  for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) acc += Elem<T>::sq(x[i]);
  double t = block_sum(acc);
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
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 8192));
}

template <typename T, int TA, int TB>
void launch_rows(const T *src, T *dst, const PermArgs &p, int op, typename Real<T>::type ar,
                 typename Real<T>::type ai, int conj, int64_t nblk, hipStream_t s) {
  if (op == PERM_COPY)
    perm_rows_kernel<T, TA, TB, PERM_COPY><<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(src, dst, p, ar, ai, conj);
  else
    perm_rows_kernel<T, TA, TB, PERM_ADD><<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(src, dst, p, ar, ai, conj);
}

template <typename T, int TA, int TB>
void launch_tile(const T *src, T *dst, const PermArgs &p, int op, typename Real<T>::type ar,
                 typename Real<T>::type ai, int conj, int64_t nblk, hipStream_t s) {
  if (op == PERM_COPY)
    perm_tile_kernel<T, TA, TB, PERM_COPY><<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(src, dst, p, ar, ai, conj);
  else
    perm_tile_kernel<T, TA, TB, PERM_ADD><<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(src, dst, p, ar, ai, conj);
}

// Fuses the permutation and launches the matching kernel.
template <typename T>
int launch_perm_typed(const T *src, T *dst, int rank, const int *dims, const int *o2n, int op, double ar_,
                      double ai_, int conj, hipStream_t s) {
  using R = typename Real<T>::type;
  const R ar = (R)ar_, ai = (R)ai_;
  int64_t vol = 1;
  for (int k = 0; k < rank; ++k) vol *= dims[k];
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
  if (r <= 1) {
    const int g = grid_for(vol);
    if (op == PERM_COPY)
      perm_flat_kernel<T, PERM_COPY><<<g, kThreads, 0, s>>>(src, dst, vol, ar, ai, conj);
    else
      perm_flat_kernel<T, PERM_ADD><<<g, kThreads, 0, s>>>(src, dst, vol, ar, ai, conj);
    return hipGetLastError() == hipSuccess ? 0 : -2;
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
  PermArgs p;
  std::memset(&p, 0, sizeof(p));
  p.ea = fe[0];
  p.eb = fe[b];
  p.isb = fis[b];
  p.osa = rows ? fos[b] : fos[0];
  p.nouter = 0;
  int64_t nouter_vol = 1;
  for (int k = 1; k < r; ++k) {
    if (k == b) continue;
    p.oe[p.nouter] = fe[k];
    p.ois[p.nouter] = fis[k];
    p.oos[p.nouter] = fos[k];
    p.nouter++;
    nouter_vol *= fe[k];
  }
  int ta, tb;
  if (rows) {
    if (p.ea >= 64) { ta = 64; tb = 16; }
    else if (p.ea >= 32) { ta = 32; tb = 32; }
    else if (p.ea >= 16) { ta = 16; tb = 64; }
    else { ta = 8; tb = 128; }
  } else {
    const int64_t m = std::min(p.ea, p.eb);
    if (m >= 24) { ta = 32; tb = 32; }
    else if (m >= 12) { ta = 16; tb = 16; }
    else { ta = 8; tb = 8; }
  }
  p.ntile_a = (p.ea + ta - 1) / ta;
  p.ntile_b = (p.eb + tb - 1) / tb;
  const int64_t nblk = p.ntile_a * p.ntile_b * nouter_vol;
  if (nblk <= 0 || nblk > 0x7fffffffLL) return -4;
  if (rows) {
    if (ta == 64) launch_rows<T, 64, 16>(src, dst, p, op, ar, ai, conj, nblk, s);
    else if (ta == 32) launch_rows<T, 32, 32>(src, dst, p, op, ar, ai, conj, nblk, s);
    else if (ta == 16) launch_rows<T, 16, 64>(src, dst, p, op, ar, ai, conj, nblk, s);
    else launch_rows<T, 8, 128>(src, dst, p, op, ar, ai, conj, nblk, s);
  } else {
    if (ta == 32) launch_tile<T, 32, 32>(src, dst, p, op, ar, ai, conj, nblk, s);
    else if (ta == 16) launch_tile<T, 16, 16>(src, dst, p, op, ar, ai, conj, nblk, s);
    else launch_tile<T, 8, 8>(src, dst, p, op, ar, ai, conj, nblk, s);
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

}  // namespace

int launch_permute(int dk, const void *src, void *dst, int rank, const int *dims, const int *o2n, int op,
                   double ar, double ai, int conj, hipStream_t s) {
  if (rank < 0 || rank > kMaxRank || (rank > 0 && (!dims || !o2n)) || !src || !dst) return -1;
  if (op != PERM_COPY && op != PERM_ADD) return -1;
  bool seen[kMaxRank] = {false};
  for (int k = 0; k < rank; ++k) {
    if (dims[k] < 0 || o2n[k] < 0 || o2n[k] >= rank || seen[o2n[k]]) return -1;
    seen[o2n[k]] = true;
  }
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

int launch_init(int dk, void *d, size_t vol, double vr, double vi, hipStream_t s) {
  if (!d) return -1;
  if (vol == 0) return 0;
  const int g = grid_for((int64_t)vol);
  switch (dk) {
    case 4: init_kernel<float><<<g, kThreads, 0, s>>>((float *)d, vol, make_val<float>(vr, vi)); break;
    case 8: init_kernel<double><<<g, kThreads, 0, s>>>((double *)d, vol, make_val<double>(vr, vi)); break;
    case 14: init_kernel<float2><<<g, kThreads, 0, s>>>((float2 *)d, vol, make_val<float2>(vr, vi)); break;
    case 18: init_kernel<double2><<<g, kThreads, 0, s>>>((double2 *)d, vol, make_val<double2>(vr, vi)); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_scale(int dk, void *d, size_t vol, double ar, double ai, hipStream_t s) {
  if (!d) return -1;
  if (vol == 0) return 0;
  const int g = grid_for((int64_t)vol);
  switch (dk) {
    case 4: scale_kernel<float><<<g, kThreads, 0, s>>>((float *)d, vol, (float)ar, (float)ai); break;
    case 8: scale_kernel<double><<<g, kThreads, 0, s>>>((double *)d, vol, ar, ai); break;
    case 14: scale_kernel<float2><<<g, kThreads, 0, s>>>((float2 *)d, vol, (float)ar, (float)ai); break;
    case 18: scale_kernel<double2><<<g, kThreads, 0, s>>>((double2 *)d, vol, ar, ai); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t norm2_scratch_elems() { return kNormBlocks; }

template <typename T>
static int launch_norm2_typed(const T *x, size_t vol, double *partials, void *out, hipStream_t s) {
  int64_t n = (int64_t)vol;
  int64_t chunk = (n + kNormBlocks - 1) / kNormBlocks;
  chunk = ((chunk + kThreads - 1) / kThreads) * kThreads;
  if (chunk == 0) chunk = kThreads;
  const int nb = (int)std::max<int64_t>(1, (n + chunk - 1) / chunk);
  norm2_stage1<T><<<nb, kThreads, 0, s>>>(x, n, chunk, partials);
  norm2_stage2<T><<<1, kThreads, 0, s>>>(partials, nb, (T *)out);
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
