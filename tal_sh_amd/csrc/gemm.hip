// MFMA GEMM kernels for the TTGT contraction stage (gfx950).
//
// Replaces the GEMM call of gpu_tensor_block_contract_dlf
// (tensor_algebra_gpu_nvidia.cu:6851-7060, cuBLAS/hipBLAS) and its in-house
// fallback gpu_matrix_multiply_tn__ (:2263-2312).  Contract:
//   C(MxN, column-major, ldc) = alpha * op(A) * op(B) + beta * C,  beta in {0,1}
// with op(A) in {A, A^T} and op(B) in {B, B^T}, each optionally conjugated,
// so that the contraction planner can feed tensor operands in whichever
// matrix form their storage already has.
//
// Kernel shape: NW wave64s per workgroup (template parameter): single GEMMs
// run 8 waves in a 2x4 grid (512 threads), batched GEMMs 4 waves in a 2x2
// grid; each wave owns a (BM/2)x(BN/(NW/2)) slab of C built from 16x16
// MFMA tiles:
//   R8/C8: v_mfma_f64_16x16x4_f64   R4/C4: v_mfma_f32_16x16x4_f32
// Operand tiles (BK=16) are double-buffered in LDS as K-contiguous rows padded
// by one 16-byte vector (conflict-free fragment reads, aligned rows); the next
// tile is loaded from HBM into registers with 16-byte accesses while the
// current one is multiplied, so each K step needs a single barrier.  Blocks
// are renumbered so that each XCD (dispatch round-robins workgroups over the
// 8 XCDs) walks a compact band of C tiles and the operand panels it streams
// stay in its own L2.  Complex products use four real MFMAs per step
// (re += ar*br - ai*bi, im += ar*bi + ai*br), conjugation folded into the sign
// of the imaginary operand.  MFMA fragment layout (16x16x4): A lane l holds
// A[l%16][l/16], B lane l holds B[l/16][l%16]; accumulator register r of lane
// l holds C[4*(l/16)+r][l%16] for f32 and C[4*r+l/16][l%16] for f64.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace tal {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename S>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = f64x4;
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<float> {
  using acc_t = f32x4;
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int kThreads = 256;
constexpr int BK = 16;
constexpr int kXcds = 8;

// Waves per workgroup of gemm_kernel (template parameter NW), in 2 rows x
// NW/2 columns: 8 (4 per SIMD) for single GEMMs — at 9216^3, against 4 waves
// of 2 x 2: FP64 64.5 -> 65.6, FP32 124.8 -> 127.3, complex FP64 65.5 -> 66.8
// TFLOP/s — and 4 (2 per SIMD) for batched ones, whose 16 x 1024^3 case ran
// 48.1 TFLOP/s with 4 waves and 42.8 with 8.
constexpr int kWavesSingle = 8, kWavesBatched = 4;
// K step (template parameter KB): 16 for FP64 and complex tiles, 32 for FP32
// 128 x 128 tiles (same LDS bytes as FP64 at 16; half the barriers per K for
// MFMAs that take half as long: 9216^3 125.8 -> 129.3 TFLOP/s).

// W scalars moved as one access (W*sizeof(S) is 16 bytes when vectorised).
template <typename S, int W>
struct Vec {
  S v[W];
};
template <typename S, int W>
__device__ inline Vec<S, W> ld(const S *p) {
  Vec<S, W> r;
  if constexpr (sizeof(S) * W == 16) {
    const uint4 u = *reinterpret_cast<const uint4 *>(p);
    __builtin_memcpy(&r, &u, 16);
  } else {
#pragma unroll
    for (int i = 0; i < W; ++i) r.v[i] = p[i];
  }
  return r;
}

// Tile t of C -> (bx, by): tiles are walked in groups of 8 tile-rows, column
// by column inside a group (a bijection of [0, gx*gy)).
__device__ inline void tile_xy(int64_t t, int64_t gx, int64_t gy, int64_t &bx, int64_t &by) {
  constexpr int64_t G = 8;  // tile-rows per group
  const int64_t group = t / (G * gy);
  const int64_t first = group * G;
  const int64_t rows = min(G, gx - first);
  const int64_t in = t - group * G * gy;
  bx = first + in % rows;
  by = in / rows;
}

// Tile of this workgroup among the `count` tiles starting at t0, renumbered
// so that consecutive tiles of one XCD are neighbours: XCD x (= launch index
// % 8) gets the x-th contiguous slice of them.
__device__ inline void tile_of(int64_t gx, int64_t gy, int64_t t0, int64_t count, int64_t &bx, int64_t &by) {
  const int64_t id = (int64_t)blockIdx.x;
  int64_t t = id;
  if (count >= kXcds * 4 && count % kXcds == 0) t = (id % kXcds) * (count / kXcds) + id / kXcds;
  tile_xy(t0 + t, gx, gy, bx, by);
}

// layout codes: 0 = N, 1 = T; CJA/CJB conjugate the complex operand; W is the
// vector width in scalars along each operand's contiguous axis.
template <typename S, int NC, int OPA, int OPB, int CJA, int CJB, int BM, int BN, int W, int NW, int KB>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 2))) void gemm_kernel(
                                                        int64_t M, int64_t N, int64_t K, const S *__restrict__ A,
                                                        int64_t lda, const S *__restrict__ B, int64_t ldb,
                                                        S *__restrict__ C, int64_t ldc, S alpha_r, S alpha_i,
                                                        int beta_one, int64_t gx, int64_t gy, int64_t bsa,
                                                        int64_t bsb, int64_t bsc, int64_t t0, int64_t count,
                                                        int ksplit, int64_t kchunk, S *__restrict__ part,
                                                        const int64_t *__restrict__ cmap) {
  // ksplit == 1: tiles [t0, t0 + count) of C, the whole K each.  ksplit > 1:
  // workgroup q*ksplit + c computes K range [c*kchunk, (c+1)*kchunk) of tile
  // t0 + q into its own BM x BN block of `part` (no alpha, no beta); a second
  // kernel sums the ksplit blocks of each tile in order.  cmap (batch 1 only):
  // element (m, n) of C is at C[cmap[m] + cmap[M + n]] instead of C[m + n*ldc].
  using MF = Mfma<S>;
  // batch entry blockIdx.y: its operands start bsa, bsb, bsc scalars further on
  A += (int64_t)blockIdx.y * bsa;
  B += (int64_t)blockIdx.y * bsb;
  C += (int64_t)blockIdx.y * bsc;
  using acc_t = typename MF::acc_t;
  // threads; columns of C per wave (waves form 2 rows x NW/2 columns)
  constexpr int NT = 64 * NW, WBN = BN / (NW / 2);
  constexpr int SM = BM / 32, SN = WBN / 16;  // 16x16 tiles per wave along M, N
  constexpr int PAD = W;                     // keeps rows 16-byte aligned
  constexpr int RS = KB + PAD;               // LDS row stride (scalars)
  // A tile: BM x KB elements (BM*KB*NC scalars) moved in vectors of W scalars
  // along the contiguous axis (m for 'N', k for 'T'; re/im interleaved).
  constexpr int ACONT = (OPA == 0 ? BM : KB) * NC, BCONT = (OPB == 0 ? KB : BN) * NC;
  constexpr int AVEC = BM * KB * NC / W / NT, BVEC = BN * KB * NC / W / NT;
  static_assert(AVEC >= 1 && BVEC >= 1, "tile too small");
  constexpr int ATPL = ACONT / W, BTPL = BCONT / W;  // threads per contiguous line
  static_assert(NT % ATPL == 0 && NT % BTPL == 0, "line split");
  __shared__ S As[2][NC][BM][RS];
  __shared__ S Bs[2][NC][BN][RS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int64_t bx, by, kbeg = 0, kend = K;
  if (ksplit > 1) {
    tile_xy(t0 + (int64_t)blockIdx.x / ksplit, gx, gy, bx, by);
    kbeg = (int64_t)(blockIdx.x % ksplit) * kchunk;
    kend = min(K, kbeg + kchunk);
  } else {
    tile_of(gx, gy, t0, count, bx, by);
  }
  const int64_t m0 = bx * BM, n0 = by * BN;

  // per-thread load coordinates: contiguous-axis scalar offset and line index
  const int a_c = (tid % ATPL) * W, a_l = tid / ATPL;
  const int b_c = (tid % BTPL) * W, b_l = tid / BTPL;
  constexpr int ALS = NT / ATPL, BLS = NT / BTPL;  // line step per vector

  Vec<S, W> ra[AVEC], rb[BVEC];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < AVEC; ++e) {
      const int line = a_l + e * ALS;
      int64_t m, k;
      if (OPA == 0) { m = m0 + a_c / NC; k = k0 + line; }
      else { m = m0 + line; k = k0 + a_c / NC; }
      const bool ok = (m < M) && (k < kend);
      const int64_t off = (OPA == 0) ? (m + k * lda) : (k + m * lda);
      if (ok) {
        ra[e] = ld<S, W>(A + off * NC + (a_c % NC));
      } else {
#pragma unroll
        for (int i = 0; i < W; ++i) ra[e].v[i] = (S)0;
      }
    }
#pragma unroll
    for (int e = 0; e < BVEC; ++e) {
      const int line = b_l + e * BLS;
      int64_t n, k;
      if (OPB == 0) { k = k0 + b_c / NC; n = n0 + line; }
      else { k = k0 + line; n = n0 + b_c / NC; }
      const bool ok = (n < N) && (k < kend);
      const int64_t off = (OPB == 0) ? (k + n * ldb) : (n + k * ldb);
      if (ok) {
        rb[e] = ld<S, W>(B + off * NC + (b_c % NC));
      } else {
#pragma unroll
        for (int i = 0; i < W; ++i) rb[e].v[i] = (S)0;
      }
    }
  };
  // scatter register vectors into the K-contiguous LDS layout [c][row][k]
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AVEC; ++e) {
      const int line = a_l + e * ALS;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int sc = a_c + i;  // scalar index along the contiguous axis
        const int c = sc % NC, x = sc / NC;
        if (OPA == 0) As[buf][c][x][line] = ra[e].v[i];
        else As[buf][c][line][x] = ra[e].v[i];
      }
    }
#pragma unroll
    for (int e = 0; e < BVEC; ++e) {
      const int line = b_l + e * BLS;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int sc = b_c + i;
        const int c = sc % NC, x = sc / NC;
        if (OPB == 0) Bs[buf][c][line][x] = rb[e].v[i];
        else Bs[buf][c][x][line] = rb[e].v[i];
      }
    }
  };

  acc_t acc[NC][SM][SN];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) acc[c][i][j] = acc_t{0, 0, 0, 0};

  const S sa = CJA ? (S)-1 : (S)1;  // sign of A's imaginary part
  const S sb = CJB ? (S)-1 : (S)1;
  const int64_t nk = (kend - kbeg + KB - 1) / KB;

  load_tiles(kbeg);
  store_tiles(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) load_tiles(kbeg + (kt + 1) * KB);
    // interleave the LDS fragment reads with the MFMAs (+2 % on FP64, FP32 and complex)
    __builtin_amdgcn_iglp_opt(1);
#pragma unroll
    for (int ks = 0; ks < KB / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      S fa[NC][SM], fb[NC][SN];
#pragma unroll
      for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c) fa[c][i] = As[buf][c][wm * (BM / 2) + i * 16 + (lane & 15)][kk];
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int c = 0; c < NC; ++c) fb[c][j] = Bs[buf][c][wn * WBN + j * 16 + (lane & 15)][kk];
#pragma unroll
      for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN; ++j) {
          if (NC == 1) {
            acc[0][i][j] = MF::op(fa[0][i], fb[0][j], acc[0][i][j]);
          } else {
            const S ai = sa * fa[NC - 1][i], bi = sb * fb[NC - 1][j];
            acc[0][i][j] = MF::op(fa[0][i], fb[0][j], acc[0][i][j]);
            acc[0][i][j] = MF::op(-ai, bi, acc[0][i][j]);
            acc[NC - 1][i][j] = MF::op(fa[0][i], bi, acc[NC - 1][i][j]);
            acc[NC - 1][i][j] = MF::op(ai, fb[0][j], acc[NC - 1][i][j]);
          }
        }
    }
    if (kt + 1 < nk) store_tiles(buf ^ 1);
    __syncthreads();
  }

  // epilogue
  if (ksplit > 1) {  // raw partial sums of this K range, BM x BN block column-major
    S *pb = part + (int64_t)blockIdx.x * BM * BN * NC;
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ml = wm * (BM / 2) + i * 16 + MF::row(lane, r), nl = wn * WBN + j * 16 + (lane & 15);
#pragma unroll
          for (int c = 0; c < NC; ++c) pb[(ml + nl * BM) * NC + c] = acc[c][i][j][r];
        }
    return;
  }
  // Short-K GEMMs (and GEMMs into plain column-major C) are store-bound:
  // there each 16 x 16 block of the wave's accumulators is turned round
  // through a wave-private LDS patch (the operand tiles' space, free after
  // the last barrier of the K loop) so that every store covers 16
  // consecutive rows (m, C's and D's contiguous axis) of 4 columns instead of
  // 4 rows, or 4 rows 4 apart, of 16 columns: runs of 128 bytes (FP64) rather
  // than 32.  Rank-6 contractions with K = 24 writing D in place ran 10-12 ->
  // 14-18 TFLOP/s, and 18.5-23 with the offsets loaded ahead of the stores;
  // a 5136^3 GEMM 47.3 -> 50.0; long-K GEMMs writing D in
  // place lost 0.6-1 % with it and keep the direct stores.  LDS accesses of
  // one wave complete in order, so the patch needs no barrier between its
  // writes and reads.
  if (!cmap || K <= 2048) {
    constexpr int TP = 17;  // patch column stride (scalars)
    static_assert(sizeof(As) >= sizeof(S) * NW * NC * 16 * TP, "patches fit the A tiles' space");
    S *patch = &As[0][0][0][0] + wave * NC * 16 * TP;
    const int tl = lane & 15, tg = lane >> 4;
    // the lane's row and column offsets into C, all loaded before the first
    // store (one table latency instead of one per store)
    int64_t roff[SM], coff[SN][4];
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      const int64_t m = m0 + wm * (BM / 2) + i * 16 + tl;
      roff[i] = m < M ? (cmap ? cmap[m] : m) : -1;
    }
#pragma unroll
    for (int j = 0; j < SN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t n = n0 + wn * WBN + j * 16 + 4 * r + tg;
        coff[j][r] = n < N ? (cmap ? cmap[M + n] : n * ldc) : -1;
      }
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) {
        S v[NC][4];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
          for (int r = 0; r < 4; ++r) patch[(c * 16 + (lane & 15)) * TP + MF::row(lane, r)] = acc[c][i][j][r];
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int r = 0; r < 4; ++r) v[c][r] = patch[(c * 16 + 4 * r + tg) * TP + tl];
          __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (roff[i] >= 0 && coff[j][r] >= 0) {
            S *c = C + (roff[i] + coff[j][r]) * NC;
            if (NC == 1) {
              const S x = alpha_r * v[0][r];
              c[0] = beta_one ? c[0] + x : x;
            } else {
              const S xr = v[0][r], xi = v[NC - 1][r];
              const S vr = alpha_r * xr - alpha_i * xi, vi = alpha_r * xi + alpha_i * xr;
              c[0] = beta_one ? c[0] + vr : vr;
              c[NC - 1] = beta_one ? c[NC - 1] + vi : vi;
            }
          }
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm * (BM / 2) + i * 16 + MF::row(lane, r);
        const int64_t n = n0 + wn * WBN + j * 16 + (lane & 15);
        if (m < M && n < N) {
          S *c = C + (cmap ? cmap[m] + cmap[M + n] : m + n * ldc) * NC;
          if (NC == 1) {
            const S v = alpha_r * acc[0][i][j][r];
            c[0] = beta_one ? c[0] + v : v;
          } else {
            const S xr = acc[0][i][j][r], xi = acc[NC - 1][i][j][r];
            const S vr = alpha_r * xr - alpha_i * xi, vi = alpha_r * xi + alpha_i * xr;
            c[0] = beta_one ? c[0] + vr : vr;
            c[NC - 1] = beta_one ? c[NC - 1] + vi : vi;
          }
        }
      }
}

// Second pass of a K-split tile range: sums the ksplit partial blocks of
// tiles [t0, t0 + ntiles) in order and applies C = alpha * sum (+ C).
template <typename S, int NC, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_tail_reduce(int64_t M, int64_t N, const S *__restrict__ part,
                                                             int ksplit, S *__restrict__ C, int64_t ldc, S alpha_r,
                                                             S alpha_i, int beta_one, int64_t gx, int64_t gy,
                                                             int64_t t0, int64_t ntiles,
                                                             const int64_t *__restrict__ cmap) {
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= ntiles * BM * BN) return;
  const int64_t q = e / (BM * BN);
  const int rem = (int)(e % (BM * BN)), ml = rem % BM, nl = rem / BM;
  int64_t bx, by;
  tile_xy(t0 + q, gx, gy, bx, by);
  const int64_t m = bx * BM + ml, n = by * BN + nl;
  if (m >= M || n >= N) return;
  S x[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) x[c] = (S)0;
  for (int ch = 0; ch < ksplit; ++ch) {
    const S *pb = part + ((q * ksplit + ch) * BM * BN + ml + (int64_t)nl * BM) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] += pb[c];
  }
  S *cp = C + (cmap ? cmap[m] + cmap[M + n] : m + n * ldc) * NC;
  if (NC == 1) {
    const S v = alpha_r * x[0];
    cp[0] = beta_one ? cp[0] + v : v;
  } else {
    const S vr = alpha_r * x[0] - alpha_i * x[NC - 1], vi = alpha_r * x[NC - 1] + alpha_i * x[0];
    cp[0] = beta_one ? cp[0] + vr : vr;
    cp[NC - 1] = beta_one ? cp[NC - 1] + vi : vi;
  }
}

// ---------------------------------------------------------------------------
// R4 GEMM as 3-term split-BF16 products (bf16x6): every FP32 operand value x
// is split exactly into x = h + m + l with h = bf16(x), m = bf16(x - h),
// l = bf16(x - h - m) (each residual is exact in FP32), and
//   a*b ~ ah*bh + ah*bm + am*bh + ah*bl + am*bm + al*bh
// is accumulated in FP32 by v_mfma_f32_16x16x32_bf16 (method of Henry, Tang,
// Heinecke, "Leveraging the bfloat16 Artificial Intelligence Datatype For
// Higher-Precision Computations", ARITH 2019).  The dropped terms are below
// 2^-24 relative; with FP32 accumulation the normwise error at K = 9216 is a
// few 1e-7 (cf. ~1e-6 for the FP32 MFMA path).  Two kernels implement it:
//  - gemm_bf16x6_kernel (launch_gemm with R4_SPLIT_BF16, no workspace):
//    128 x 128 tiles; FP32 tiles come from HBM as for the FP32 kernel and the
//    split happens once per element when the tile is written to LDS (three
//    BF16 planes, rows padded to 80 bytes so the 16-byte fragment reads of 16
//    lanes hit all 64 banks once);
//  - gemm_bf16x6p_kernel (launch_gemm_split, used by talshTensorContract):
//    both operands are split once into swizzled BF16 planes in a workspace by
//    split_planes_kernel, then 256 x 128 tiles of 8 waves stream the planes
//    with no VALU split work in the K loop; the last, mostly idle round of
//    tiles is split along K into partials summed by gemm_tail_reduce.
// MFMA 16x16x32 bf16
// layout: A lane l holds A[l%16][8*(l/16)..+7], B lane l holds
// B[8*(l/16)..+7][l%16], accumulator register r of lane l holds
// C[4*(l/16)+r][l%16].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)x;
  const float r = x - (float)h;
  m = (__bf16)r;
  l = (__bf16)(r - (float)m);
}

// Block tile BM x BN with WM x WN waves (each wave (BM/WM) x (BN/WN)); BK = 32
// (one MFMA depth).  Per K step each wave sums the five small products of a
// 16x16 tile into a zero-started register and the dominant ah*bh product
// into another, then adds both to the running FP32 accumulator, so the
// accumulator is rounded once per K step instead of six times.
template <int OPA, int OPB, int W, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2))) void gemm_bf16x6_kernel(
    int64_t M, int64_t N, int64_t K, const float *__restrict__ A, int64_t lda, const float *__restrict__ B,
    int64_t ldb, float *__restrict__ C, int64_t ldc, float alpha, int beta_one, int64_t gx, int64_t gy) {
  constexpr int NT = 64 * WM * WN, BKS = 32, RS = BKS + 8;
  constexpr int TM = BM / WM, TN = BN / WN, SM = TM / 16, SN = TN / 16;
  constexpr int ACONT = (OPA == 0 ? BM : BKS), BCONT = (OPB == 0 ? BKS : BN);
  constexpr int AVEC = BM * BKS / W / NT, BVEC = BN * BKS / W / NT;
  constexpr int ATPL = ACONT / W, BTPL = BCONT / W;
  constexpr int ALS = NT / ATPL, BLS = NT / BTPL;
  static_assert(AVEC >= 1 && BVEC >= 1 && NT % ATPL == 0 && NT % BTPL == 0, "tile split");
  __shared__ __attribute__((aligned(16))) __bf16 As[3][BM][RS];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][BN][RS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int64_t bx, by;
  tile_of(gx, gy, 0, gx * gy, bx, by);
  const int64_t m0 = bx * BM, n0 = by * BN;
  const int a_c = (tid % ATPL) * W, a_l = tid / ATPL;
  const int b_c = (tid % BTPL) * W, b_l = tid / BTPL;
  Vec<float, W> ra[AVEC], rb[BVEC];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < AVEC; ++e) {
      const int line = a_l + e * ALS;
      const int64_t m = OPA == 0 ? m0 + a_c : m0 + line, k = OPA == 0 ? k0 + line : k0 + a_c;
      const int64_t off = OPA == 0 ? (m + k * lda) : (k + m * lda);
      if (m < M && k < K) {
        ra[e] = ld<float, W>(A + off);
      } else {
#pragma unroll
        for (int i = 0; i < W; ++i) ra[e].v[i] = 0.0f;
      }
    }
#pragma unroll
    for (int e = 0; e < BVEC; ++e) {
      const int line = b_l + e * BLS;
      const int64_t n = OPB == 0 ? n0 + line : n0 + b_c, k = OPB == 0 ? k0 + b_c : k0 + line;
      const int64_t off = OPB == 0 ? (k + n * ldb) : (n + k * ldb);
      if (n < N && k < K) {
        rb[e] = ld<float, W>(B + off);
      } else {
#pragma unroll
        for (int i = 0; i < W; ++i) rb[e].v[i] = 0.0f;
      }
    }
  };
  // split each value into three BF16 planes; along K the W terms of a plane
  // are contiguous and go out as one packed store
  auto put = [&](__bf16 (*P)[BM > BN ? BM : BN][RS], const Vec<float, W> &v, int row, int col, bool kcont) {
    __bf16 h[W], m[W], l[W];
#pragma unroll
    for (int i = 0; i < W; ++i) split3(v.v[i], h[i], m[i], l[i]);
    if (kcont && W == 4) {
      uint2 u;
      __builtin_memcpy(&u, h, 8);
      *reinterpret_cast<uint2 *>(&P[0][row][col]) = u;
      __builtin_memcpy(&u, m, 8);
      *reinterpret_cast<uint2 *>(&P[1][row][col]) = u;
      __builtin_memcpy(&u, l, 8);
      *reinterpret_cast<uint2 *>(&P[2][row][col]) = u;
    } else {
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int r = kcont ? row : row + i, c = kcont ? col + i : col;
        P[0][r][c] = h[i];
        P[1][r][c] = m[i];
        P[2][r][c] = l[i];
      }
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < AVEC; ++e) {
      const int line = a_l + e * ALS;
      if (OPA == 1) put(reinterpret_cast<__bf16(*)[BM > BN ? BM : BN][RS]>(As), ra[e], line, a_c, true);
      else put(reinterpret_cast<__bf16(*)[BM > BN ? BM : BN][RS]>(As), ra[e], a_c, line, false);
    }
#pragma unroll
    for (int e = 0; e < BVEC; ++e) {
      const int line = b_l + e * BLS;
      if (OPB == 0) put(reinterpret_cast<__bf16(*)[BM > BN ? BM : BN][RS]>(Bs), rb[e], line, b_c, true);
      else put(reinterpret_cast<__bf16(*)[BM > BN ? BM : BN][RS]>(Bs), rb[e], b_c, line, false);
    }
  };

  f32x4 acc[SM][SN];
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  const int64_t nk = (K + BKS - 1) / BKS;
  const int kof = 8 * (lane >> 4);
  const f32x4 zero = {0, 0, 0, 0};
  load_tiles(0);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BKS);
    bf16x8 fb[3][SN];
#pragma unroll
    for (int j = 0; j < SN; ++j)
#pragma unroll
      for (int t = 0; t < 3; ++t)
        fb[t][j] = *reinterpret_cast<const bf16x8 *>(&Bs[t][wn * TN + j * 16 + (lane & 15)][kof]);
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      bf16x8 fa[3];
#pragma unroll
      for (int t = 0; t < 3; ++t)
        fa[t] = *reinterpret_cast<const bf16x8 *>(&As[t][wm * TM + i * 16 + (lane & 15)][kof]);
#pragma unroll
      for (int j = 0; j < SN; ++j) {
        f32x4 sm = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[2], fb[0][j], zero, 0, 0, 0);
        sm = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[1][j], sm, 0, 0, 0);
        sm = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[2][j], sm, 0, 0, 0);
        sm = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[0][j], sm, 0, 0, 0);
        sm = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[1][j], sm, 0, 0, 0);
        const f32x4 hh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[0][j], zero, 0, 0, 0);
        acc[i][j] += hh + sm;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm * TM + i * 16 + 4 * (lane >> 4) + r;
        const int64_t n = n0 + wn * TN + j * 16 + (lane & 15);
        if (m < M && n < N) {
          float *c = C + m + n * ldc;
          const float v = alpha * acc[i][j][r];
          *c = beta_one ? *c + v : v;
        }
      }
}

// ---------------------------------------------------------------------------
// Pre-split GEMMs.  split_planes_kernel writes the planes of an operand once
// into a workspace (row r, column k of a plane at r*Kp + k; rows padded to the
// GEMM tile, K padded to 32, zeros in the padding), so the GEMM's inner loop
// only moves bytes and issues MFMAs: no split arithmetic (which the on-the-fly
// kernel above repeats for every C tile that reads an element) and no bounds
// checks.  The dominant product of the leading terms goes to one accumulator
// and the small products to a second, so the main accumulator is rounded once
// per K step of 32 without any VALU adds; the two are summed in the epilogue.
//
// Two splits share the kernel body (NP = planes per operand):
//  - NP = 3, bf16x6: three BF16 planes h, m, l (x = h + m + l exactly) and six
//    products per block, as described above;
//  - NP = 2, fp16x3: every row of op(A) and every column of op(B) is scaled
//    by a power of two 2^-e, e the exponent of the row's largest magnitude
//    (so the row lies in (-1, 1)), then split into two FP16 planes
//    h = fp16(y), m = fp16(y - h) (y - h is exact in FP32).  FP16 keeps 11
//    significant bits, so |y - h - m| <= 2^-22 |y| for normal m, and the
//    three products ah*bh + ah*bm + am*bh (v_mfma_f32_16x16x32_f16, the same
//    rate as BF16) carry about 22 bits of each operand: normwise relative
//    error a few 1e-7, like bf16x6, for half of its MFMA work.  The dropped
//    am*bm is below 2^-22 relative.  An element below 2^-24 of its row's
//    largest keeps only an absolute accuracy of 2^-25 of that largest (FP16
//    subnormals), so rows that mix magnitudes over more than about 2^10 lose
//    relative accuracy on their small elements; bf16x6, whose planes keep
//    FP32's exponent range, has no such limit.  The epilogue multiplies by
//    2^(ea + eb), which is exact.
//
// Block tile 256 x 128, 8 waves of 64 x 64 (two per SIMD), BK = 32.  LDS
// holds two stages of both operands (2 x (16 + 8) KB per plane); a plane row
// is 64 bytes (four 16-byte chunks) and chunk c of row r is stored at slot
// c ^ (((r >> 1) ^ (r >> 3)) & 3), so the 16-byte fragment reads are free of
// bank conflicts whether the LDS serves them 8 lanes over 32 banks or 16 lanes
// over 64 banks per cycle.
// The next K step is loaded into registers while the current one is multiplied;
// one barrier per K step.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ inline void split3_finite(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)x;
  if (!__builtin_isfinite(x)) {  // Inf/NaN carried by the leading term alone
    m = l = (__bf16)0.0f;
    return;
  }
  const float r = x - (float)h;
  m = (__bf16)r;
  l = (__bf16)(r - (float)m);
}

// Exponent e with max * 2^-e in [0.5, 1) for the bit pattern u of a row's
// largest magnitude; 0 for an all-zero row or one holding Inf/NaN.
__device__ inline int row_exp(uint32_t u) {
  const float v = __uint_as_float(u);
  if (!(v > 0.0f) || !__builtin_isfinite(v)) return 0;
  int e;
  (void)frexpf(v, &e);
  return e;
}

__device__ inline void split2_scaled(float x, int e, _Float16 &h, _Float16 &m) {
  const float y = ldexpf(x, -e);
  h = (_Float16)y;
  m = __builtin_isfinite(y) ? (_Float16)(y - (float)h) : (_Float16)0.0f;  // Inf/NaN: leading term alone
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t *__restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// U[r] = max over k < K of the bit pattern of |X(r, k)| (for non-negative
// floats the pattern orders as the value; NaN patterns sort above Inf), U
// zeroed beforehand.  Block = 4 waves; grid.x = row groups, grid.y = K
// segments of kseg.  KC = 1 (X(r, k) at X[r*ld + k]): one wave per row, lanes
// along k.  KC = 0 (X(r, k) at X[r + k*ld]): lanes along r, the four waves
// share the segment.
// vec (KC = 1): X 16-byte aligned and ld a multiple of 4, so whole rounds of
// 256 floats of a segment (kseg a multiple of 256) are read as float4.
__device__ inline uint32_t abs_bits(float x) { return __float_as_uint(x) & 0x7fffffffu; }

template <int KC>
__global__ __launch_bounds__(256) void row_absmax_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                         int64_t K, int64_t kseg, int vec, uint32_t *__restrict__ U) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t k0 = (int64_t)blockIdx.y * kseg, k1 = min(K, k0 + kseg);
  uint32_t m = 0;
  if (KC) {
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    if (r < R) {
      const float *x = X + r * ld;
      int64_t k = k0;
      if (vec) {
        const int64_t kv = k0 + ((k1 - k0) & ~(int64_t)255);
#pragma unroll 4
        for (int64_t q = k0 + 4 * lane; q < kv; q += 256) {
          const float4 v = *reinterpret_cast<const float4 *>(x + q);
          m = max(m, max(max(abs_bits(v.x), abs_bits(v.y)), max(abs_bits(v.z), abs_bits(v.w))));
        }
        k = kv;
      }
#pragma unroll 4
      for (k += lane; k < k1; k += 64) m = max(m, abs_bits(x[k]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if (lane == 0 && r < R && m) atomicMax(U + r, m);
  } else {
    const int64_t r = (int64_t)blockIdx.x * 64 + lane;
    if (r < R) {
#pragma unroll 8
      for (int64_t k = k0 + wave; k < k1; k += 4) m = max(m, abs_bits(X[r + k * ld]));
      if (m) atomicMax(U + r, m);
    }
  }
}

// fp16x3 planes of op(X) when its rows are K-contiguous (X(r, k) at
// X[r*ld + k]): one block per padded row finds the row's largest magnitude
// (U[r], 0 for padding rows), then splits the row, whose second read comes
// from L2, so HBM sees each element once.  vec as for row_absmax_kernel.
__global__ __launch_bounds__(256) void split_rows_f16_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                             int64_t K, int64_t Kp, uint16_t *__restrict__ P,
                                                             int64_t plane, int vec, uint32_t *__restrict__ U) {
  __shared__ uint32_t red[4];
  const int64_t r = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float *x = X + r * ld;
  uint32_t m = 0;
  if (r < R) {
    int64_t k = 0;
    if (vec) {
      const int64_t kv = K & ~(int64_t)3;
#pragma unroll 8
      for (int64_t q = 4 * tid; q < kv; q += 1024) {
        const float4 v = *reinterpret_cast<const float4 *>(x + q);
        m = max(m, max(max(abs_bits(v.x), abs_bits(v.y)), max(abs_bits(v.z), abs_bits(v.w))));
      }
      k = kv;
    }
    for (k += tid; k < K; k += 256) m = max(m, abs_bits(x[k]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if (lane == 0) red[wave] = m;
  __syncthreads();
  m = max(max(red[0], red[1]), max(red[2], red[3]));
  if (tid == 0) U[r] = m;
  const int e = row_exp(m);
#pragma unroll 4
  for (int64_t kc = tid; kc < Kp / 8; kc += 256) {
    const int64_t k0 = kc * 8;
    float v[8];
    if (vec && r < R && k0 + 8 <= K) {
      const float4 a = *reinterpret_cast<const float4 *>(x + k0);
      const float4 b = *reinterpret_cast<const float4 *>(x + k0 + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (r < R && k0 + i < K) ? x[k0 + i] : 0.0f;
    }
    _Float16 h[8], l[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) split2_scaled(v[i], e, h[i], l[i]);
    uint4 u;
    __builtin_memcpy(&u, h, 16);
    *reinterpret_cast<uint4 *>(P + r * Kp + k0) = u;
    __builtin_memcpy(&u, l, 16);
    *reinterpret_cast<uint4 *>(P + plane + r * Kp + k0) = u;
  }
}

// KC = 1: source element (r, k) at X[r*ld + k]; KC = 0: at X[r + k*ld].
// One thread per 8 consecutive k of one row (KC = 1: lanes along k, KC = 0:
// lanes along r, so source reads are coalesced either way).  NP = 3: BF16
// planes; NP = 2: FP16 planes of the row scaled by the exponent from U.
template <int KC, int NP>
__global__ __launch_bounds__(256) void split_planes_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                           int64_t K, int64_t Rp, int64_t Kp,
                                                           uint16_t *__restrict__ P, int64_t plane, int vec,
                                                           const uint32_t *__restrict__ U) {
  const int64_t kcn = Kp / 8;
  const int64_t item = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (item >= Rp * kcn) return;
  const int64_t r = KC ? item / kcn : item % Rp, kc = KC ? item % kcn : item / Rp;
  const int64_t k0 = kc * 8;
  float v[8];
  if (KC && vec && r < R && k0 + 8 <= K) {
    const float4 a = *reinterpret_cast<const float4 *>(X + r * ld + k0);
    const float4 b = *reinterpret_cast<const float4 *>(X + r * ld + k0 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = k0 + i;
      v[i] = (r < R && k < K) ? (KC ? X[r * ld + k] : X[r + k * ld]) : 0.0f;
    }
  }
  uint4 u;
  const int64_t off = r * Kp + k0;
  if constexpr (NP == 3) {
    __bf16 h[8], m[8], l[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) split3_finite(v[i], h[i], m[i], l[i]);
    __builtin_memcpy(&u, h, 16);
    *reinterpret_cast<uint4 *>(P + off) = u;
    __builtin_memcpy(&u, m, 16);
    *reinterpret_cast<uint4 *>(P + plane + off) = u;
    __builtin_memcpy(&u, l, 16);
    *reinterpret_cast<uint4 *>(P + 2 * plane + off) = u;
  } else {
    const int e = r < R ? row_exp(U[r]) : 0;
    _Float16 h[8], m[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) split2_scaled(v[i], e, h[i], m[i]);
    __builtin_memcpy(&u, h, 16);
    *reinterpret_cast<uint4 *>(P + off) = u;
    __builtin_memcpy(&u, m, 16);
    *reinterpret_cast<uint4 *>(P + plane + off) = u;
  }
}

constexpr int PBM = 256, PBN = 128, PBK = 32;

__device__ inline int pswz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
__device__ inline int pslot(int r, int c) { return r * 4 + (c ^ pswz(r)); }

template <int NP>
__device__ inline f32x4 pmfma(uint4 a, uint4 b, f32x4 c) {
  if constexpr (NP == 3)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                  0, 0);
}

// aplane / bplane / row strides in 16-byte chunks.  ksplit > 1: workgroup
// q*ksplit + c computes K steps [c*ksteps, (c+1)*ksteps) of tile t0 + q into
// its own PBM x PBN block of `part` (big + small, scaled by 2^(ea + eb) for
// NP = 2, no alpha, no beta).  UA / UB (NP = 2): row maxima of op(A) and
// column maxima of op(B), Mp and Np entries.
template <int NP>
__device__ __forceinline__ void presplit_gemm(int64_t M, int64_t N, int64_t Kp, const uint4 *__restrict__ Ap,
                                              int64_t aplane, const uint4 *__restrict__ Bp, int64_t bplane,
                                              float *__restrict__ C, int64_t ldc, float alpha, int beta_one,
                                              int64_t gx, int64_t gy, int64_t t0, int64_t count, int ksplit,
                                              int64_t ksteps, float *__restrict__ part,
                                              const uint32_t *__restrict__ UA, const uint32_t *__restrict__ UB,
                                              const int64_t *__restrict__ cmap) {
  __shared__ uint4 As[2][NP][PBM * 4];
  __shared__ uint4 Bs[2][NP][PBN * 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  int64_t bx, by, kt0 = 0, kt1 = Kp / PBK;
  if (ksplit > 1) {
    tile_xy(t0 + (int64_t)blockIdx.x / ksplit, gx, gy, bx, by);
    kt0 = (int64_t)(blockIdx.x % ksplit) * ksteps;
    kt1 = min(kt1, kt0 + ksteps);
  } else {
    tile_of(gx, gy, t0, count, bx, by);
  }
  const int64_t kpc = Kp / 8;  // chunks per plane row
  // thread tid loads chunk (tid & 3) of row (tid >> 2) [+128] of each plane
  const int lr = tid >> 2, lc = tid & 3;
  // block-uniform bases (scalar registers) plus one 32-bit per-thread byte offset
  const char *ag = reinterpret_cast<const char *>(Ap + bx * PBM * kpc);
  const char *bg = reinterpret_cast<const char *>(Bp + by * PBN * kpc);
  const uint32_t voff = (uint32_t)((lr * kpc + lc) * 16);
  const int ls = pslot(lr, lc);
  const int64_t a1 = 128 * kpc;
  // prefetch registers: A planes t rows lr and lr+128; B planes t row lr
  uint4 a0, a1r, a2, a3, a4, a5, b0, b1, b2;
#define TAL_PG(base, chunk) (*reinterpret_cast<const uint4 *>((base) + (chunk) * 16 + voff))
#define TAL_PLOAD(kc)                                                                  \
  a0 = TAL_PG(ag, (kc));                                                               \
  a1r = TAL_PG(ag, a1 + (kc));                                                         \
  a2 = TAL_PG(ag, aplane + (kc));                                                      \
  a3 = TAL_PG(ag, aplane + a1 + (kc));                                                 \
  if constexpr (NP == 3) {                                                             \
    a4 = TAL_PG(ag, 2 * aplane + (kc));                                                \
    a5 = TAL_PG(ag, 2 * aplane + a1 + (kc));                                           \
  }                                                                                    \
  b0 = TAL_PG(bg, (kc));                                                               \
  b1 = TAL_PG(bg, bplane + (kc));                                                      \
  if constexpr (NP == 3) b2 = TAL_PG(bg, 2 * bplane + (kc));
#define TAL_PSTORE(st)                                                                 \
  As[st][0][ls] = a0;                                                                  \
  As[st][0][ls + 512] = a1r;                                                           \
  As[st][1][ls] = a2;                                                                  \
  As[st][1][ls + 512] = a3;                                                            \
  if constexpr (NP == 3) {                                                             \
    As[st][NP - 1][ls] = a4;                                                           \
    As[st][NP - 1][ls + 512] = a5;                                                     \
  }                                                                                    \
  Bs[st][0][ls] = b0;                                                                  \
  Bs[st][1][ls] = b1;                                                                  \
  if constexpr (NP == 3) Bs[st][NP - 1][ls] = b2;

  f32x4 big[4][4], small[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) big[i][j] = small[i][j] = f32x4{0, 0, 0, 0};

  // fragment slot of row 16*x + (lane & 15), chunk lane >> 4: the swizzle of
  // that row is pswz(lane & 15), with bit 1 flipped when x is odd
  const int fs[2] = {pslot(lane & 15, lane >> 4), (lane & 15) * 4 + ((lane >> 4) ^ pswz(lane & 15) ^ 2)};
  const int64_t nk = kt1 - kt0;  // may be <= 0 for a trailing K range: its partial sums are zeros
  if (nk > 0) {
    TAL_PLOAD(kt0 * (PBK / 8))
    TAL_PSTORE(0)
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt & 1);
    if (kt + 1 < nk) {
      TAL_PLOAD((kt0 + kt + 1) * (PBK / 8))
    }
    uint4 fb[NP][4];
#pragma unroll
    for (int t = 0; t < NP; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) fb[t][j] = Bs[cur][t][(wn * 64 + j * 16) * 4 + fs[j & 1]];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint4 fa[NP];
#pragma unroll
      for (int t = 0; t < NP; ++t) fa[t] = As[cur][t][(wm * 64 + i * 16) * 4 + fs[i & 1]];
      if constexpr (NP == 3) {
#pragma unroll
        for (int j = 0; j < 4; ++j) small[i][j] = pmfma<NP>(fa[2], fb[0][j], small[i][j]);
#pragma unroll
        for (int j = 0; j < 4; ++j) small[i][j] = pmfma<NP>(fa[1], fb[1][j], small[i][j]);
#pragma unroll
        for (int j = 0; j < 4; ++j) small[i][j] = pmfma<NP>(fa[0], fb[2][j], small[i][j]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) small[i][j] = pmfma<NP>(fa[1], fb[0][j], small[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) small[i][j] = pmfma<NP>(fa[0], fb[1][j], small[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) big[i][j] = pmfma<NP>(fa[0], fb[0][j], big[i][j]);
    }
    if (kt + 1 < nk) {
      TAL_PSTORE(cur ^ 1)
    }
    __syncthreads();
  }
#undef TAL_PG
#undef TAL_PLOAD
#undef TAL_PSTORE
  // NP = 2: exponents of this lane's 16 rows (i, r) and 4 columns (j); every
  // index is below Mp / Np, the sizes of UA / UB
  int ea[4][4] = {}, eb[4] = {};
  if constexpr (NP == 2) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) ea[i][r] = row_exp(UA[bx * PBM + wm * 64 + i * 16 + 4 * (lane >> 4) + r]);
#pragma unroll
    for (int j = 0; j < 4; ++j) eb[j] = row_exp(UB[by * PBN + wn * 64 + j * 16 + (lane & 15)]);
  }
  if (ksplit > 1) {
    float *pb = part + (int64_t)blockIdx.x * PBM * PBN;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ml = wm * 64 + i * 16 + 4 * (lane >> 4) + r, nl = wn * 64 + j * 16 + (lane & 15);
          const float s = big[i][j][r] + small[i][j][r];
          pb[ml + nl * PBM] = NP == 2 ? ldexpf(s, ea[i][r] + eb[j]) : s;
        }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = bx * PBM + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
        const int64_t n = by * PBN + wn * 64 + j * 16 + (lane & 15);
        if (m < M && n < N) {
          float *c = C + (cmap ? cmap[m] + cmap[M + n] : m + n * ldc);
          const float s = big[i][j][r] + small[i][j][r];
          const float v = alpha * (NP == 2 ? ldexpf(s, ea[i][r] + eb[j]) : s);
          *c = beta_one ? *c + v : v;
        }
      }
}

__global__ __launch_bounds__(512) void gemm_bf16x6p_kernel(int64_t M, int64_t N, int64_t Kp,
                                                           const uint4 *__restrict__ Ap, int64_t aplane,
                                                           const uint4 *__restrict__ Bp, int64_t bplane,
                                                           float *__restrict__ C, int64_t ldc, float alpha,
                                                           int beta_one, int64_t gx, int64_t gy, int64_t t0,
                                                           int64_t count, int ksplit, int64_t ksteps,
                                                           float *__restrict__ part, const int64_t *__restrict__ cmap) {
  presplit_gemm<3>(M, N, Kp, Ap, aplane, Bp, bplane, C, ldc, alpha, beta_one, gx, gy, t0, count, ksplit, ksteps,
                   part, nullptr, nullptr, cmap);
}

__global__ __launch_bounds__(512) void gemm_f16x3p_kernel(int64_t M, int64_t N, int64_t Kp,
                                                          const uint4 *__restrict__ Ap, int64_t aplane,
                                                          const uint4 *__restrict__ Bp, int64_t bplane,
                                                          float *__restrict__ C, int64_t ldc, float alpha,
                                                          int beta_one, int64_t gx, int64_t gy, int64_t t0,
                                                          int64_t count, int ksplit, int64_t ksteps,
                                                          float *__restrict__ part, const uint32_t *__restrict__ UA,
                                                          const uint32_t *__restrict__ UB,
                                                          const int64_t *__restrict__ cmap) {
  presplit_gemm<2>(M, N, Kp, Ap, aplane, Bp, bplane, C, ldc, alpha, beta_one, gx, gy, t0, count, ksplit, ksteps,
                   part, UA, UB, cmap);
}

int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

__global__ __launch_bounds__(256) void offset_map_kernel(OffsetMap map, int64_t *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= map.M + map.N) return;
  const bool row = e < map.M;
  int64_t x = row ? e : e - map.M, off = 0;
  const int t0 = row ? 0 : map.nm, t1 = row ? map.nm : map.nm + map.nn;
  for (int t = t0; t < t1; ++t) {
    off += (x % map.ext[t]) * map.str[t];
    x /= map.ext[t];
  }
  out[e] = off;
}

// Splits op(X) (rows R, depth K) into NP planes at P (NP = 2: also the row
// maxima into U, Rp entries, used for the scaling); returns 0 or a launch error.
template <int NP>
int split_operand(bool kcont, const float *X, int64_t ld, int64_t R, int64_t K, int64_t Rp, int64_t Kp,
                  uint16_t *P, uint32_t *U, hipStream_t s) {
  const int64_t items = Rp * (Kp / 8), nblk = (items + 255) / 256;
  if (nblk > 0x7fffffffLL) return -4;
  const int vec = ((uintptr_t)X % 16 == 0) && ld % 4 == 0;
  if (NP == 2 && kcont) {  // maxima and planes in one pass over HBM
    if (Rp > 0x7fffffffLL) return -4;
    split_rows_f16_kernel<<<dim3((unsigned)Rp), dim3(256), 0, s>>>(X, ld, R, K, Kp, P, Rp * Kp, vec, U);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  if constexpr (NP == 2) {
    fill_u32_kernel<<<dim3((unsigned)((Rp + 255) / 256)), dim3(256), 0, s>>>(U, Rp);
    if (R > 0 && K > 0) {
      const int64_t rows = kcont ? 4 : 64;
      int64_t kseg = kcont ? 2048 : 512;
      kseg = std::max(kseg, round_up((K + 65534) / 65535, 256));
      if ((R + rows - 1) / rows > 0x7fffffffLL) return -4;
      const dim3 grid((unsigned)((R + rows - 1) / rows), (unsigned)((K + kseg - 1) / kseg));
      if (kcont)
        row_absmax_kernel<1><<<grid, dim3(256), 0, s>>>(X, ld, R, K, kseg, vec, U);
      else
        row_absmax_kernel<0><<<grid, dim3(256), 0, s>>>(X, ld, R, K, kseg, vec, U);
    }
  }
  if (kcont)
    split_planes_kernel<1, NP><<<dim3((unsigned)nblk), dim3(256), 0, s>>>(X, ld, R, K, Rp, Kp, P, Rp * Kp, vec, U);
  else
    split_planes_kernel<0, NP><<<dim3((unsigned)nblk), dim3(256), 0, s>>>(X, ld, R, K, Rp, Kp, P, Rp * Kp, vec, U);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int OPA, int OPB>
int launch_bf16x6(int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                  int64_t ldc, double ar, int beta_one, hipStream_t s) {
  const int64_t acont = OPA == 0 ? M : K, bcont = OPB == 0 ? K : N;
  const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && acont % 4 == 0 && bcont % 4 == 0 &&
                   lda % 4 == 0 && ldb % 4 == 0;
  if (vec) {  // 128 x 128 tiles, 4 waves of 64 x 64
    const int64_t gx = (M + 127) / 128, gy = (N + 127) / 128, nblk = gx * gy;
    if (nblk > 0x7fffffffLL) return -4;
    gemm_bf16x6_kernel<OPA, OPB, 4, 128, 128, 2, 2><<<dim3((unsigned)nblk), dim3(256), 0, s>>>(
        M, N, K, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, (float)ar, beta_one, gx, gy);
  } else {  // scalar accesses: 128 x 128 tiles, 4 waves of 64 x 64
    const int64_t gx = (M + 127) / 128, gy = (N + 127) / 128, nblk = gx * gy;
    if (nblk > 0x7fffffffLL) return -4;
    gemm_bf16x6_kernel<OPA, OPB, 1, 128, 128, 2, 2><<<dim3((unsigned)nblk), dim3(256), 0, s>>>(
        M, N, K, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, (float)ar, beta_one, gx, gy);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Workgroups of gemm_kernel<S, NC, ..., BM, BN> resident at once on the
// current device (CUs x blocks per CU), cached per device.
template <typename S, int NC, int BM, int BN, int NWS, int KB>
int64_t gemm_slots() {
  static int64_t cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!cache[dev]) {
    int cus = 0, per = 0;
    constexpr int WV = (int)(16 / sizeof(S));
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_kernel<S, NC, 1, 0, 0, 0, BM, BN, WV, NWS, KB>,
                                                     64 * NWS, 0) !=
            hipSuccess)
      return 0;
    cache[dev] = (int64_t)cus * std::max(per, 1);
  }
  return cache[dev];
}

// The tiles of C are dealt to the resident workgroups in rounds.  When the
// last round is at most a quarter full (or the whole GEMM is), its `rem`
// tiles are computed as ksplit K ranges each (filling the GPU) plus a
// summing pass; ksplit = 1 means no split.
struct TailSplit {
  int64_t rem = 0;
  int ksplit = 1;
};
template <typename S, int NC, int BM, int BN, int NWS, int KB>
TailSplit tail_split(int64_t M, int64_t N, int64_t K, int64_t batch) {
  TailSplit t;
  if (batch != 1 || M <= 0 || N <= 0) return t;
  const int64_t T = ((M + BM - 1) / BM) * ((N + BN - 1) / BN), slots = gemm_slots<S, NC, BM, BN, NWS, KB>();
  if (slots <= 0) return t;
  const int64_t rem = T % slots;
  if (rem == 0 || rem * 4 > slots) return t;
  // up to 64 K ranges per tile (a 312 x 312 x 97,344 product: 16 ranges ran
  // 30.7 TFLOP/s, 20 ran 36.5); the partials never exceed one round of tiles
  const int64_t ks = std::min<int64_t>(std::min<int64_t>(64, slots / rem), ((K + KB - 1) / KB) / 8);
  if (ks < 2) return t;
  t.rem = rem;
  t.ksplit = (int)ks;
  return t;
}
template <typename S, int NC, int BM, int BN>
size_t tail_bytes(const TailSplit &t) {
  return t.ksplit > 1 ? (size_t)t.rem * t.ksplit * BM * BN * NC * sizeof(S) : 0;
}

template <typename S, int NC, int OPA, int OPB, int CJA, int CJB, int BM, int BN, int NWS, int KB>
int launch_typed(int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B, int64_t ldb,
                 void *C, int64_t ldc, double ar, double ai, int beta_one, int64_t batch, int64_t sa, int64_t sb,
                 int64_t sc, void *ws, size_t ws_bytes, const int64_t *cmap, hipStream_t s) {
  const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
  const int64_t nblk = gx * gy;
  if (nblk > 0x7fffffffLL) return -4;
  // 16-byte accesses need every contiguous run (and every base, every batch entry) to split into whole vectors
  constexpr int WV = (int)(16 / sizeof(S));
  const int64_t acont = (OPA == 0 ? M : K) * NC, bcont = (OPB == 0 ? K : N) * NC;
  const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && acont % WV == 0 &&
                   bcont % WV == 0 && (lda * NC) % WV == 0 && (ldb * NC) % WV == 0 &&
                   (batch == 1 || ((sa * NC) % WV == 0 && (sb * NC) % WV == 0));
  TailSplit ts = tail_split<S, NC, BM, BN, NWS, KB>(M, N, K, batch);
  if (!ws || tail_bytes<S, NC, BM, BN>(ts) > ws_bytes) ts = TailSplit{};
  const int64_t full = nblk - ts.rem;  // tiles computed whole, in the first launch
  const int64_t kchunk = ts.ksplit > 1 ? (((K + KB - 1) / KB + ts.ksplit - 1) / ts.ksplit) * KB : 0;
  // batch entries go to grid.y, at most 65535 per launch
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const unsigned nb = (unsigned)std::min<int64_t>(65535, batch - b0);
    const S *a = (const S *)A + b0 * sa * NC, *b = (const S *)B + b0 * sb * NC;
    S *c = (S *)C + b0 * sc * NC;
    if (full > 0) {
      const dim3 grid((unsigned)full, nb);
      const int64_t bsa = sa * NC, bsb = sb * NC, bsc = sc * NC;
      if (batch > 1 && vec)
        gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, WV, kWavesBatched, KB><<<grid, dim3(64 * kWavesBatched), 0, s>>>(
            M, N, K, a, lda, b, ldb, c, ldc, (S)ar, (S)ai, beta_one, gx, gy, bsa, bsb, bsc, 0, full, 1, 0, nullptr,
            cmap);
      else if (batch > 1)
        gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, 1, kWavesBatched, KB><<<grid, dim3(64 * kWavesBatched), 0, s>>>(
            M, N, K, a, lda, b, ldb, c, ldc, (S)ar, (S)ai, beta_one, gx, gy, bsa, bsb, bsc, 0, full, 1, 0, nullptr,
            cmap);
      else if (vec)
        gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, WV, NWS, KB><<<grid, dim3(64 * NWS), 0, s>>>(
            M, N, K, a, lda, b, ldb, c, ldc, (S)ar, (S)ai, beta_one, gx, gy, bsa, bsb, bsc, 0, full, 1, 0, nullptr,
            cmap);
      else
        gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, 1, NWS, KB><<<grid, dim3(64 * NWS), 0, s>>>(
            M, N, K, a, lda, b, ldb, c, ldc, (S)ar, (S)ai, beta_one, gx, gy, bsa, bsb, bsc, 0, full, 1, 0, nullptr,
            cmap);
    }
  }
  if (ts.ksplit > 1) {
    const unsigned nw = (unsigned)(ts.rem * ts.ksplit);
    if (vec)
      gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, WV, NWS, KB><<<dim3(nw), dim3(64 * NWS), 0, s>>>(
          M, N, K, (const S *)A, lda, (const S *)B, ldb, (S *)C, ldc, (S)ar, (S)ai, beta_one, gx, gy, 0, 0, 0, full,
          ts.rem, ts.ksplit, kchunk, (S *)ws, cmap);
    else
      gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN, 1, NWS, KB><<<dim3(nw), dim3(64 * NWS), 0, s>>>(
          M, N, K, (const S *)A, lda, (const S *)B, ldb, (S *)C, ldc, (S)ar, (S)ai, beta_one, gx, gy, 0, 0, 0, full,
          ts.rem, ts.ksplit, kchunk, (S *)ws, cmap);
    const int64_t nel = ts.rem * BM * BN;
    gemm_tail_reduce<S, NC, BM, BN><<<dim3((unsigned)((nel + kThreads - 1) / kThreads)), dim3(kThreads), 0, s>>>(
        M, N, (const S *)ws, ts.ksplit, (S *)C, ldc, (S)ar, (S)ai, beta_one, gx, gy, full, ts.rem, cmap);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename S, int NC, int BM, int BN, int NWS, int KB>
int dispatch_ops(int opa, int opb, int cja, int cjb, int64_t M, int64_t N, int64_t K, const void *A, int64_t lda,
                 const void *B, int64_t ldb, void *C, int64_t ldc, double ar, double ai, int beta_one,
                 int64_t batch, int64_t sa, int64_t sb, int64_t sc, void *ws, size_t ws_bytes,
                 const int64_t *cmap, hipStream_t s) {
#define TAL_GEMM_CASE(a, b, ca, cb)                                                                      \
  if (opa == a && opb == b && cja == ca && cjb == cb)                                                    \
    return launch_typed<S, NC, a, b, ca, cb, BM, BN, NWS, KB>(M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, \
                                                              batch, sa, sb, sc, ws, ws_bytes, cmap, s);
#define TAL_GEMM_LAYOUTS(ca, cb) \
  TAL_GEMM_CASE(0, 0, ca, cb)    \
  TAL_GEMM_CASE(0, 1, ca, cb)    \
  TAL_GEMM_CASE(1, 0, ca, cb)    \
  TAL_GEMM_CASE(1, 1, ca, cb)
  if constexpr (NC == 1) {
    TAL_GEMM_LAYOUTS(0, 0)
  } else {
    TAL_GEMM_LAYOUTS(0, 0)
    TAL_GEMM_LAYOUTS(0, 1)
    TAL_GEMM_LAYOUTS(1, 0)
    TAL_GEMM_LAYOUTS(1, 1)
  }
#undef TAL_GEMM_LAYOUTS
#undef TAL_GEMM_CASE
  return -1;
}

int op_code(char t) {
  switch (t) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    default: return -1;
  }
}

}  // namespace

bool gemm_split_bf16_available() { return true; }

// Resident workgroups of the pre-split GEMM kernel with NP planes on the
// current device, cached per device.
template <int NP>
static int64_t presplit_slots() {
  static int64_t cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!cache[dev]) {
    int cus = 0, per = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = NP == 3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_bf16x6p_kernel, 512, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_f16x3p_kernel, 512, 0);
    if (e != hipSuccess) return 0;
    cache[dev] = (int64_t)cus * std::max(per, 1);
  }
  return cache[dev];
}

// K split of a mostly idle last round of tiles (as for the FP32/FP64 GEMM).
template <int NP>
static TailSplit presplit_tail(int64_t M, int64_t N, int64_t K) {
  TailSplit t;
  if (M <= 0 || N <= 0) return t;
  const int64_t T = (round_up(M, PBM) / PBM) * (round_up(N, PBN) / PBN), slots = presplit_slots<NP>();
  if (slots <= 0) return t;
  const int64_t rem = T % slots;
  if (rem == 0 || rem * 4 > slots) return t;
  const int64_t ksteps = round_up(std::max<int64_t>(K, 1), PBK) / PBK;
  const int64_t ks = std::min<int64_t>(std::min<int64_t>(16, slots / rem), ksteps / 8);
  if (ks < 2) return t;
  t.rem = rem;
  t.ksplit = (int)ks;
  return t;
}

// Workspace of the pre-split GEMM: the planes of op(A) and op(B) (NP planes
// of 2-byte values each), for NP = 2 the row / column maxima, then the
// partial sums of a K-split last round; each part starts on 256 bytes.
struct SplitLayout {
  int64_t Mp, Np, Kp;
  size_t a, b, ua, ub, part, total;
  TailSplit tail;
};

template <int NP>
static SplitLayout split_layout(int64_t M, int64_t N, int64_t K) {
  SplitLayout l;
  l.Mp = round_up(std::max<int64_t>(M, 1), PBM);
  l.Np = round_up(std::max<int64_t>(N, 1), PBN);
  l.Kp = round_up(std::max<int64_t>(K, 1), PBK);
  l.tail = presplit_tail<NP>(M, N, K);
  l.a = 0;
  l.b = l.a + (size_t)round_up(2 * NP * l.Mp * l.Kp, 256);
  l.ua = l.b + (size_t)round_up(2 * NP * l.Np * l.Kp, 256);
  l.ub = l.ua + (NP == 2 ? (size_t)round_up(4 * l.Mp, 256) : 0);
  l.part = l.ub + (NP == 2 ? (size_t)round_up(4 * l.Np, 256) : 0);
  l.total = l.part + (l.tail.ksplit > 1 ? (size_t)l.tail.rem * l.tail.ksplit * PBM * PBN * sizeof(float) : 0);
  return l;
}

template <int NP>
static int launch_presplit(int opa, int opb, int64_t M, int64_t N, int64_t K, double ar, const void *A, int64_t lda,
                           const void *B, int64_t ldb, int beta_one, void *C, int64_t ldc, void *ws,
                           const int64_t *cmap, hipStream_t s) {
  const SplitLayout l = split_layout<NP>(M, N, K);
  char *w = (char *)ws;
  uint16_t *pa = (uint16_t *)(w + l.a), *pb = (uint16_t *)(w + l.b);
  uint32_t *ua = NP == 2 ? (uint32_t *)(w + l.ua) : nullptr, *ub = NP == 2 ? (uint32_t *)(w + l.ub) : nullptr;
  float *part = (float *)(w + l.part);
  // A 'T' (KxM) and B 'N' (KxN) already have K contiguous per row of the plane
  if (int e = split_operand<NP>(opa == 1, (const float *)A, lda, M, K, l.Mp, l.Kp, pa, ua, s)) return e;
  if (int e = split_operand<NP>(opb == 0, (const float *)B, ldb, N, K, l.Np, l.Kp, pb, ub, s)) return e;
  const int64_t gx = l.Mp / PBM, gy = l.Np / PBN;
  if (gx * gy > 0x7fffffffLL) return -4;
  const TailSplit ts = l.tail;
  const int64_t full = gx * gy - ts.rem, nsteps = l.Kp / PBK;
  const int64_t ksteps = ts.ksplit > 1 ? (nsteps + ts.ksplit - 1) / ts.ksplit : 0;
  const int64_t apl = l.Mp * l.Kp / 8, bpl = l.Np * l.Kp / 8;
  auto gemm = [&](int64_t blocks, int64_t t0, int64_t count, int ksplit, int64_t kst, float *pt) {
    if (NP == 3)
      gemm_bf16x6p_kernel<<<dim3((unsigned)blocks), dim3(512), 0, s>>>(
          M, N, l.Kp, (const uint4 *)pa, apl, (const uint4 *)pb, bpl, (float *)C, ldc, (float)ar, beta_one, gx, gy,
          t0, count, ksplit, kst, pt, cmap);
    else
      gemm_f16x3p_kernel<<<dim3((unsigned)blocks), dim3(512), 0, s>>>(
          M, N, l.Kp, (const uint4 *)pa, apl, (const uint4 *)pb, bpl, (float *)C, ldc, (float)ar, beta_one, gx, gy,
          t0, count, ksplit, kst, pt, ua, ub, cmap);
  };
  if (full > 0) gemm(full, 0, full, 1, 0, nullptr);
  if (ts.ksplit > 1) {
    gemm(ts.rem * ts.ksplit, full, ts.rem, ts.ksplit, ksteps, part);
    const int64_t nel = ts.rem * PBM * PBN;
    gemm_tail_reduce<float, 1, PBM, PBN><<<dim3((unsigned)((nel + kThreads - 1) / kThreads)), dim3(kThreads), 0, s>>>(
        M, N, part, ts.ksplit, (float *)C, ldc, (float)ar, 0.0f, beta_one, gx, gy, full, ts.rem, cmap);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_offset_map(const OffsetMap &map, int64_t *out, hipStream_t s) {
  if (map.nm < 0 || map.nn < 0 || map.nm + map.nn > kMaxMapDims || !out) return -1;
  for (int t = 0; t < map.nm + map.nn; ++t)
    if (map.ext[t] < 1) return -1;
  const int64_t n = map.M + map.N, nblk = (n + 255) / 256;
  if (nblk > 0x7fffffffLL) return -4;
  offset_map_kernel<<<dim3((unsigned)nblk), dim3(256), 0, s>>>(map, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t gemm_split_workspace(int method, int64_t M, int64_t N, int64_t K) {
  return method == SPLIT_FP16X3 ? split_layout<2>(M, N, K).total : split_layout<3>(M, N, K).total;
}

int launch_gemm_split(int method, char ta, char tb, int64_t M, int64_t N, int64_t K, double ar, const void *A,
                      int64_t lda, const void *B, int64_t ldb, int beta_one, void *C, int64_t ldc, void *ws,
                      hipStream_t s, const int64_t *cmap) {
  const int opa = op_code(ta), opb = op_code(tb);
  if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0 || !A || !B || !C || !ws || (uintptr_t)ws % 16) return -1;
  if (method != SPLIT_BF16X6 && method != SPLIT_FP16X3) return -1;
  if (M == 0 || N == 0) return 0;
  if (ldc < std::max<int64_t>(1, M)) return -1;
  if (lda < std::max<int64_t>(1, opa == 0 ? M : K)) return -1;
  if (ldb < std::max<int64_t>(1, opb == 0 ? K : N)) return -1;
  return method == SPLIT_FP16X3
             ? launch_presplit<2>(opa, opb, M, N, K, ar, A, lda, B, ldb, beta_one, C, ldc, ws, cmap, s)
             : launch_presplit<3>(opa, opb, M, N, K, ar, A, lda, B, ldb, beta_one, C, ldc, ws, cmap, s);
}

// Real single GEMMs whose 128 x 128 tiles would not fill one round of the
// resident workgroups run 64 x 64 tiles: four times the tiles, so a 1024^3
// product spreads over the whole GPU.  FP64 uses 8 waves (32 x 16 of C each)
// and a K step of 32, one 256-tile round with no K split (1024^3: 54 us,
// against 62 us for 4 waves with a K step of 16 split in two plus the
// summing pass); FP32 keeps 4 waves (32 x 32 each) and a K step of 16.
template <typename S>
bool small_real(int64_t M, int64_t N, int64_t batch) {
  if (batch != 1) return false;
  const int64_t t128 = ((M + 127) / 128) * ((N + 127) / 128);
  return t128 < gemm_slots<S, 1, 128, 128, kWavesSingle, sizeof(S) == 4 ? 2 * BK : BK>();
}

// Real GEMMs with at most 32 columns and many 128-row tiles (the N = 24
// contractions of the reference's benchmark list) are latency-bound
// on their A tile loads: 128 x 32 tiles of 4 waves (64 x 16 of C each) keep
// more workgroups, and so more of A, in flight per CU than 128 x 128 tiles
// of which three quarters of the columns are empty.  Those GEMMs ran 6-10 ->
// 16-28 TFLOP/s in FP64 and 9-20 -> 20-49 in FP32 (K = 48-312, M = 97,344
// to 5,308,416); up to 96 columns, three narrow tiles still beat one wide
// one from 4,096 rows (5184 x 72 x 5184: 12.8 -> 19.2 TFLOP/s per step).
bool skinny_real(int64_t M, int64_t N, int64_t batch) {
  return batch == 1 && ((N <= 32 && M >= 128 * 512) || (N <= 96 && M >= 4096));
}

size_t gemm_tail_workspace(int dk, int64_t M, int64_t N, int64_t K, int64_t batch) {
  switch (dk) {
    case 8:
      if (skinny_real(M, N, batch))
        return tail_bytes<double, 1, 128, 32>(tail_split<double, 1, 128, 32, 4, BK>(M, N, K, batch));
      return small_real<double>(M, N, batch)
                 ? tail_bytes<double, 1, 64, 64>(tail_split<double, 1, 64, 64, kWavesSingle, 2 * BK>(M, N, K, batch))
                 : tail_bytes<double, 1, 128, 128>(tail_split<double, 1, 128, 128, kWavesSingle, BK>(M, N, K, batch));
    case 4:
      if (skinny_real(M, N, batch))
        return tail_bytes<float, 1, 128, 32>(tail_split<float, 1, 128, 32, 4, 2 * BK>(M, N, K, batch));
      return small_real<float>(M, N, batch)
                 ? tail_bytes<float, 1, 64, 64>(tail_split<float, 1, 64, 64, 4, BK>(M, N, K, batch))
                 : tail_bytes<float, 1, 128, 128>(
                       tail_split<float, 1, 128, 128, kWavesSingle, 2 * BK>(M, N, K, batch));
    case 18: return tail_bytes<double, 2, 64, 64>(tail_split<double, 2, 64, 64, kWavesSingle, BK>(M, N, K, batch));
    case 14: return tail_bytes<float, 2, 64, 64>(tail_split<float, 2, 64, 64, kWavesSingle, BK>(M, N, K, batch));
    default: return 0;
  }
}

int launch_gemm(int dk, int r4mode, char ta, char tb, int conja, int conjb, int64_t M, int64_t N, int64_t K,
                double ar, double ai, const void *A, int64_t lda, const void *B, int64_t ldb, int beta_one, void *C,
                int64_t ldc, hipStream_t s, int64_t batch, int64_t sa, int64_t sb, int64_t sc, void *ws,
                size_t ws_bytes, const int64_t *cmap) {
  const bool cplx = (dk == 14 || dk == 18);
  const int opa = op_code(ta), opb = op_code(tb);
  const int cja = (cplx && conja) ? 1 : 0, cjb = (cplx && conjb) ? 1 : 0;
  if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0 || !A || !B || !C || batch < 1) return -1;
  if (M == 0 || N == 0) return 0;
  if (ldc < std::max<int64_t>(1, M)) return -1;
  if (lda < std::max<int64_t>(1, opa == 0 ? M : K)) return -1;
  if (ldb < std::max<int64_t>(1, opb == 0 ? K : N)) return -1;
  if (batch > 1 && (sa < 0 || sb < 0 || sc < ldc * N)) return -1;  // batch entries of C may not overlap
  if (cmap && (batch > 1 || (dk == 4 && r4mode == R4_SPLIT_BF16))) return -1;
  if (dk == 4 && r4mode == R4_SPLIT_BF16) {
    for (int64_t q = 0; q < batch; ++q) {
      const float *a = (const float *)A + q * sa, *b = (const float *)B + q * sb;
      float *c = (float *)C + q * sc;
      int e;
      if (opa == 0 && opb == 0) e = launch_bf16x6<0, 0>(M, N, K, a, lda, b, ldb, c, ldc, ar, beta_one, s);
      else if (opa == 0 && opb == 1) e = launch_bf16x6<0, 1>(M, N, K, a, lda, b, ldb, c, ldc, ar, beta_one, s);
      else if (opa == 1 && opb == 0) e = launch_bf16x6<1, 0>(M, N, K, a, lda, b, ldb, c, ldc, ar, beta_one, s);
      else e = launch_bf16x6<1, 1>(M, N, K, a, lda, b, ldb, c, ldc, ar, beta_one, s);
      if (e) return e;
    }
    return 0;
  }
#define TAL_GEMM_KIND(S, NC, BM, BN, NWS, KB)                                                                    \
  dispatch_ops<S, NC, BM, BN, NWS, KB>(opa, opb, cja, cjb, M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, batch, \
                                       sa, sb, sc, ws, ws_bytes, cmap, s)
  switch (dk) {
    case 8:
      if (skinny_real(M, N, batch)) return TAL_GEMM_KIND(double, 1, 128, 32, 4, BK);
      return small_real<double>(M, N, batch) ? TAL_GEMM_KIND(double, 1, 64, 64, kWavesSingle, 2 * BK)
                                             : TAL_GEMM_KIND(double, 1, 128, 128, kWavesSingle, BK);
    case 4:
      if (skinny_real(M, N, batch)) return TAL_GEMM_KIND(float, 1, 128, 32, 4, 2 * BK);
      return small_real<float>(M, N, batch) ? TAL_GEMM_KIND(float, 1, 64, 64, 4, BK)
                                            : TAL_GEMM_KIND(float, 1, 128, 128, kWavesSingle, 2 * BK);
    case 18: return TAL_GEMM_KIND(double, 2, 64, 64, kWavesSingle, BK);
    case 14: return TAL_GEMM_KIND(float, 2, 64, 64, kWavesSingle, BK);
    default: return -1;
  }
#undef TAL_GEMM_KIND
}

}  // namespace tal
