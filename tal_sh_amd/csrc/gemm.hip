// MFMA GEMM kernels for the TTGT contraction stage (gfx950).
//
// Replaces the GEMM call of gpu_tensor_block_contract_dlf
// (tensor_algebra_gpu_nvidia.cu:6851-7060, cuBLAS/hipBLAS) and its in-house
// fallback gpu_matrix_multiply_tn__ (:2263-2312).  Contract:
//   C(MxN, column-major, ldc) = alpha * op(A) * op(B) + beta * C,  beta in {0,1}
// with op(A) in {A, A^T} and op(B) in {B, B^T}, each optionally conjugated, so that the
// contraction planner can feed tensor operands in whichever matrix form their
// storage already has.
//
// Kernel shape: 256 threads = 4 wave64s in a 2x2 grid; each wave owns a
// (BM/2)x(BN/2) slab of C built from 16x16 MFMA tiles:
//   R8/C8: v_mfma_f64_16x16x4_f64   R4/C4: v_mfma_f32_16x16x4_f32
// Operand tiles of BK=16 are staged through LDS in K-contiguous rows padded by
// one element; the next K tile is prefetched into registers while the current
// one is multiplied.  Complex products use four real MFMAs per step
// (re += ar*br - ai*bi, im += ar*bi + ai*br), with conjugation folded into the
// sign of the imaginary operand.  MFMA fragment layout (16x16x4): A lane l
// holds A[l%16][l/16], B lane l holds B[l/16][l%16]; accumulator register r of
// lane l holds C[4*(l/16)+r][l%16] for f32 and C[4*r+l/16][l%16] for f64.
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
constexpr int PAD = 1;

// layout codes: 0 = N, 1 = T; CJA/CJB conjugate the complex operand
template <typename S, int NC, int OPA, int OPB, int CJA, int CJB, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_kernel(int64_t M, int64_t N, int64_t K, const S *__restrict__ A,
                                                        int64_t lda, const S *__restrict__ B, int64_t ldb,
                                                        S *__restrict__ C, int64_t ldc, S alpha_r, S alpha_i,
                                                        int beta_one) {
  using MF = Mfma<S>;
  using acc_t = typename MF::acc_t;
  constexpr int SM = BM / 32, SN = BN / 32;  // 16x16 tiles per wave along M, N
  constexpr int EA = BM * BK / kThreads, EB = BN * BK / kThreads;
  static_assert(EA >= 1 && EB >= 1, "tile too small");
  __shared__ S As[NC][BM][BK + PAD];
  __shared__ S Bs[NC][BN][BK + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;

  // load mapping
  // A: OPA==0 (N): A(m,k) = A[m + k*lda], m contiguous; else A(m,k) = A[k + m*lda], k contiguous.
  int am, ak, am_step, ak_step;
  if (OPA == 0) { am = tid % BM; ak = tid / BM; am_step = 0; ak_step = kThreads / BM; }
  else { ak = tid % BK; am = tid / BK; ak_step = 0; am_step = kThreads / BK; }
  // B: OPB==0 (N): B(k,n) = B[k + n*ldb], k contiguous; else B(k,n) = B[n + k*ldb], n contiguous.
  int bn, bk, bn_step, bk_step;
  if (OPB == 0) { bk = tid % BK; bn = tid / BK; bk_step = 0; bn_step = kThreads / BK; }
  else { bn = tid % BN; bk = tid / BN; bn_step = 0; bk_step = kThreads / BN; }

  S ra[NC][EA], rb[NC][EB];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int64_t m = m0 + am + e * am_step, k = k0 + ak + e * ak_step;
      const bool ok = (m < M) && (k < K);
      const int64_t off = (OPA == 0) ? (m + k * lda) : (k + m * lda);
#pragma unroll
      for (int c = 0; c < NC; ++c) ra[c][e] = ok ? A[off * NC + c] : (S)0;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int64_t n = n0 + bn + e * bn_step, k = k0 + bk + e * bk_step;
      const bool ok = (n < N) && (k < K);
      const int64_t off = (OPB == 0) ? (k + n * ldb) : (n + k * ldb);
#pragma unroll
      for (int c = 0; c < NC; ++c) rb[c][e] = ok ? B[off * NC + c] : (S)0;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < EA; ++e)
#pragma unroll
      for (int c = 0; c < NC; ++c) As[c][am + e * am_step][ak + e * ak_step] = ra[c][e];
#pragma unroll
    for (int e = 0; e < EB; ++e)
#pragma unroll
      for (int c = 0; c < NC; ++c) Bs[c][bn + e * bn_step][bk + e * bk_step] = rb[c][e];
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

  load_tiles(0);
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + BK < K) load_tiles(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      S fa[NC][SM], fb[NC][SN];
#pragma unroll
      for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c) fa[c][i] = As[c][wm * (BM / 2) + i * 16 + (lane & 15)][kk];
#pragma unroll
      for (int j = 0; j < SN; ++j)
#pragma unroll
        for (int c = 0; c < NC; ++c) fb[c][j] = Bs[c][wn * (BN / 2) + j * 16 + (lane & 15)][kk];
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
  }

  // epilogue
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm * (BM / 2) + i * 16 + MF::row(lane, r);
        const int64_t n = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
        if (m < M && n < N) {
          S *c = C + (m + n * ldc) * NC;
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

template <typename S, int NC, int OPA, int OPB, int CJA, int CJB, int BM, int BN>
int launch_typed(int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B, int64_t ldb,
                 void *C, int64_t ldc, double ar, double ai, int beta_one, hipStream_t s) {
  const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
  if (gx > 0x7fffffffLL || gy > 65535) return -4;
  gemm_kernel<S, NC, OPA, OPB, CJA, CJB, BM, BN><<<dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, s>>>(
      M, N, K, (const S *)A, lda, (const S *)B, ldb, (S *)C, ldc, (S)ar, (S)ai, beta_one);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename S, int NC, int BM, int BN>
int dispatch_ops(int opa, int opb, int cja, int cjb, int64_t M, int64_t N, int64_t K, const void *A, int64_t lda,
                 const void *B, int64_t ldb, void *C, int64_t ldc, double ar, double ai, int beta_one,
                 hipStream_t s) {
#define TAL_GEMM_CASE(a, b, ca, cb)                                                                      \
  if (opa == a && opb == b && cja == ca && cjb == cb)                                                    \
    return launch_typed<S, NC, a, b, ca, cb, BM, BN>(M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, s);
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

bool gemm_split_bf16_available() { return false; }

int launch_gemm(int dk, int r4mode, char ta, char tb, int conja, int conjb, int64_t M, int64_t N, int64_t K,
                double ar, double ai, const void *A, int64_t lda, const void *B, int64_t ldb, int beta_one, void *C,
                int64_t ldc, hipStream_t s) {
  (void)r4mode;
  const bool cplx = (dk == 14 || dk == 18);
  const int opa = op_code(ta), opb = op_code(tb);
  const int cja = (cplx && conja) ? 1 : 0, cjb = (cplx && conjb) ? 1 : 0;
  if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0 || !A || !B || !C) return -1;
  if (M == 0 || N == 0) return 0;
  if (ldc < std::max<int64_t>(1, M)) return -1;
  if (lda < std::max<int64_t>(1, opa == 0 ? M : K)) return -1;
  if (ldb < std::max<int64_t>(1, opb == 0 ? K : N)) return -1;
  switch (dk) {
    case 8: return dispatch_ops<double, 1, 128, 128>(opa, opb, cja, cjb, M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, s);
    case 4: return dispatch_ops<float, 1, 128, 128>(opa, opb, cja, cjb, M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, s);
    case 18: return dispatch_ops<double, 2, 64, 64>(opa, opb, cja, cjb, M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, s);
    case 14: return dispatch_ops<float, 2, 64, 64>(opa, opb, cja, cjb, M, N, K, A, lda, B, ldb, C, ldc, ar, ai, beta_one, s);
    default: return -1;
  }
}

}  // namespace tal
