// Device-kernel launchers of libtalsh_amd (gfx950).  All launchers are
// asynchronous on the given stream and validate their shapes on the host
// before launching; a nonzero return means nothing was launched.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tal {

// Optional dispatch timestamps around the kernels of one operation (per host
// thread).  While `on`, launch_clocked gives the first kernel `t0` as its
// start event and every kernel `t1` as its stop event, so the span (t0, t1)
// covers the kernels and what runs between them, but not the host's launch
// latency before the first one.
struct KernelClock {
  bool on = false, started = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
};
KernelClock &kernel_clock();

template <typename F, typename... A>
inline void launch_clocked(F kernel, dim3 grid, dim3 block, hipStream_t s, A... args) {
  KernelClock &c = kernel_clock();
  if (c.on) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, c.started ? nullptr : c.t0, c.t1, 0, args...);
    c.started = true;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
  }
}

enum PermOp { PERM_COPY = 0, PERM_ADD = 1 };

// dst = perm(src) (PERM_COPY) or dst += alpha * perm(src) (PERM_ADD) for a
// dlf tensor of `rank` dims with extents `dims` (input order) and a 0-based
// O2N permutation `o2n`.  Complex kinds may conjugate src.  Bit-exact.
int launch_permute(int dk, const void *src, void *dst, int rank, const int *dims, const int *o2n, int op,
                   double ar, double ai, int conj, hipStream_t s);
// Forgets launch_permute's per-shape kernel-variant choices and frees their
// timing events (at shutdown).
void perm_tune_clear();
// Two independent permutation copies (dst_j = perm_j(src_j), arguments as
// launch_permute's with op PERM_COPY; neither destination may overlap the
// other copy's source or destination) in one launch when both take the same
// kernel, else in two, on stream s.
int launch_permute_pair(int dk, const void *src0, void *dst0, int rank0, const int *dims0, const int *o2n0,
                        const void *src1, void *dst1, int rank1, const int *dims1, const int *o2n1, hipStream_t s);

// Slice (mode 0): small = big[offsets + i] (small += ... when accumulative);
// insert (mode 1): big[offsets + i] = small (+=).  small has extents sdims,
// big has extents bdims, both dlf with the same dimension order.  Bit-exact.
int launch_slice(int dk, int mode, const void *src, void *dst, int rank, const int *sdims, const int *bdims,
                 const int *offsets, int accumulative, hipStream_t s);

// d = val (all elements)
int launch_init(int dk, void *d, size_t vol, double vr, double vi, hipStream_t s);
// d = d * alpha
int launch_scale(int dk, void *d, size_t vol, double ar, double ai, hipStream_t s);
// *out (data kind dk, rank-0 body) = sum |x|^2 over src, FP64 two-stage
// deterministic reduction; `partials` is device scratch of norm2_scratch_elems() doubles.
size_t norm2_scratch_elems();
int launch_norm2(int dk, const void *src, size_t vol, double *partials, void *out, hipStream_t s);
// The same reduction with its partials in a static per-GPU device array
// instead of caller scratch.  noalloc_scratch_setup(hip_dev), with that device
// current, looks up the array's address on it (once per device, at init);
// launch_norm2_noalloc returns -3 on a device not set up.  Calls that overlap
// on one GPU share the array and race.
constexpr int kNoAllocMaxDevices = 64;
int noalloc_scratch_setup(int hip_dev);
size_t norm2_noalloc_scratch_bytes();
int launch_norm2_noalloc(int hip_dev, int dk, const void *src, size_t vol, void *out, hipStream_t s);

// GEMM precision modes for real single precision.
enum R4Mode { R4_FP32 = 0, R4_SPLIT_BF16 = 1 };

// Whether the split-BF16 R4 GEMM is built into this library.
bool gemm_split_bf16_available();

// Split methods of the R4 GEMM on lower-precision MFMA: three BF16 planes per
// operand and six products (bf16x6), or two FP16 planes of the row-scaled
// operand and three products (fp16x3).
enum SplitMethod { SPLIT_BF16X6 = 0, SPLIT_FP16X3 = 1 };

// Split R4 GEMM with pre-split operands (same contract as launch_gemm for
// dk = R4, batch 1, including cmap); ws is device scratch of
// gemm_split_workspace(method, M, N, K) bytes, 16-byte aligned, that holds
// the planes of both operands.
size_t gemm_split_workspace(int method, int64_t M, int64_t N, int64_t K);
int launch_gemm_split(int method, char ta, char tb, int64_t M, int64_t N, int64_t K, double ar, const void *A,
                      int64_t lda, const void *B, int64_t ldb, int beta_one, void *C, int64_t ldc, void *ws,
                      hipStream_t s, const int64_t *cmap = nullptr);

// C(MxN, col-major, ldc) = alpha * op(A) * op(B) + beta * C, beta in {0,1}.
// ta: 'N' A is MxK (lda>=M); 'T' A is KxM (lda>=K).
// tb: 'N' B is KxN (ldb>=K); 'T' B is NxK (ldb>=N).
// conja/conjb conjugate complex operands.  Complex kinds use interleaved (re,im) storage.
// batch > 1: that many independent products, entry q at A + q*sa, B + q*sb,
// C + q*sc (strides in elements; C entries may not overlap).
// ws (ws_bytes >= gemm_tail_workspace(...)) lets a GEMM whose last round of
// tiles would leave most of the GPU idle compute those tiles in K ranges;
// without it every tile runs whole.  The result depends only on the shapes.
// cmap (device, M + N entries; batch 1): element (m, n)
// of C is at C[cmap[m] + cmap[M + n]] (elements), so the GEMM can read and
// write a tensor whose M and N indices interleave in memory (ldc unused).
size_t gemm_tail_workspace(int dk, int64_t M, int64_t N, int64_t K, int64_t batch);
int launch_gemm(int dk, int r4mode, char ta, char tb, int conja, int conjb, int64_t M, int64_t N, int64_t K,
                double ar, double ai, const void *A, int64_t lda, const void *B, int64_t ldb, int beta_one, void *C,
                int64_t ldc, hipStream_t s, int64_t batch = 1, int64_t sa = 0, int64_t sb = 0, int64_t sc = 0,
                void *ws = nullptr, size_t ws_bytes = 0, const int64_t *cmap = nullptr);

// Offsets of a C map: for the mixed-radix index m over extents ext[0..nm)
// (first fastest), out[m] = sum of digit_t * str[t]; out[M + n] likewise
// over ext[nm..nm+nn) with M = prod ext[0..nm), N = prod ext[nm..nm+nn).
constexpr int kMaxMapDims = 64;
struct OffsetMap {
  int nm = 0, nn = 0;
  int64_t M = 1, N = 1;
  int64_t ext[kMaxMapDims], str[kMaxMapDims];
};
int launch_offset_map(const OffsetMap &map, int64_t *out, hipStream_t s);

}  // namespace tal
