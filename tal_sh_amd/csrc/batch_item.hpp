// Work items of the batched-contraction kernel.
//
// Shared by the device kernel (batch.hip), the host planner (batch_plan.cpp)
// and the CPU emulation of the kernel in the tests, so that every address the
// kernel forms comes from the inline functions below.
//
// An item is one workgroup tile of one contraction D(m,n) (+)= alpha * sum_k
// L(m,k) * R(k,n): 32 BI rows x 32 BJ columns (BI, BJ in {1, 2}) computed by
// the workgroup's four waves, each a quadrant of BI x BJ blocks of 16 x 16
// with 16x16x4 MFMAs.  Rows m, columns n and contracted indices k are
// mixed-radix numbers over up to three (fused) tensor dimensions each; a
// row's byte offset in L and in D, a column's in R and in D and a contracted
// index's in L and in R are sums of digit * stride.  Offsets are 32-bit byte
// offsets into buffer resources whose range is the operand's size: an index
// past the edge (m >= M, n >= N, k >= K) gets the offset kOutside, so its
// loads return 0 and its stores are dropped by the hardware range check, with
// no masks in the kernel.
//
// Operands are staged through LDS kBatchKC contracted indices at a time,
// loaded along whichever index is the operand's contiguous one (its rows, or
// its contracted indices), so that consecutive lanes read consecutive
// addresses; the MFMAs then read their fragments from LDS.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define TAL_HD __host__ __device__
#else
#define TAL_HD
#endif

namespace tal {

// A byte offset outside every batch operand.  Two of them, or one of them
// plus a valid offset, add up to at most 2 * kOutside < 2^32: no wrap-around.
constexpr uint32_t kOutside = 0x7FFFFF00u;
// Largest operand (bytes) a batch accepts, so that every valid offset is below kOutside.
constexpr uint64_t kBatchMaxBytes = kOutside;
// Contracted indices per LDS chunk (four MFMA steps of 4).
constexpr int kBatchKC = 16;
// Chunks whose loads a workgroup keeps in flight ahead of the chunk its MFMAs use.
constexpr int kBatchDepthF64 = 1;
constexpr int kBatchDepthF32 = 1;

// Division of n < 2^31 by a fixed d >= 1 with a multiply-high (round-up
// method of Granlund and Montgomery, PLDI 1994): with l = ceil(log2 d) and
// magic = floor(2^32 * (2^l - d) / d) + 1, q = (mulhi(n, magic) + n) >> l.
// The sum stays below 2^32 because mulhi(n, magic) < n < 2^31.
struct Divider {
  uint32_t magic, shift;
};
inline Divider make_divider(uint32_t d) {
  uint32_t l = 0;
  while ((uint64_t{1} << l) < d) ++l;
  const uint64_t m = ((uint64_t{1} << 32) * ((uint64_t{1} << l) - d)) / d + 1;
  return Divider{(uint32_t)m, l};
}
TAL_HD inline uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
TAL_HD inline uint32_t divide(uint32_t n, uint32_t magic, uint32_t shift) { return (mulhi32(n, magic) + n) >> shift; }

// One index group: x -> digits (d0, d1, d2) over extents (e0, e1, any);
// byte offsets in the group's two tensors o_t = sum_i d_i * s[t][i].  Unused
// dimensions have extent 1 and stride 0.
struct BatchGroup {
  uint32_t e0, e1;
  uint32_t mg0, mg1;  // dividers of e0 and e1
  uint32_t sh0, sh1;
  int32_t s[2][3];
};

// `shape`: bits 0-1 BI - 1, bits 2-3 BJ - 1 (blocks of 16 rows / columns per wave),
// bit 4 set when D is accumulated (D += ...), else D = ..., bit 5 / bit 6 set
// when L / R is staged along its contracted indices (else along its rows /
// columns).
struct BatchItem {
  uint64_t L, R, D;                 // device addresses of the operands
  uint32_t lbytes, rbytes, dbytes;  // their sizes: the buffer ranges
  int32_t M, N, K;
  int32_t m0, n0;  // first row and column of the tile
  int32_t shape;
  int32_t pad0;
  double alpha;
  BatchGroup gm;  // rows: tensors (L, D)
  BatchGroup gn;  // columns: tensors (R, D)
  BatchGroup gk;  // contracted: tensors (L, R)
  int32_t pad1[10];
};
static_assert(sizeof(BatchGroup) == 48, "BatchGroup layout");
static_assert(sizeof(BatchItem) == 256, "BatchItem is four 64-byte lines");

TAL_HD inline int item_bi(const BatchItem &it) { return (it.shape & 3) + 1; }
TAL_HD inline int item_bj(const BatchItem &it) { return ((it.shape >> 2) & 3) + 1; }
TAL_HD inline bool item_accumulates(const BatchItem &it) { return (it.shape >> 4) & 1; }
TAL_HD inline bool item_a_along_k(const BatchItem &it) { return (it.shape >> 5) & 1; }
TAL_HD inline bool item_b_along_k(const BatchItem &it) { return (it.shape >> 6) & 1; }

TAL_HD inline void group_offsets(const BatchGroup &g, uint32_t x, uint32_t *o0, uint32_t *o1) {
  const uint32_t q0 = divide(x, g.mg0, g.sh0), d0 = x - q0 * g.e0;
  const uint32_t d2 = divide(q0, g.mg1, g.sh1), d1 = q0 - d2 * g.e1;
  *o0 = (uint32_t)((int32_t)d0 * g.s[0][0] + (int32_t)d1 * g.s[0][1] + (int32_t)d2 * g.s[0][2]);
  *o1 = (uint32_t)((int32_t)d0 * g.s[1][0] + (int32_t)d1 * g.s[1][1] + (int32_t)d2 * g.s[1][2]);
}

// Byte offsets of an index past its group's extent `count`: kOutside.
// Row m: offset in L (t = 0) or in D (t = 1).
TAL_HD inline uint32_t row_offset(const BatchItem &it, int m, int t) {
  if (m >= it.M) return kOutside;
  uint32_t ol, od;
  group_offsets(it.gm, (uint32_t)m, &ol, &od);
  return t ? od : ol;
}
// Column n: offset in R (t = 0) or in D (t = 1).
TAL_HD inline uint32_t col_offset(const BatchItem &it, int n, int t) {
  if (n >= it.N) return kOutside;
  uint32_t orr, od;
  group_offsets(it.gn, (uint32_t)n, &orr, &od);
  return t ? od : orr;
}
// Contracted index k: offsets in L and in R.
TAL_HD inline void k_offsets(const BatchItem &it, int k, uint32_t *ol, uint32_t *orr) {
  if (k >= it.K) {
    *ol = *orr = kOutside;
    return;
  }
  group_offsets(it.gk, (uint32_t)k, ol, orr);
}

// Staging of an operand tile of `tr` rows x kBatchKC contracted indices by
// `nt` threads: in its load j (0 <= j < tr * kBatchKC / nt) thread t loads
// tile row *r, chunk-local contracted index *c.  Along k, 16 consecutive
// threads run over the contracted indices of one row; along the rows, `tr`
// consecutive threads run over the rows of one contracted index.
TAL_HD inline void stage_coords(bool along_k, int tr, int nt, int t, int j, int *r, int *c) {
  if (along_k) {
    *c = t % kBatchKC;
    *r = t / kBatchKC + j * (nt / kBatchKC);
  } else {
    *r = t % tr;
    *c = t / tr + j * (nt / tr);
  }
}

// Wave v of the workgroup computes the quadrant at tile row 16 BI (v & 1),
// tile column 16 BJ (v >> 1).
TAL_HD inline int quad_row(int bi, int wave) { return 16 * bi * (wave & 1); }
TAL_HD inline int quad_col(int bj, int wave) { return 16 * bj * (wave >> 1); }

// MFMA 16x16x4 operand layout, read from the staged tiles: in step s of a
// chunk, lane l takes A(row 16 i + l % 16, index 4 s + l / 16) of block row i
// and B(index 4 s + l / 16, column 16 j + l % 16) of block column j (rows and
// columns relative to the wave's quadrant).
TAL_HD inline int frag_row(int blk, int lane) { return 16 * blk + (lane & 15); }
TAL_HD inline int frag_k(int step, int lane) { return 4 * step + (lane >> 4); }

// Accumulator layout of the 16x16x4 MFMAs: register r of lane l holds
// C(row, l % 16) with row 4 * r + l / 16 (FP64) or 4 * (l / 16) + r (FP32).
TAL_HD inline int acc_row(int esz, int lane, int r) { return esz == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r; }

}  // namespace tal
