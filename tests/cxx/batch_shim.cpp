// C entry points over the batched-contraction planner (tal_sh_amd/csrc/
// batch_plan.cpp) and a CPU emulation of the batch kernel's workgroups, for the
// CPU test tests/test_batch_plan.py, which builds this file together with
// batch_plan.cpp and plan.cpp using g++.
//
// The emulation forms every operand, result and range address with the same
// inline functions of batch_item.hpp that the kernel (batch.hip) uses, and
// models the hardware the kernel relies on: 16x16x4 MFMA operand and
// accumulator layouts, and buffer accesses at or past the range returning 0
// (loads) or being dropped (stores).  An offset between an operand's end and
// kOutside, or a misaligned one, is counted as a planner error.
#include <cstdint>
#include <cstring>
#include <vector>

#include "batch_plan.hpp"

namespace {

template <typename S>
S load(const void *base, uint32_t bytes, uint32_t off, long long *bad) {
  if (off % sizeof(S) != 0) ++*bad;
  if ((uint64_t)off + sizeof(S) <= bytes) {
    S v;
    std::memcpy(&v, (const char *)base + off, sizeof(S));
    return v;
  }
  if (off < tal::kOutside) ++*bad;
  return S(0);
}

template <typename S>
void store(void *base, uint32_t bytes, uint32_t off, S v, long long *bad, int *count) {
  if (off % sizeof(S) != 0) ++*bad;
  if ((uint64_t)off + sizeof(S) <= bytes) {
    std::memcpy((char *)base + off, &v, sizeof(S));
    if (count) ++count[off / sizeof(S)];
    return;
  }
  if (off < tal::kOutside) ++*bad;
}

// One workgroup computing one item, thread by thread: every chunk of
// kBatchKC contracted indices is loaded into the staged tiles with the
// kernel's thread mapping (each tile entry written exactly once per chunk,
// checked), then each wave reads its quadrant's MFMA fragments.
template <typename S>
void emulate_item(const tal::BatchItem &it, long long *bad, int *count) {
  using namespace tal;
  constexpr int NT = 256;
  const int BI = item_bi(it), BJ = item_bj(it), TA = 32 * BI, TB = 32 * BJ;
  const bool ak = item_a_along_k(it), bk = item_b_along_k(it);
  const void *L = (const void *)it.L, *R = (const void *)it.R;
  void *D = (void *)it.D;
  // C[wave][block i][block j][row][col]
  std::vector<S> C((size_t)4 * BI * BJ * 256, S(0));
  auto cidx = [&](int w, int i, int j, int row, int col) {
    return ((((size_t)w * BI + i) * BJ + j) * 16 + row) * 16 + col;
  };
  const int nch = (it.K + kBatchKC - 1) / kBatchKC;
  for (int ch = 0; ch < nch; ++ch) {
    const int k0 = ch * kBatchKC;
    S a[64][kBatchKC], b[64][kBatchKC];
    int wa[64][kBatchKC] = {}, wb[64][kBatchKC] = {};
    for (int t = 0; t < NT; ++t) {
      for (int j = 0; j < TA * kBatchKC / NT; ++j) {
        int r, c;
        stage_coords(ak, TA, NT, t, j, &r, &c);
        uint32_t kl, kr;
        k_offsets(it, k0 + c, &kl, &kr);
        a[r][c] = load<S>(L, it.lbytes, row_offset(it, it.m0 + r, 0) + kl, bad);
        ++wa[r][c];
      }
      for (int j = 0; j < TB * kBatchKC / NT; ++j) {
        int r, c;
        stage_coords(bk, TB, NT, t, j, &r, &c);
        uint32_t kl, kr;
        k_offsets(it, k0 + c, &kl, &kr);
        b[r][c] = load<S>(R, it.rbytes, kr + col_offset(it, it.n0 + r, 0), bad);
        ++wb[r][c];
      }
    }
    for (int r = 0; r < 64; ++r)
      for (int c = 0; c < kBatchKC; ++c) {
        if (wa[r][c] != (r < TA ? 1 : 0)) ++*bad;
        if (wb[r][c] != (r < TB ? 1 : 0)) ++*bad;
      }
    // MFMA 16x16x4 steps per wave: A(row, k) in lane row + 16 k, B(k, col) in lane col + 16 k
    for (int w = 0; w < 4; ++w) {
      const int qr = quad_row(BI, w), qc = quad_col(BJ, w);
      for (int st = 0; st < kBatchKC / 4; ++st) {
        S fa[2][64], fb[2][64];
        for (int lane = 0; lane < 64; ++lane) {
          for (int i = 0; i < BI; ++i) fa[i][lane] = a[qr + frag_row(i, lane)][frag_k(st, lane)];
          for (int j = 0; j < BJ; ++j) fb[j][lane] = b[qc + frag_row(j, lane)][frag_k(st, lane)];
        }
        for (int i = 0; i < BI; ++i)
          for (int j = 0; j < BJ; ++j)
            for (int row = 0; row < 16; ++row)
              for (int col = 0; col < 16; ++col) {
                S s = C[cidx(w, i, j, row, col)];
                for (int k = 0; k < 4; ++k) s += fa[i][row + 16 * k] * fb[j][col + 16 * k];
                C[cidx(w, i, j, row, col)] = s;
              }
      }
    }
  }
  const S alpha = (S)it.alpha;
  for (int w = 0; w < 4; ++w) {
    const int qr = quad_row(BI, w), qc = quad_col(BJ, w);
    for (int lane = 0; lane < 64; ++lane)
      for (int i = 0; i < BI; ++i)
        for (int r = 0; r < 4; ++r) {
          const int row = acc_row((int)sizeof(S), lane, r);
          const uint32_t drow = row_offset(it, it.m0 + qr + 16 * i + row, 1);
          for (int j = 0; j < BJ; ++j) {
            const uint32_t off = drow + col_offset(it, it.n0 + qc + frag_row(j, lane), 1);
            S v = alpha * C[cidx(w, i, j, row, lane & 15)];
            if (item_accumulates(it)) v = load<S>(D, it.dbytes, off, bad) + v;
            store<S>(D, it.dbytes, off, v, bad, count);
          }
        }
  }
}

}  // namespace

extern "C" {

// Exhaustive check of the multiply-high division for d in [1, dmax] and n in
// [0, nmax), plus n near 2^31 and near multiples of d; returns mismatches.
long long batch_divider_check(uint32_t dmax, uint32_t nmax) {
  long long bad = 0;
  for (uint32_t d = 1; d <= dmax; ++d) {
    const tal::Divider v = tal::make_divider(d);
    auto check = [&](uint32_t n) {
      if (tal::divide(n, v.magic, v.shift) != n / d) ++bad;
    };
    for (uint32_t n = 0; n < nmax; ++n) check(n);
    for (uint32_t n = 0x7FFFFFFFu; n > 0x7FFFFFFFu - 4096; --n) check(n);
    for (uint64_t q = 1; q * d < 0x7FFFFFFFu; q = q * 3 + 1) {
      const uint64_t n = q * d;
      check((uint32_t)n);
      check((uint32_t)(n - 1));
      if (n + 1 < 0x80000000u) check((uint32_t)(n + 1));
    }
  }
  return bad;
}

// Plans D (+)= alpha * L * R with wave tile tm x tn and runs every item on the
// CPU emulation.  Returns batch_task_geometry's code (0 = planned and run; 3 = not a kernel tile);
// *bad = addressing errors, count[e] = stores to D element e, *items = item count.
int batch_emulate(const char *pattern, const int *ddims, const int *ldims, const int *rdims, int esz, const void *L,
                  const void *R, void *D, double alpha, int accumulate, int tm, int tn, long long *bad, int *count,
                  long long *items, int *ndims) {
  if ((tm != 32 && tm != 64) || (tn != 32 && tn != 64)) return 3;  // not a workgroup tile of the kernel
  tal::DigitalPattern p;
  if (tal::parse_pattern(pattern, &p) != 0) return 1;
  tal::BatchTaskGeom g;
  const int e = tal::batch_task_geometry(p, ddims, ldims, rdims, esz, &g);
  if (e) return e;
  ndims[0] = g.ndims_m;
  ndims[1] = g.ndims_n;
  ndims[2] = g.ndims_k;
  auto vol = [](int rank, const int *dims) {
    uint64_t v = 1;
    for (int i = 0; i < rank; ++i) v *= (uint64_t)dims[i];
    return v;
  };
  std::vector<tal::BatchItem> its;
  tal::batch_task_items(g, (uint64_t)L, (uint64_t)R, (uint64_t)D, vol(p.lrank, ldims) * esz, vol(p.rrank, rdims) * esz,
                        vol(p.drank, ddims) * esz, alpha, accumulate != 0, tal::WaveTile{tm, tn}, &its);
  *items = (long long)its.size();
  *bad = 0;
  for (const auto &it : its) {
    if (esz == 8)
      emulate_item<double>(it, bad, count);
    else
      emulate_item<float>(it, bad, count);
  }
  return 0;
}

// Queues `ntask` synthetic contractions (M[t] x N[t] x K[t], wave tile
// tm[t] x tn[t]) over nq work queues.  Outputs: for each queued position its
// task and tile origin and estimated cost, and the queue start offsets.
// Returns the number of items.
long long batch_queue_probe(int ntask, const int *M, const int *N, const int *K, const int *tm, const int *tn, int esz,
                            int nq, int *out_task, int *out_m0, int *out_n0, double *icost, int *qstart) {
  std::vector<tal::BatchItem> items;
  std::vector<int> task_of;
  for (int t = 0; t < ntask; ++t) {
    tal::BatchTaskGeom g;
    g.M = M[t];
    g.N = N[t];
    g.K = K[t];
    const size_t before = items.size();
    tal::batch_task_items(g, (uint64_t)(t + 1) << 20, 0, 0, 0, 0, 0, 1.0, false, tal::WaveTile{tm[t], tn[t]}, &items);
    task_of.insert(task_of.end(), items.size() - before, t);
  }
  std::vector<tal::BatchItem> ordered;
  std::vector<int> qs;
  tal::batch_queues(items, task_of, esz, nq, &ordered, &qs);
  for (size_t i = 0; i < ordered.size(); ++i) {
    out_task[i] = (int)(ordered[i].L >> 20) - 1;
    out_m0[i] = ordered[i].m0;
    out_n0[i] = ordered[i].n0;
    icost[i] = tal::batch_item_cost(ordered[i], esz);
  }
  for (size_t q = 0; q < qs.size(); ++q) qstart[q] = qs[q];
  return (long long)ordered.size();
}
}
