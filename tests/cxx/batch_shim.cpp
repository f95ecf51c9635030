// C entry points over the batched-contraction planner (tal_sh_amd/csrc/
// batch_plan.cpp) and a CPU emulation of the batch kernel's waves, for the
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

// One wave computing one item, lane by lane.
template <typename S>
void emulate_item(const tal::BatchItem &it, long long *bad, int *count) {
  using namespace tal;
  const int BI = item_bi(it), BJ = item_bj(it);
  const void *L = (const void *)it.L, *R = (const void *)it.R;
  void *D = (void *)it.D;
  uint32_t arow[4][64], bcol[4][64], dcol[4][64];
  for (int lane = 0; lane < 64; ++lane) {
    for (int i = 0; i < BI; ++i) arow[i][lane] = a_row_offset(it, i, lane);
    for (int j = 0; j < BJ; ++j) b_col_offsets(it, j, lane, &bcol[j][lane], &dcol[j][lane]);
  }
  std::vector<S> C((size_t)BI * BJ * 256, S(0));  // block (i, j): C[((i * BJ + j) * 16 + row) * 16 + col]
  const int nsteps = (it.K + 3) / 4;
  for (int c = 0; c < nsteps; ++c) {
    S a[4][64], b[4][64];
    for (int lane = 0; lane < 64; ++lane) {
      uint32_t kl, kr;
      k_offsets(it, c, lane, &kl, &kr);
      for (int i = 0; i < BI; ++i) a[i][lane] = load<S>(L, it.lbytes, arow[i][lane] + kl, bad);
      for (int j = 0; j < BJ; ++j) b[j][lane] = load<S>(R, it.rbytes, kr + bcol[j][lane], bad);
    }
    // MFMA 16x16x4: A(row, k) in lane row + 16 k, B(k, col) in lane col + 16 k
    for (int i = 0; i < BI; ++i)
      for (int j = 0; j < BJ; ++j)
        for (int row = 0; row < 16; ++row)
          for (int col = 0; col < 16; ++col) {
            S s = C[((size_t)(i * BJ + j) * 16 + row) * 16 + col];
            for (int k = 0; k < 4; ++k) s += a[i][row + 16 * k] * b[j][col + 16 * k];
            C[((size_t)(i * BJ + j) * 16 + row) * 16 + col] = s;
          }
  }
  const S alpha = (S)it.alpha;
  for (int lane = 0; lane < 64; ++lane)
    for (int i = 0; i < BI; ++i)
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row((int)sizeof(S), lane, r);
        const uint32_t drow = d_row_offset(it, i, row);
        for (int j = 0; j < BJ; ++j) {
          const uint32_t off = drow + dcol[j][lane];
          S v = alpha * C[((size_t)(i * BJ + j) * 16 + row) * 16 + (lane & 15)];
          if (item_accumulates(it)) v = load<S>(D, it.dbytes, off, bad) + v;
          store<S>(D, it.dbytes, off, v, bad, count);
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
// CPU emulation.  Returns batch_task_geometry's code (0 = planned and run);
// *bad = addressing errors, count[e] = stores to D element e, *items = item count.
int batch_emulate(const char *pattern, const int *ddims, const int *ldims, const int *rdims, int esz, const void *L,
                  const void *R, void *D, double alpha, int accumulate, int tm, int tn, long long *bad, int *count,
                  long long *items, int *ndims) {
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

// Schedules `ntask` synthetic contractions (M[t] x N[t] x K[t], wave tile
// tm[t] x tn[t]) over nblocks x waves workers on nxcd XCDs.  Outputs: for each
// scheduled position its task and tile origin, the worker start offsets, the
// estimated cost of each scheduled item and per worker.  Returns the number of items.
long long batch_schedule_probe(int ntask, const int *M, const int *N, const int *K, const int *tm, const int *tn,
                               int esz, int nblocks, int waves, int nxcd, int *out_task, int *out_m0, int *out_n0,
                               int *start, double *icost, double *wcost) {
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
  std::vector<int> st;
  tal::batch_schedule(items, task_of, esz, nblocks, waves, nxcd, &ordered, &st);
  for (size_t i = 0; i < ordered.size(); ++i) {
    out_task[i] = (int)(ordered[i].L >> 20) - 1;
    out_m0[i] = ordered[i].m0;
    out_n0[i] = ordered[i].n0;
    icost[i] = tal::batch_item_cost(ordered[i], esz);
  }
  for (size_t w = 0; w < st.size(); ++w) start[w] = st[w];
  for (size_t w = 0; w + 1 < st.size(); ++w) {
    wcost[w] = 0;
    for (int i = st[w]; i < st[w + 1]; ++i) wcost[w] += tal::batch_item_cost(ordered[(size_t)i], esz);
  }
  return (long long)ordered.size();
}
}
