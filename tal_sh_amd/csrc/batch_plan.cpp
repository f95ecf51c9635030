// Host planning of batched contractions (see batch_plan.hpp).
#include "batch_plan.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <utility>

namespace tal {

const WaveTile kWaveTiles[4] = {{32, 32}, {32, 64}, {64, 32}, {64, 64}};

namespace {

struct Dim {
  int64_t ext, s0, s1;  // extent, element strides in the group's two tensors
};

std::vector<int64_t> dlf_strides(int rank, const int *dims) {
  std::vector<int64_t> s((size_t)rank);
  int64_t acc = 1;
  for (int i = 0; i < rank; ++i) {
    s[(size_t)i] = acc;
    acc *= dims[i];
  }
  return s;
}

// Drops unit extents, orders by the stride in tensor `by` (0 or 1), fuses
// neighbours whose strides chain in both tensors; false if more than 3 remain.
bool make_group(std::vector<Dim> dims, int by, int esz, BatchGroup *g, int *total, int *nd) {
  dims.erase(std::remove_if(dims.begin(), dims.end(), [](const Dim &d) { return d.ext == 1; }), dims.end());
  std::stable_sort(dims.begin(), dims.end(),
                   [by](const Dim &a, const Dim &b) { return (by ? a.s1 : a.s0) < (by ? b.s1 : b.s0); });
  std::vector<Dim> f;
  for (const Dim &d : dims) {
    if (!f.empty() && d.s0 == f.back().s0 * f.back().ext && d.s1 == f.back().s1 * f.back().ext)
      f.back().ext *= d.ext;
    else
      f.push_back(d);
  }
  *nd = (int)f.size();
  if (f.size() > 3) return false;
  int64_t n = 1;
  for (const Dim &d : f) n *= d.ext;
  *total = (int)n;
  while (f.size() < 3) f.push_back(Dim{1, 0, 0});
  const Divider v0 = make_divider((uint32_t)f[0].ext), v1 = make_divider((uint32_t)f[1].ext);
  g->e0 = (uint32_t)f[0].ext;
  g->e1 = (uint32_t)f[1].ext;
  g->mg0 = v0.magic;
  g->sh0 = v0.shift;
  g->mg1 = v1.magic;
  g->sh1 = v1.shift;
  for (int i = 0; i < 3; ++i) {
    g->s[0][i] = (int32_t)(f[(size_t)i].s0 * esz);
    g->s[1][i] = (int32_t)(f[(size_t)i].s1 * esz);
  }
  return true;
}

// First dimension of extent > 1 (the fastest one that moves), or -1.
int fastest(int rank, const int *dims) {
  for (int i = 0; i < rank; ++i)
    if (dims[i] > 1) return i;
  return -1;
}

}  // namespace

int batch_task_geometry(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims, int esz,
                        BatchTaskGeom *g) {
  ContractPlan cp;  // validates the pattern against the extents
  if (plan_contraction(p, ddims, ldims, rdims, 1, &cp) != 0) return 1;
  if (cp.batch > 1) return 2;
  const std::vector<int64_t> ls = dlf_strides(p.lrank, ldims), rs = dlf_strides(p.rrank, rdims),
                             ds = dlf_strides(p.drank, ddims);
  std::vector<Dim> m, n, k;
  for (int i = 0; i < p.lrank; ++i) {
    const int c = p.dig[(size_t)i];
    if (c > 0)
      m.push_back(Dim{ldims[i], ls[(size_t)i], ds[(size_t)(c - 1)]});
    else
      k.push_back(Dim{ldims[i], ls[(size_t)i], rs[(size_t)(-c - 1)]});
  }
  for (int j = 0; j < p.rrank; ++j) {
    const int c = p.dig[(size_t)(p.lrank + j)];
    if (c > 0) n.push_back(Dim{rdims[j], rs[(size_t)j], ds[(size_t)(c - 1)]});
  }
  const int lf = fastest(p.lrank, ldims), rf = fastest(p.rrank, rdims);
  const bool l_kfast = lf >= 0 && p.dig[(size_t)lf] < 0, r_kfast = rf >= 0 && p.dig[(size_t)(p.lrank + rf)] < 0;
  const int kby = (r_kfast && !l_kfast) ? 1 : 0;
  g->l_along_k = l_kfast;
  g->r_along_k = r_kfast;
  if (!make_group(m, 0, esz, &g->gm, &g->M, &g->ndims_m) || !make_group(n, 0, esz, &g->gn, &g->N, &g->ndims_n) ||
      !make_group(k, kby, esz, &g->gk, &g->K, &g->ndims_k))
    return 2;
  return 0;
}

void batch_task_items(const BatchTaskGeom &g, uint64_t L, uint64_t R, uint64_t D, uint64_t lbytes, uint64_t rbytes,
                      uint64_t dbytes, double alpha, bool accumulate, WaveTile tile, std::vector<BatchItem> *out) {
  BatchItem it{};
  it.L = L;
  it.R = R;
  it.D = D;
  it.lbytes = (uint32_t)lbytes;
  it.rbytes = (uint32_t)rbytes;
  it.dbytes = (uint32_t)dbytes;
  it.M = g.M;
  it.N = g.N;
  it.K = g.K;
  it.shape = (tile.tm / 32 - 1) | ((tile.tn / 32 - 1) << 2) | (accumulate ? 16 : 0) | (g.l_along_k ? 32 : 0) |
             (g.r_along_k ? 64 : 0);
  it.alpha = alpha;
  it.gm = g.gm;
  it.gn = g.gn;
  it.gk = g.gk;
  for (int n0 = 0; n0 < g.N; n0 += tile.tn)
    for (int m0 = 0; m0 < g.M; m0 += tile.tm) {
      it.m0 = m0;
      it.n0 = n0;
      out->push_back(it);
    }
}

double batch_item_cost(const BatchItem &it, int esz) {
  // The kernel is bound by memory latency: an item costs about three round
  // trips (descriptor, first loads, stores) plus one per kBatchDepth chunks in
  // flight, plus the issue time of its loads and MFMAs.
  const double bi = item_bi(it), bj = item_bj(it), chunks = (it.K + kBatchKC - 1) / kBatchKC;
  const double depth = esz == 8 ? kBatchDepthF64 : kBatchDepthF32;
  return 3.0 + chunks / depth + chunks * (0.1 * (bi + bj) + bi * bj * (esz == 8 ? 0.2 : 0.1));
}

void batch_queues(const std::vector<BatchItem> &items, const std::vector<int> &task_of, int esz, int nq,
                  std::vector<BatchItem> *ordered, std::vector<int> *qstart) {
  nq = std::max(1, nq);
  std::vector<double> cost(items.size());
  double total = 0;
  int ntask = 0;
  for (size_t i = 0; i < items.size(); ++i) {
    cost[i] = batch_item_cost(items[i], esz);
    total += cost[i];
    ntask = std::max(ntask, task_of[i] + 1);
  }
  std::vector<std::vector<int>> of_task((size_t)ntask);
  std::vector<double> tcost((size_t)ntask, 0.0);
  for (size_t i = 0; i < items.size(); ++i) {
    of_task[(size_t)task_of[i]].push_back((int)i);
    tcost[(size_t)task_of[i]] += cost[i];
  }
  // units: a contraction, or for a large one runs of consecutive items
  // (neighbouring tiles share operand panels) of about `chunk` cost each
  const double chunk = total / nq / 8;
  struct Unit {
    double cost;
    std::vector<int> items;
  };
  std::vector<Unit> units;
  for (int t = 0; t < ntask; ++t) {
    const std::vector<int> &its = of_task[(size_t)t];
    if (its.empty()) continue;
    const int pieces = std::max(1, (int)std::floor(tcost[(size_t)t] / std::max(chunk, 1e-30)));
    const double per = tcost[(size_t)t] / pieces;
    Unit u{0.0, {}};
    int emitted = 0;
    for (int i : its) {
      u.items.push_back(i);
      u.cost += cost[(size_t)i];
      if (emitted + 1 < pieces && u.cost >= per) {  // the last piece takes the rest
        units.push_back(u);
        u = Unit{0.0, {}};
        ++emitted;
      }
    }
    if (!u.items.empty()) units.push_back(u);
  }
  std::vector<int> order(units.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return units[(size_t)a].cost > units[(size_t)b].cost; });
  std::vector<double> qload((size_t)nq, 0.0);
  std::vector<std::vector<int>> qunits((size_t)nq);
  for (int u : order) {
    const int q = (int)(std::min_element(qload.begin(), qload.end()) - qload.begin());
    qunits[(size_t)q].push_back(u);
    qload[(size_t)q] += units[(size_t)u].cost;
  }
  ordered->clear();
  ordered->reserve(items.size());
  qstart->assign((size_t)nq + 1, 0);
  for (int q = 0; q < nq; ++q) {
    (*qstart)[(size_t)q] = (int)ordered->size();
    for (int u : qunits[(size_t)q])
      for (int i : units[(size_t)u].items) ordered->push_back(items[(size_t)i]);
  }
  (*qstart)[(size_t)nq] = (int)ordered->size();
}

}  // namespace tal
