// Host planning of batched contractions: index groups, work items and the
// static schedule of items over the persistent kernel's waves.  Pure C++ (no
// HIP calls), so the CPU tests build it with g++ and emulate the kernel on it.
#pragma once

#include <cstdint>
#include <vector>

#include "batch_item.hpp"
#include "plan.hpp"

namespace tal {

// One contraction of a batch as the planner sees it.
struct BatchTaskGeom {
  BatchGroup gm, gn, gk;
  int M = 1, N = 1, K = 1;
  int ndims_m = 0, ndims_n = 0, ndims_k = 0;  // dimensions per group after fusion
};

// Builds the index groups of D (+)= L * R from its digital pattern and
// extents.  Rows are L's uncontracted dimensions ordered by their stride in
// L, columns R's ordered by their stride in R (so lanes that load
// consecutive rows / columns read consecutive addresses when such an index
// is the operand's fastest); contracted dimensions follow L's strides, or
// R's when only R has a contracted index as its fastest.  Dimensions of
// extent 1 are dropped and neighbours whose strides chain in both tensors
// are fused.  esz = element bytes.  Returns 0, 1 for an invalid
// pattern / shape, 2 when a group keeps more than three dimensions or a
// hyper index is present (no batch form; the single-contraction path runs it).
int batch_task_geometry(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims, int esz,
                        BatchTaskGeom *g);

// Candidate wave tiles (rows x columns of 16-multiples, at most four 16 x 16 blocks).
struct WaveTile {
  int tm, tn;
};
extern const WaveTile kWaveTiles[6];

// Items of one contraction for wave tile (tm, tn), appended to `out`.
void batch_task_items(const BatchTaskGeom &g, uint64_t L, uint64_t R, uint64_t D, uint64_t lbytes, uint64_t rbytes,
                      uint64_t dbytes, double alpha, bool accumulate, WaveTile tile, std::vector<BatchItem> *out);

// Estimated cost of an item (arbitrary units: issue slots of its loads and MFMAs plus a fixed part).
double batch_item_cost(const BatchItem &it, int esz);

// Static schedule of items over `nblocks` workgroups of `waves` waves each;
// block b runs on XCD b % nxcd.  Contractions are kept on one XCD (their
// items share operand panels in its L2) unless a contraction is a sizeable
// part of the whole; items are then spread over the XCD's waves largest
// first, each to the least loaded wave.  `task_of[i]` is the contraction of
// item i.  Outputs: the items in worker order and, for worker w = b * waves
// + wave, its items at [start[w], start[w + 1]).
void batch_schedule(const std::vector<BatchItem> &items, const std::vector<int> &task_of, int esz, int nblocks,
                    int waves, int nxcd, std::vector<BatchItem> *ordered, std::vector<int> *start);

}  // namespace tal
