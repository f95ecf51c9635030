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
  bool l_along_k = false, r_along_k = false;  // the operand's contiguous index is a contracted one
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

// Candidate workgroup tiles (rows x columns: 32 or 64 each).
struct WaveTile {
  int tm, tn;
};
extern const WaveTile kWaveTiles[4];

// Items of one contraction for workgroup tile (tm, tn), appended to `out`.
void batch_task_items(const BatchTaskGeom &g, uint64_t L, uint64_t R, uint64_t D, uint64_t lbytes, uint64_t rbytes,
                      uint64_t dbytes, double alpha, bool accumulate, WaveTile tile, std::vector<BatchItem> *out);

// Estimated cost of an item (arbitrary units: issue slots of its loads and MFMAs plus a fixed part).
double batch_item_cost(const BatchItem &it, int esz);

// Work queues of the persistent kernel, one per XCD (workgroup b takes its
// items from queue b % nq while it has any, then from the others in turn).
// Contractions are kept whole in one queue (their items share operand panels
// in the XCD's L2) unless a contraction is a sizeable part of the whole; it
// is then cut into runs of neighbouring tiles.  Units go to the least loaded
// queue, largest first, and each queue holds its units largest first with a
// unit's items side by side, so the waves that pull them concurrently share
// the unit's lines and the queue ends with its smallest items.  `task_of[i]`
// is the contraction of item i.  Outputs: the items in queue order and, for
// queue q, its items at [qstart[q], qstart[q + 1]).
void batch_queues(const std::vector<BatchItem> &items, const std::vector<int> &task_of, int esz, int nq,
                  std::vector<BatchItem> *ordered, std::vector<int> *qstart);

}  // namespace tal
