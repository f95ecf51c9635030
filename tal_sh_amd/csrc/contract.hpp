// TTGT schedule of one contraction on device buffers, shared by
// talshTensorContract and talshTensorContractXL.  Internal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <initializer_list>

#include "plan.hpp"
#include "talsh.h"

namespace tal {

// Bytes of the temporaries a plan needs (0 = not needed): transposed A, B,
// D', the split-BF16 planes of both GEMM operands, and the partial sums of a
// GEMM whose last round of tiles is split along K.
// `split` is the fast-math split method of an R4 GEMM (-1: plain GEMM), read
// once per operation so the workspace and the kernel always agree.
struct TtgtTemps {
  size_t a = 0, b = 0, d = 0, w = 0, s = 0;
  int split = -1;
};
struct TtgtBuffers {
  void *a = nullptr, *b = nullptr, *d = nullptr, *w = nullptr, *s = nullptr;
  size_t s_bytes = 0;
  int split = -1;
};
TtgtTemps ttgt_temp_bytes(const ContractPlan &plan, int dk, size_t lbytes, size_t rbytes, size_t dbytes);

// Enqueues D (+)= alpha * L * R on stream s: operand transposes, MFMA GEMM
// (bracketed by gemm_start / gemm_end when given), back-transpose of D.
// All pointers are device bodies on the stream's GPU; dims in dlf order.
int enqueue_ttgt(const DigitalPattern &p, const ContractPlan &plan, int dk, bool acc, double ar, double ai,
                 void *D, const void *L, const void *R, const int *ddims, const int *ldims, const int *rdims,
                 const TtgtBuffers &tmp, hipStream_t s, hipEvent_t gemm_start, hipEvent_t gemm_end);

// Frees the device tables enqueue_ttgt keeps between calls (at shutdown,
// after the devices are idle).
void ttgt_cache_clear();

// Copies the box [off, off + sd) of a dlf tensor with extents bd (element
// size esz) between a host body and a packed device buffer (h2d: host to
// device) with strided DMA copies on stream s.  Pageable or pinned host memory.
int box_dma(bool h2d, char *host, char *dev, int rank, const int *bd, const int *off, const int *sd, size_t esz,
            hipStream_t s);

// Execution GPU index for (dev_id, dev_kind); operands decide when the caller
// leaves the choice to the runtime.  Host execution is not provided.
int exec_gpu(int dev_id, int dev_kind, std::initializer_list<const talsh_tens_t *> tens, int *g);

}  // namespace tal
