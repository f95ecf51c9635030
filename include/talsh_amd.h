/* MI355X extensions of libtalsh_amd.
 *
 * These entry points have no counterpart in the reference C API; they expose
 * configuration and the batched/serving forms of the contraction path that
 * only this build provides.  Everything here is optional: a caller that only
 * uses talsh.h gets the reference behaviour with the defaults below.
 */
#ifndef TALSH_AMD_H_
#define TALSH_AMD_H_

#include <stddef.h>

#include "talsh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fast math (talshEnableFastMath etc., declared in talsh.h): on this build
 * "fast math" selects a split GEMM on lower-precision MFMA for R4
 * contractions (FP32 accumulation); disabled, R4/C4 use the FP32 MFMA GEMM.
 * R8/C8 always use the FP64 MFMA GEMM.  The split method is one of: */
#define TALSH_AMD_SPLIT_BF16X6 0 /* default: 3-term BF16 split, 6 products; no range limit */
#define TALSH_AMD_SPLIT_FP16X3 1 /* rows/columns scaled by powers of two, 2-term FP16 split, 3 products; \
                                    small elements keep 2^-25 of their row's largest magnitude */

/* Selects the split method used while fast math is enabled (default
 * TALSH_AMD_SPLIT_BF16X6).  May be called at any time; applies to
 * contractions issued afterwards. */
int talshAmdSetFastMathSplit(int method);
int talshAmdGetFastMathSplit(void);

/* Size in bytes of the per-GPU buffer pool taken at talshInit (default: 80 %
 * of the free device memory).  Must be called before talshInit. */
int talshAmdSetDeviceBufferSize(size_t bytes);

/* Number of HIP streams per GPU used for tasks (default 4, at most 32).
 * Must be called before talshInit. */
int talshAmdSetNumStreams(int n);

/* Device memory (bytes) one talshTensorContractXL call may use for its pieces
 * (two are in flight); 0 (default) = 90 % of the largest free range of the
 * device buffer.  May be called at any time. */
int talshAmdSetContractXLMemory(size_t bytes);

/* Blocks until all work queued on GPU `gpu` (kind-specific number) is done. */
int talshAmdDeviceSynchronize(int gpu);

/* Symbolic -> digital contraction pattern (the reference's C-callable
 * converter, talshf.F90:619-667, declared in talshc.cpp:109): 0 on success. */
int talsh_get_contr_ptrn_str2dig(const char *c_str, int *dig_ptrn, int *drank, int *lrank, int *rrank,
                                 int *conj_bits);

/* The TTGT plan talshTensorContract would use for a pattern and shapes:
 * mnk = {M, N, K} (of one GEMM; a hyper-contraction runs one per combination
 * of its hyper-index values); flags = {A is R, A transposed, B transposed, D transposed};
 * ops = GEMM forms of A and B ('N'/'T'); perms = O2N maps of the transposed
 * operands (A, then B, then D->D'; -1 where not transposed); moved = elements
 * copied by the transposes.  When D' begins with D's first index (extent at
 * least 16) and there is no hyper index, the GEMM reads and writes D in place
 * through per-row and per-column offsets instead of copying it; `moved` still
 * counts D.  Host-only query, no device work. */
int talshAmdPlanContraction(const char *pattern, const int *ddims, const int *ldims, const int *rdims,
                            int accumulative, long long *mnk, int *flags, char *ops, int *perms, double *moved);

/* ---- Batched contractions (one persistent launch, replayed as a hipGraph) ----
 * A batch holds many contractions D (+)= alpha * L * R (same pattern syntax as
 * talshTensorContract) on tensors resident on the execution GPU, all of one
 * real data kind (R4 or R8).  Outputs must be distinct tensors and no output
 * may be an input of another contraction in the batch.  The first execution
 * plans the batch (gather tables, autotuned tile shapes, captured graph); an
 * unchanged batch re-executes by replaying the graph.  Execution is
 * asynchronous when talsh_task is given (talshTaskWait etc. apply), blocking
 * otherwise. */
typedef struct talshAmdBatch talshAmdBatch;
int talshAmdBatchCreate(talshAmdBatch **batch);
int talshAmdBatchAdd(talshAmdBatch *batch, const char *pattern, talsh_tens_t *dtens, talsh_tens_t *ltens,
                     talsh_tens_t *rtens, double scale_real, double scale_imag, int accumulative);
int talshAmdBatchSize(const talshAmdBatch *batch);
int talshAmdBatchExecute(talshAmdBatch *batch, int dev_id, int dev_kind, talsh_task_t *talsh_task);
/* FLOP per execution, work items (tiles), shape classes tuned by this batch,
 * tuning time (ms) and number of executions so far. */
int talshAmdBatchStats(const talshAmdBatch *batch, double *flops, long long *work_items, int *tuned_classes,
                       double *tune_ms, int *executions);
int talshAmdBatchDestroy(talshAmdBatch *batch);

/* ---- Serving: continuous admission of contraction requests ----
 * Requests are talshTensorContract operations submitted at any time.  They
 * are admitted onto the GPU, in arrival order (a waiting head may be
 * overtaken by smaller requests at most 64 times), whenever fewer than
 * max_inflight requests are running and the request's transient device
 * memory (operands it has to bring to the GPU + transposed temporaries) fits
 * in what is left of hbm_budget.  Admitted requests run asynchronously and
 * overlap on the GPU's streams; talshAmdServerProgress reports completions
 * (tag, latency from submission, task status) and admits more work; it
 * reports at most max_out completions per call (the rest on later calls).
 * A request that cannot run even on an idle device is reported with status
 * TALSH_TASK_ERROR.  Submit returns DEVICE_UNABLE for a request whose
 * transient need exceeds the whole budget.
 * Operands already resident on the GPU are used in place and cost no budget.
 * The caller keeps every tensor of a request alive and untouched until its
 * completion is reported. */
typedef struct talshAmdServer talshAmdServer;
typedef struct {
  const char *pattern;
  talsh_tens_t *dtens;
  talsh_tens_t *ltens;
  talsh_tens_t *rtens;
  double scale_real;
  double scale_imag;
  int copy_ctrl;    /* as for talshTensorContract */
  int accumulative; /* YEP / NOPE */
  long long tag;    /* caller's id, returned on completion */
} talshAmdRequest;
int talshAmdServerCreate(talshAmdServer **server, int gpu, size_t hbm_budget, int max_inflight);
int talshAmdServerSubmit(talshAmdServer *server, const talshAmdRequest *request);
int talshAmdServerProgress(talshAmdServer *server, int max_out, long long *tags, double *latency_s, int *status,
                           int *nout);
int talshAmdServerPending(const talshAmdServer *server, int *queued, int *inflight);
/* out[12] = {submitted, completed, failed, deferred, p50 latency s, p99 latency s, mean latency s,
 *            p50 queue wait s, p99 queue wait s, TFLOP/s over first arrival..last completion,
 *            peak transient bytes, budget bytes} */
int talshAmdServerStats(const talshAmdServer *server, double *out, int nout);
int talshAmdServerDestroy(talshAmdServer *server);

/* Allocation-free norm2.  No counterpart in the reference: this is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernels and result
 * as talshTensorNorm2, on raw device pointers: *result (a device address
 * holding one element of data_kind, like talshTensorNorm2's rank-0 output)
 * = sum of |x|^2 over the `volume` elements of kind data_kind (R4, R8, C4,
 * C8) at device address `body` on GPU `gpu`, enqueued on `stream` (a
 * hipStream_t of that GPU; NULL = its null stream).  The reduction's partial
 * sums live in a fixed per-GPU device array of 4 KiB (512 doubles, set up by
 * talshInit; enough for any volume); no pool buffer, task, event or device
 * allocation is made.  The caller orders the call against other work on
 * `body` and `result` (no hazard tracking).  Calls that overlap on one GPU
 * share the array and race; serialise them (one stream, or events).
 * Returns TALSH_NOT_INITIALIZED before talshInit, TALSH_INVALID_ARGS for a
 * bad kind, GPU, pointer or volume 0. */
int talshAmdTensorNorm2NoAlloc(int data_kind, const void *body, size_t volume, void *result, int gpu,
                               void *stream);

/* Copies a short build description into buf. */
int talshAmdBuildInfo(char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* TALSH_AMD_H_ */
