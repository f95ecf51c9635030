/* TAL-SH device-unified C API, MI355X implementation (libtalsh_amd.so).
 *
 * This header is the drop-in boundary: the entry points below carry the
 * reference's names, argument meanings, return codes and struct layouts
 * (talsh.h:25-597 of TAL-SH), so that code written against TAL-SH's C API
 * (test.cpp, talshxx.hpp, talshf.F90's bind(c) interfaces) links against this
 * library unchanged.  Each declaration cites the reference declaration it
 * replaces.  Execution goes to HIP kernels on gfx950; there is no CPU
 * execution path: an operation scheduled on DEV_HOST returns
 * TALSH_NOT_AVAILABLE.  Host images are still supported as data sources and
 * destinations (host<->device copies follow the coherence letters).
 *
 * Semantic choices where the reference is silent or inconsistent:
 *  - talshTensorNorm2 (documented in TALSH_manual.txt:893-913, never
 *    declared in the reference talsh.h) returns the SQUARED Frobenius norm,
 *    as the reference's CPU kernels and tests do (tensor_algebra_cpu.F90:
 *    2032-2129, test.cpp:1073), accumulated in FP64 for every data kind.
 *  - GPUs may be addressed with either the kind-1 numbering (flat ids 1..8,
 *    the numbering the reference's own HIP build uses) or DEV_AMD_GPU
 *    (flat ids 17..24); both name the same devices.
 */
#ifndef TALSH_H_
#define TALSH_H_

#include "tensor_algebra.h"

#ifdef __cplusplus
#define TALSH_DEFAULT(x) = x
#else
#define TALSH_DEFAULT(x)
#endif

/* parameters (talsh.h:17-23) */
#define TALSH_MAX_DEV_PRESENT 16
#define TALSH_MEM_ALLOC_POLICY_HOST MEM_ALLOC_TMP_BUF /* default Host allocation policy */
#define TALSH_MEM_ALLOC_FALLBACK_HOST 1               /* default fallback to regular memory: {0|1} */
#define TALSH_CPTAL_MIN_BUF_SIZE 1073741824           /* smallest Host buffer CP-TAL uses effectively */
#define TALSH_NO_HOST_BUFFER 16777216
#define TALSH_GFLOP_THRESH_GPU 4

/* tensor operation kinds and stages (talsh.h:53-82) */
#define TALSH_TENSOR_NOOP -1
#define TALSH_TENSOR_INIT 68
#define TALSH_TENSOR_NORM1 69
#define TALSH_TENSOR_NORM2 70
#define TALSH_TENSOR_MIN 71
#define TALSH_TENSOR_MAX 72
#define TALSH_TENSOR_FOLD 73
#define TALSH_TENSOR_UNFOLD 74
#define TALSH_TENSOR_SLICE 75
#define TALSH_TENSOR_INSERT 76
#define TALSH_TENSOR_COPY 77
#define TALSH_TENSOR_PERMUTE 78
#define TALSH_TENSOR_SCALE 79
#define TALSH_TENSOR_ADD 80
#define TALSH_TENSOR_TRACE 81
#define TALSH_TENSOR_CONTRACT 82
#define TALSH_TENSOR_HADAMARD 83
#define TALSH_TENSOR_KHATRIRAO 84
#define TALSH_OP_UNDEFINED -1
#define TALSH_OP_EMPTY 0
#define TALSH_OP_PARTIAL 1
#define TALSH_OP_DEFINED 2
#define TALSH_OP_RESOURCED 3
#define TALSH_OP_LOADED 4
#define TALSH_OP_SCHEDULED 5
#define TALSH_OP_COMPLETED 6
#define TALSH_OP_STORED 7
#define TALSH_OP_RETIRED 8

/* error codes (talsh.h:25-41) */
#define TALSH_SUCCESS 0
#define TALSH_FAILURE -666
#define TALSH_NOT_AVAILABLE -888
#define TALSH_NOT_IMPLEMENTED -999
#define TALSH_NOT_INITIALIZED 1000000
#define TALSH_ALREADY_INITIALIZED 1000001
#define TALSH_INVALID_ARGS 1000002
#define TALSH_INTEGER_OVERFLOW 1000003
#define TALSH_OBJECT_NOT_EMPTY 1000004
#define TALSH_OBJECT_IS_EMPTY 1000005
#define TALSH_IN_PROGRESS 1000006
#define TALSH_NOT_ALLOWED 1000007
#define TALSH_LIMIT_EXCEEDED 1000008
#define TALSH_NOT_FOUND 1000009
#define TALSH_OBJECT_BROKEN 1000010
#define TALSH_INVALID_REQUEST 1000011

/* task status (talsh.h:43-50) */
#define TALSH_TASK_ERROR 1999999
#define TALSH_TASK_EMPTY 2000000
#define TALSH_TASK_SCHEDULED 2000001
#define TALSH_TASK_STARTED 2000002
#define TALSH_TASK_INPUT_READY 2000003
#define TALSH_TASK_OUTPUT_READY 2000004
#define TALSH_TASK_COMPLETED 2000005

/* Dense tensor with up to TALSH_MAX_DEV_PRESENT body images (talsh.h:87-94, 40 B). */
typedef struct {
  talsh_tens_shape_t *shape_p;
  talsh_dev_rsc_t *dev_rsc;
  int *data_kind;
  int *avail;
  int dev_rsc_len;
  int ndev;
} talsh_tens_t;

/* Tensor operation argument (talsh.h:104-107, 16 B). */
typedef struct {
  talsh_tens_t *tens_p;
  int source_image;
} talshTensArg_t;

/* Task handle (talsh.h:110-121, 120 B); task_p points at a HIP task object. */
typedef struct {
  void *task_p;
  int task_error;
  int dev_kind;
  int data_kind;
  int coherence;
  int num_args;
  talshTensArg_t tens_args[MAX_TENSOR_OPERANDS];
  double data_vol;
  double flops;
  double exec_time;
} talsh_task_t;

/* Tensor slice view (talsh.h:96-101): a box of an existing tensor, not owning it. */
typedef struct {
  talsh_tens_t *tensor;         /* non-owning */
  talsh_tens_signature_t bases; /* base offsets of the box inside the tensor */
  talsh_tens_shape_t shape;     /* extents of the box */
} talsh_tens_slice_t;

/* Basic tensor operation (talsh.h:124-140): slices of existing tensors as formal operands, the actual
 * operands built for execution, and the stage it has reached (TALSH_OP_*). */
typedef struct {
  int opkind;
  int data_kind;
  unsigned int num_args;
  talsh_tens_slice_t tens_slice[MAX_TENSOR_OPERANDS];
  const char *symb_pattern; /* non-owning */
  double alpha_real;
  double alpha_imag;
  talsh_tens_t tens_arg[MAX_TENSOR_OPERANDS];
  talsh_task_t task_handle;
  int exec_dev_id;
  int stage;
  double time_started;
  double time_scheduled;
  double time_completed;
  double time_finished;
} talsh_tens_op_t;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- helpers and control (talsh.h:149-229) ---- */
int talshValidDataKind(int datk, int *datk_size);
int talshInit(size_t *host_buf_size, int *host_arg_max, int ngpus, int gpu_list[], int nmics, int mic_list[],
              int namds, int amd_list[]);
int talshShutdown(void);
int talshDeviceCount(int dev_kind, int *dev_count);
int talshFlatDevId(int dev_kind, int dev_num);
int talshKindDevId(int dev_id, int *dev_kind);
int talshDeviceState(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshDeviceState_(int dev_num, int dev_kind);
int talshDeviceBusyLeast(int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshDeviceBusyLeast_(int dev_kind);
size_t talshDeviceMemorySize(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
size_t talshDeviceMemorySize_(int dev_num, int dev_kind);
size_t talshDeviceBufferSize(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
size_t talshDeviceBufferSize_(int dev_num, int dev_kind);
size_t talshDeviceTensorSize(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
size_t talshDeviceTensorSize_(int dev_num, int dev_kind);
size_t talshDeviceBufferFreeSize(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
size_t talshDeviceBufferFreeSize_(int dev_num, int dev_kind);
double talshDeviceGetFlops(int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int dev_id TALSH_DEFAULT(DEV_DEFAULT));
int talshStats(int dev_id TALSH_DEFAULT(-1), int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshStats_(int dev_id, int dev_kind);
/* Host memory allocation policy (talsh.h:163-165, talshf.F90:882-893): where new Host tensor bodies come from.
 * MEM_ALLOC_REGULAR: ordinary (pageable) memory; MEM_ALLOC_TMP_BUF / MEM_ALLOC_ALL_BUF: the pinned Host buffer
 * given to talshInit, with fallback != 0 letting a full buffer fall back to ordinary memory (otherwise the
 * construction returns TRY_LATER).  ierr = 0 on success.  Defaults: TMP_BUF with fallback, as in the reference. */
void talshSetMemAllocPolicyHost(int mem_policy, int fallback, int *ierr);
/* Fast math (talsh.h:167-175): selects the split-BF16 MFMA GEMM for R4/C4 contractions. */
int talshEnableFastMath(int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int dev_id TALSH_DEFAULT(DEV_DEFAULT));
int talshDisableFastMath(int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int dev_id TALSH_DEFAULT(DEV_DEFAULT));
int talshQueryFastMath(int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int dev_id TALSH_DEFAULT(DEV_DEFAULT));
/* Execution device for given operands (talsh.h:192-194): the device already holding the most operand data;
 * operations run on GPUs only, so the Host is never returned while a GPU is initialised. */
int talshDetermineOptimalDevice(const talsh_tens_t *tens0, const talsh_tens_t *tens1 TALSH_DEFAULT(NULL),
                                const talsh_tens_t *tens2 TALSH_DEFAULT(NULL));
/* Base pointer of a device's buffer (talsh.h:212-214): the GPU pool region or the pinned Host buffer. */
void *talshDeviceBufferBasePtr(int dev_num, int dev_kind TALSH_DEFAULT(DEV_NULL));
void *talshDeviceBufferBasePtr_(int dev_num, int dev_kind);
/* Logging switches (talsh.h:219-225): buffer allocations / tensor operations printed to stdout. */
void talshMemManagerLogStart(void);
void talshMemManagerLogFinish(void);
void talshTensorOpLogStart(void);
void talshTensorOpLogFinish(void);

/* ---- tensor blocks (talsh.h:232-311) ---- */
int talshTensorCreate(talsh_tens_t **tens_block);
int talshTensorClean(talsh_tens_t *tens_block);
int talshTensorIsEmpty(const talsh_tens_t *tens_block);
int talshTensorConstruct(talsh_tens_t *tens_block, int data_kind, int tens_rank, const int tens_dims[],
                         int dev_id TALSH_DEFAULT(0), void *ext_mem TALSH_DEFAULT(NULL),
                         int in_hab TALSH_DEFAULT(-1), talsh_tens_init_i init_method TALSH_DEFAULT(NULL),
                         double init_val_real TALSH_DEFAULT(0.0), double init_val_imag TALSH_DEFAULT(0.0));
int talshTensorConstruct_(talsh_tens_t *tens_block, int data_kind, int tens_rank, const int tens_dims[], int dev_id,
                          void *ext_mem, int in_hab, talsh_tens_init_i init_method, double init_val_real,
                          double init_val_imag);
int talshTensorImportData(talsh_tens_t *tens_block, int data_kind, const void *ext_data);
int talshTensorDestruct(talsh_tens_t *tens_block);
int talshTensorDestroy(talsh_tens_t *tens_block);
int talshTensorRank(const talsh_tens_t *tens_block);
size_t talshTensorVolume(const talsh_tens_t *tens_block);
size_t talshTensorSizeAllImages(const talsh_tens_t *tens_block, int *num_images);
const int *talshTensorDimExtents(const talsh_tens_t *tens_block, int *rank);
int talshTensorShape(const talsh_tens_t *tens_block, talsh_tens_shape_t *tens_shape);
int talshTensorDataKind(const talsh_tens_t *tens_block, int *num_images, int *data_kinds);
int talshTensorReshape(talsh_tens_t *tens_block, int tens_rank, const int tens_dims[]);
int talshTensorInUse(const talsh_tens_t *tens_block);
int talshTensorPresence(const talsh_tens_t *tens_block, int *ncopies, int copies[], int data_kinds[],
                        int dev_kind TALSH_DEFAULT(DEV_NULL), int dev_id TALSH_DEFAULT(-1));
int talshTensorPresence_(const talsh_tens_t *tens_block, int *ncopies, int copies[], int data_kinds[],
                         int dev_kind, int dev_id);
int talshTensorGetBodyAccess(talsh_tens_t *tens_block, void **body_p, int data_kind, int dev_id,
                             int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshTensorGetBodyAccess_(talsh_tens_t *tens_block, void **body_p, int data_kind, int dev_id, int dev_kind);
int talshTensorGetBodyAccessConst(const talsh_tens_t *tens_block, const void **body_p, int data_kind, int dev_id,
                                  int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshTensorGetScalar(talsh_tens_t *tens_block, double *scalar_real, double *scalar_imag);
void talshTensorPrint(const talsh_tens_t *tens_block);
void talshTensorPrintInfo(const talsh_tens_t *tens_block);
/* Elements of the Host image with absolute value >= thresh, with their multi-indices (talsh.h:311). */
void talshTensorPrintBody(const talsh_tens_t *tens_block, double thresh);

/* ---- tasks (talsh.h:332-372) ---- */
int talshTaskCreate(talsh_task_t **talsh_task);
int talshTaskClean(talsh_task_t *talsh_task);
int talshTaskIsEmpty(const talsh_task_t *talsh_task);
int talshTaskDestruct(talsh_task_t *talsh_task);
int talshTaskDestroy(talsh_task_t *talsh_task);
int talshTaskDevId(talsh_task_t *talsh_task, int *dev_kind TALSH_DEFAULT(NULL));
int talshTaskDevId_(talsh_task_t *talsh_task, int *dev_kind);
int talshTaskArgCoherence(const talsh_task_t *talsh_task);
const talshTensArg_t *talshTaskTensArgs(const talsh_task_t *talsh_task, int *num_args);
int talshTaskStatus(talsh_task_t *talsh_task);
int talshTaskComplete(talsh_task_t *talsh_task, int *stats, int *ierr);
int talshTaskWait(talsh_task_t *talsh_task, int *stats);
int talshTasksWait(int ntasks, talsh_task_t talsh_tasks[], int stats[]);
int talshTaskTime(talsh_task_t *talsh_task, double *total, double *comput TALSH_DEFAULT(NULL),
                  double *input TALSH_DEFAULT(NULL), double *output TALSH_DEFAULT(NULL),
                  double *mmul TALSH_DEFAULT(NULL));
int talshTaskTime_(talsh_task_t *talsh_task, double *total, double *comput, double *input, double *output,
                   double *mmul);
void talshTaskPrint(const talsh_task_t *talsh_task);
/* Tensor slice views (talsh.h:312-329). */
int talshTensorSliceCreate(talsh_tens_slice_t **slice);
int talshTensorSliceClean(talsh_tens_slice_t *slice);
int talshTensorSliceConstruct(talsh_tens_slice_t *slice, const talsh_tens_t *tensor, const size_t *offsets,
                              const int *dims, const int *divs TALSH_DEFAULT(NULL), const int *grps TALSH_DEFAULT(NULL));
size_t talshTensorSliceVolume(const talsh_tens_slice_t *slice);
int talshTensorSliceDestruct(talsh_tens_slice_t *slice);
int talshTensorSliceDestroy(talsh_tens_slice_t *slice);
/* Tensor operations (talsh.h:373-431): a contraction of tensor slices staged as
 * DEFINED -> RESOURCED (operand tensors on the execution GPU) -> LOADED (input boxes copied in, from a GPU
 * image by the slice kernel or from the Host image by strided DMA) -> SCHEDULED (talshTensorContract, async)
 * -> COMPLETED -> STORED (result box accumulated into the destination tensor's image) -> RETIRED.
 * Only TALSH_TENSOR_CONTRACT executes, as in the reference. */
int talshTensorOpCreate(talsh_tens_op_t **tens_op);
int talshTensorOpClean(talsh_tens_op_t *tens_op);
int talshTensorOpSetArgument(talsh_tens_op_t *tens_op, const talsh_tens_t *tensor, const size_t *offsets,
                             const int *dims);
int talshTensorOpSpecify(talsh_tens_op_t *tens_op, int operation_kind, int data_kind,
                         const char *symbolic_pattern TALSH_DEFAULT(NULL), double prefactor_real TALSH_DEFAULT(1.0),
                         double prefactor_imag TALSH_DEFAULT(0.0));
int talshTensorOpSetExecDevice(talsh_tens_op_t *tens_op, int dev_id, int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorOpActivate(talsh_tens_op_t *tens_op);
int talshTensorOpLoadInput(talsh_tens_op_t *tens_op);
int talshTensorOpExecute(talsh_tens_op_t *tens_op, int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                         int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorOpTest(talsh_tens_op_t *tens_op, int *completed, int wait TALSH_DEFAULT(NOPE));
int talshTensorOpStoreOutput(talsh_tens_op_t *tens_op);
int talshTensorOpDeactivate(talsh_tens_op_t *tens_op);
int talshTensorOpDestruct(talsh_tens_op_t *tens_op);
int talshTensorOpDestroy(talsh_tens_op_t *tens_op);
int talshTensorOpProgress(talsh_tens_op_t *tens_op, int *done);
size_t talshTensorOpGetArgVolume(const talsh_tens_op_t *tens_op, unsigned int arg_num);
size_t talshTensorOpGetArgSize(const talsh_tens_op_t *tens_op, unsigned int arg_num);
double talshTensorOpGetByteCount(const talsh_tens_op_t *tens_op, unsigned int element_size TALSH_DEFAULT(1));
double talshTensorOpGetFlopCount(const talsh_tens_op_t *tens_op);
double talshTensorOpGetIntensity(const talsh_tens_op_t *tens_op);
/* Splits a contraction in two (talsh.h:427-430): halves the largest extent of the batch (hyper) indices if
 * any, else of the largest of the right, left and contracted index groups (by volume, in that order of
 * preference), as the reference does (talshc.cpp:3150-3317), so callers get the same sub-operations. */
int talshTensorOpDecompose2(const talsh_tens_op_t *tens_op, talsh_tens_op_t *child_op1, talsh_tens_op_t *child_op2);
void talshTensorOpPrint(const talsh_tens_op_t *tens_op);

/* ---- placement (talsh.h:437-453) ---- */
int talshTensorPlace(talsh_tens_t *tens, int dev_id, int dev_kind TALSH_DEFAULT(DEV_NULL),
                     void *dev_mem TALSH_DEFAULT(NULL), int copy_ctrl TALSH_DEFAULT(COPY_M),
                     talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorPlace_(talsh_tens_t *tens, int dev_id, int dev_kind, void *dev_mem, int copy_ctrl,
                      talsh_task_t *talsh_task);
int talshTensorDiscard(talsh_tens_t *tens, int dev_id, int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshTensorDiscard_(talsh_tens_t *tens, int dev_id, int dev_kind);
int talshTensorDiscardOther(talsh_tens_t *tens, int dev_id, int dev_kind TALSH_DEFAULT(DEV_NULL));
int talshTensorDiscardOther_(talsh_tens_t *tens, int dev_id, int dev_kind);

/* ---- tensor operations (talsh.h:455-530) ---- */
int talshTensorInit(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                    int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int copy_ctrl TALSH_DEFAULT(COPY_M),
                    talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorInit_(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *talsh_task);
int talshTensorScale(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                     int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int copy_ctrl TALSH_DEFAULT(COPY_M),
                     talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorScale_(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind,
                      int copy_ctrl, talsh_task_t *talsh_task);
/* Slice: dtens (+)= block of ltens at 0-based offsets; insert: block of dtens at offsets (+)= ltens
 * (talsh.h:472-493; both tensors keep their dimension order, ranks equal). */
int talshTensorSlice(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets,
                     int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT),
                     int copy_ctrl TALSH_DEFAULT(COPY_MT), int accumulative TALSH_DEFAULT(NOPE),
                     talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorSlice_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                      int copy_ctrl, int accumulative, talsh_task_t *talsh_task);
int talshTensorInsert(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets,
                      int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT),
                      int copy_ctrl TALSH_DEFAULT(COPY_MT), int accumulative TALSH_DEFAULT(NOPE),
                      talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorInsert_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                       int copy_ctrl, int accumulative, talsh_task_t *talsh_task);
int talshTensorCopy(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens,
                    int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT),
                    int copy_ctrl TALSH_DEFAULT(COPY_MT), talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorCopy_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *talsh_task);
int talshTensorAdd(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double scale_real TALSH_DEFAULT(1.0),
                   double scale_imag TALSH_DEFAULT(0.0), int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                   int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int copy_ctrl TALSH_DEFAULT(COPY_MT),
                   talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorAdd_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double scale_real,
                    double scale_imag, int dev_id, int dev_kind, int copy_ctrl, talsh_task_t *talsh_task);
int talshTensorContract(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                        double scale_real TALSH_DEFAULT(1.0), double scale_imag TALSH_DEFAULT(0.0),
                        int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT),
                        int copy_ctrl TALSH_DEFAULT(COPY_MTT), int accumulative TALSH_DEFAULT(YEP),
                        talsh_task_t *talsh_task TALSH_DEFAULT(NULL));
int talshTensorContract_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                         double scale_real, double scale_imag, int dev_id, int dev_kind, int copy_ctrl,
                         int accumulative, talsh_task_t *talsh_task);
/* Extra-large contraction (talsh.h:531-542): blocking; the operation is cut into pieces that fit the GPU's
 * device buffer (or the limit set by talshAmdSetContractXLMemory), operands may live on the host; on return
 * dtens holds one image, on the host.  DEV_HOST execution returns TALSH_NOT_AVAILABLE. */
int talshTensorContractXL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                          double scale_real TALSH_DEFAULT(1.0), double scale_imag TALSH_DEFAULT(0.0),
                          int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT),
                          int accumulative TALSH_DEFAULT(YEP));
int talshTensorContractXL_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                           double scale_real, double scale_imag, int dev_id, int dev_kind, int accumulative);
/* Squared 2-norm into a rank-0 tensor dtens (TALSH_manual.txt:893-913; see the header comment). */
int talshTensorNorm2(talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                     int dev_kind TALSH_DEFAULT(DEV_DEFAULT), int copy_ctrl TALSH_DEFAULT(COPY_TT),
                     talsh_task_t *talsh_task TALSH_DEFAULT(NULL));

/* ---- decompositions (talsh.h:540-592): outside the contraction path; these validate their arguments and
 * return TALSH_NOT_IMPLEMENTED (the reference computes them on the Host with LAPACK). ---- */
int talshTensorDecomposeSVD(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                            talsh_tens_t *stens, const char absorb TALSH_DEFAULT('N'),
                            int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorDecomposeSVDL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorDecomposeSVDR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorDecomposeSVDLR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                              int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorOrthogonalizeSVD(const char *cptrn, talsh_tens_t *dtens, int dev_id TALSH_DEFAULT(DEV_DEFAULT),
                                int dev_kind TALSH_DEFAULT(DEV_DEFAULT));
int talshTensorOrthogonalizeMGS(talsh_tens_t *dtens, int num_iso_dims, int *iso_dims,
                                int dev_id TALSH_DEFAULT(DEV_DEFAULT), int dev_kind TALSH_DEFAULT(DEV_DEFAULT));

/* ---- debugging (talsh.h:593) ---- */
double talshTensorImageNorm1_cpu(const talsh_tens_t *talsh_tens);

#ifdef __cplusplus
}
#endif

#endif /* TALSH_H_ */
