// Tensor slice views and tensor operations of the C API
// (talshc.cpp:1855-1935 slices, :2599-3317 operations).
//
// An operation names boxes of existing tensors as its formal operands.  Its
// actual operands are built on the execution GPU (the reference builds them
// in Host memory and slices there): activation takes device buffers from the
// pool, loading copies each input box in -- by the slice kernel when the
// source tensor has an image on that GPU, by strided DMA from its Host image
// otherwise -- execution is an asynchronous talshTensorContract on the
// resident operands, and storing accumulates the result box into the
// destination tensor's image (GPU: insert kernel; Host: box in, add on the
// GPU, box out), after which the destination keeps that image only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "contract.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "talsh_amd.h"

using namespace tal;

namespace {

size_t shape_vol(const talsh_tens_shape_t &s) {
  if (s.num_dim < 0) return 0;
  size_t v = 1;
  for (int i = 0; i < s.num_dim; ++i) v *= (size_t)s.dims[i];
  return v;
}

// Execution GPU of an operation: the preset one, else the GPU that holds its
// operands (or the least busy one).
int op_gpu(const talsh_tens_op_t *op, int *g) {
  if (op->exec_dev_id != DEV_NULL) {
    const int gi = gpu_index(op->exec_dev_id);
    if (gi < 0) return TALSH_NOT_AVAILABLE;
    *g = gi;
    return TALSH_SUCCESS;
  }
  const talsh_tens_t *t[3] = {nullptr, nullptr, nullptr};
  for (unsigned i = 0; i < op->num_args && i < 3; ++i) t[i] = op->tens_slice[i].tensor;
  return exec_gpu(DEV_DEFAULT, DEV_DEFAULT, {t[0], t[1], t[2]}, g);
}

std::vector<int> ivec(const int *p, int n) { return std::vector<int>(p, p + (n > 0 ? n : 0)); }
std::vector<int> offsets_of(const talsh_tens_slice_t &s) {
  std::vector<int> o(s.bases.num_dim > 0 ? s.bases.num_dim : 0);
  for (size_t i = 0; i < o.size(); ++i) o[i] = (int)s.bases.offsets[i];
  return o;
}

// Body of the operand tensor built on GPU g.
void *arg_body(const talsh_tens_t &t, int g) {
  const int i = image_find(&t, flat_gpu(g), NO_TYPE, true);
  return i >= 0 ? t.dev_rsc[i].gmem_p : nullptr;
}

// Copies the box of `s` into the packed device buffer dst on GPU g (blocking).
int load_box(const talsh_tens_slice_t &s, int dk, void *dst, int g) {
  talsh_tens_t *big = s.tensor;
  Device &dev = rt().gpus[g];
  hipStream_t st = dev.streams[0];
  const int rank = big->shape_p->num_dim;
  std::vector<int> bd = ivec(big->shape_p->dims, rank), sd = ivec(s.shape.dims, rank), off = offsets_of(s);
  int i = image_find(big, flat_gpu(g), dk, true);
  if (i >= 0) {
    hazard_wait_read(big->dev_rsc[i].gmem_p, image_bytes(big, i), st);
    if (launch_slice(dk, 0, big->dev_rsc[i].gmem_p, dst, rank, sd.data(), bd.data(), off.data(), 0, st))
      return TALSH_FAILURE;
  } else {
    i = image_find(big, DEV_HOST, dk, true);
    if (i < 0) {  // only on another GPU: bring a Host copy first
      if (int e = talshTensorPlace(big, 0, DEV_HOST, nullptr, COPY_K, nullptr)) return e;
      i = image_find(big, DEV_HOST, dk, true);
      if (i < 0) return TALSH_NOT_AVAILABLE;
    }
    hazard_drain(big->dev_rsc[i].gmem_p, image_bytes(big, i));
    if (box_dma(true, (char *)big->dev_rsc[i].gmem_p, (char *)dst, rank, bd.data(), off.data(), sd.data(),
                (size_t)dk_size(dk), st))
      return TALSH_FAILURE;
  }
  return hipStreamSynchronize(st) == hipSuccess ? TALSH_SUCCESS : TALSH_FAILURE;
}

// big[box] += src (packed device buffer on GPU g), then big keeps only the
// image that was written (blocking).
int store_box(const talsh_tens_slice_t &s, int dk, const void *src, int g) {
  talsh_tens_t *big = s.tensor;
  Device &dev = rt().gpus[g];
  hipStream_t st = dev.streams[0];
  const int rank = big->shape_p->num_dim;
  std::vector<int> bd = ivec(big->shape_p->dims, rank), sd = ivec(s.shape.dims, rank), off = offsets_of(s);
  int i = image_find(big, flat_gpu(g), dk, true);
  if (i >= 0) {
    void *p = big->dev_rsc[i].gmem_p;
    hazard_wait_write(p, image_bytes(big, i), st);
    if (launch_slice(dk, 1, src, p, rank, sd.data(), bd.data(), off.data(), 1, st)) return TALSH_FAILURE;
    if (hipStreamSynchronize(st) != hipSuccess) return TALSH_FAILURE;
    image_discard_other(big, image_find(big, flat_gpu(g), dk, true));
    return TALSH_SUCCESS;
  }
  i = image_find(big, DEV_HOST, dk, true);
  if (i < 0) {
    if (int e = talshTensorPlace(big, 0, DEV_HOST, nullptr, COPY_K, nullptr)) return e;
    i = image_find(big, DEV_HOST, dk, true);
    if (i < 0) return TALSH_NOT_AVAILABLE;
  }
  char *host = (char *)big->dev_rsc[i].gmem_p;
  hazard_drain(host, image_bytes(big, i));
  const size_t esz = (size_t)dk_size(dk), bytes = shape_vol(s.shape) * esz;
  void *tmp = nullptr;
  int h = -1;
  if (int e = dev_alloc(g, bytes, &tmp, &h)) return e;
  std::vector<int> o2n(rank);
  for (int k = 0; k < rank; ++k) o2n[k] = k;
  int err = TALSH_SUCCESS;
  if (box_dma(true, host, (char *)tmp, rank, bd.data(), off.data(), sd.data(), esz, st) ||
      launch_permute(dk, src, tmp, rank, sd.data(), o2n.data(), PERM_ADD, 1.0, 0.0, 0, st) ||
      box_dma(false, host, (char *)tmp, rank, bd.data(), off.data(), sd.data(), esz, st))
    err = TALSH_FAILURE;
  if (hipStreamSynchronize(st) != hipSuccess) err = TALSH_FAILURE;
  dev_free(g, tmp, h);
  if (err == TALSH_SUCCESS) image_discard_other(big, image_find(big, DEV_HOST, dk, true));
  return err;
}

}  // namespace

// ------------------------------------------------------------------ slices
int talshTensorSliceCreate(talsh_tens_slice_t **slice) {
  if (!slice) return TALSH_INVALID_ARGS;
  *slice = (talsh_tens_slice_t *)std::malloc(sizeof(talsh_tens_slice_t));
  if (!*slice) return TRY_LATER;
  return talshTensorSliceClean(*slice);
}

int talshTensorSliceClean(talsh_tens_slice_t *slice) {
  if (!slice) return TALSH_INVALID_ARGS;
  slice->tensor = nullptr;
  slice->bases.num_dim = -1;
  slice->bases.offsets = nullptr;
  slice->shape.num_dim = -1;
  slice->shape.dims = slice->shape.divs = slice->shape.grps = nullptr;
  return TALSH_SUCCESS;
}

int talshTensorSliceConstruct(talsh_tens_slice_t *slice, const talsh_tens_t *tensor, const size_t *offsets,
                              const int *dims, const int *divs, const int *grps) {
  (void)divs, (void)grps;  // dimension divisions / groups: not used by dense dlf slices
  if (!slice || !tensor || talshTensorIsEmpty(tensor) == YEP) return TALSH_INVALID_ARGS;
  const int rank = tensor->shape_p->num_dim;
  if (rank < 0 || (rank > 0 && (!offsets || !dims))) return TALSH_INVALID_ARGS;
  for (int i = 0; i < rank; ++i)
    if (dims[i] <= 0 || offsets[i] + (size_t)dims[i] > (size_t)tensor->shape_p->dims[i]) return TALSH_INVALID_ARGS;
  talshTensorSliceClean(slice);
  if (rank > 0) {
    slice->bases.offsets = (size_t *)std::malloc(sizeof(size_t) * rank);
    slice->shape.dims = (int *)std::malloc(sizeof(int) * rank);
    if (!slice->bases.offsets || !slice->shape.dims) {
      talshTensorSliceDestruct(slice);
      return TRY_LATER;
    }
    std::memcpy(slice->bases.offsets, offsets, sizeof(size_t) * rank);
    std::memcpy(slice->shape.dims, dims, sizeof(int) * rank);
  }
  slice->bases.num_dim = rank;
  slice->shape.num_dim = rank;
  slice->tensor = const_cast<talsh_tens_t *>(tensor);
  return TALSH_SUCCESS;
}

size_t talshTensorSliceVolume(const talsh_tens_slice_t *slice) { return slice ? shape_vol(slice->shape) : 0; }

int talshTensorSliceDestruct(talsh_tens_slice_t *slice) {
  if (!slice) return TALSH_INVALID_ARGS;
  std::free(slice->bases.offsets);
  std::free(slice->shape.dims);
  return talshTensorSliceClean(slice);
}

int talshTensorSliceDestroy(talsh_tens_slice_t *slice) {
  if (!slice) return TALSH_INVALID_ARGS;
  const int e = talshTensorSliceDestruct(slice);
  std::free(slice);
  return e;
}

// ------------------------------------------------------------ operations
int talshTensorOpCreate(talsh_tens_op_t **tens_op) {
  if (!tens_op) return TALSH_INVALID_ARGS;
  *tens_op = (talsh_tens_op_t *)std::malloc(sizeof(talsh_tens_op_t));
  if (!*tens_op) return TRY_LATER;
  return talshTensorOpClean(*tens_op);
}

int talshTensorOpClean(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  op->opkind = TALSH_TENSOR_NOOP;
  op->data_kind = NO_TYPE;
  op->num_args = 0;
  op->symb_pattern = nullptr;
  op->alpha_real = op->alpha_imag = 0.0;
  op->exec_dev_id = DEV_NULL;
  op->time_started = op->time_scheduled = op->time_completed = op->time_finished = -1.0;
  talshTaskClean(&op->task_handle);
  for (int i = 0; i < MAX_TENSOR_OPERANDS; ++i) {
    talshTensorSliceClean(&op->tens_slice[i]);
    talshTensorClean(&op->tens_arg[i]);
  }
  op->stage = TALSH_OP_EMPTY;
  return TALSH_SUCCESS;
}

int talshTensorOpSetArgument(talsh_tens_op_t *op, const talsh_tens_t *tensor, const size_t *offsets,
                             const int *dims) {
  if (!op || !tensor) return TALSH_INVALID_ARGS;
  if (op->opkind != TALSH_TENSOR_NOOP) return TALSH_NOT_ALLOWED;
  if (op->num_args >= MAX_TENSOR_OPERANDS) return TALSH_LIMIT_EXCEEDED;
  if (int e = talshTensorSliceConstruct(&op->tens_slice[op->num_args], tensor, offsets, dims)) return e;
  op->num_args++;
  op->stage = TALSH_OP_PARTIAL;
  return TALSH_SUCCESS;
}

int talshTensorOpSpecify(talsh_tens_op_t *op, int operation_kind, int data_kind, const char *symbolic_pattern,
                         double prefactor_real, double prefactor_imag) {
  int dks = 0;
  if (!op || talshValidDataKind(data_kind, &dks) != YEP) return TALSH_INVALID_ARGS;
  if (op->opkind != TALSH_TENSOR_NOOP) return TALSH_NOT_ALLOWED;
  op->symb_pattern = symbolic_pattern;
  op->alpha_real = prefactor_real;
  op->alpha_imag = prefactor_imag;
  op->data_kind = data_kind;
  op->opkind = operation_kind;
  op->stage = TALSH_OP_DEFINED;
  return TALSH_SUCCESS;
}

int talshTensorOpSetExecDevice(talsh_tens_op_t *op, int dev_id, int dev_kind) {
  if (!op || dev_id < 0) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_DEFINED) return TALSH_NOT_ALLOWED;
  op->exec_dev_id = dev_kind == DEV_DEFAULT ? dev_id : talshFlatDevId(dev_kind, dev_id);
  return TALSH_SUCCESS;
}

// DEFINED -> RESOURCED: the actual operands, on the execution GPU (bodies left
// as allocated; inputs are loaded, the output is overwritten by the contraction).
int talshTensorOpActivate(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  if (op->opkind == TALSH_TENSOR_NOOP || op->stage != TALSH_OP_DEFINED) return TALSH_NOT_ALLOWED;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  int g;
  if (int e = op_gpu(op, &g)) return e;
  op->time_started = wall_time();
  for (unsigned i = 0; i < op->num_args; ++i) {
    const talsh_tens_slice_t &s = op->tens_slice[i];
    talshTensorClean(&op->tens_arg[i]);
    const int e = talshTensorConstruct(&op->tens_arg[i], op->data_kind, s.shape.num_dim, s.shape.dims, flat_gpu(g),
                                       nullptr, 0, talsh_tens_no_init, 0.0, 0.0);
    if (e) {
      for (unsigned j = 0; j <= i; ++j)
        if (talshTensorIsEmpty(&op->tens_arg[j]) == NOPE) talshTensorDestruct(&op->tens_arg[j]);
      return e;
    }
  }
  if (op->exec_dev_id == DEV_NULL) op->exec_dev_id = flat_gpu(g);
  op->stage = TALSH_OP_RESOURCED;
  return TALSH_SUCCESS;
}

// RESOURCED -> LOADED: input boxes copied into the actual operands (blocking).
int talshTensorOpLoadInput(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_RESOURCED) return TALSH_NOT_ALLOWED;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  const int g = gpu_index(op->exec_dev_id);
  if (g < 0) return TALSH_NOT_AVAILABLE;
  for (unsigned i = 1; i < op->num_args; ++i) {
    if (op->tens_slice[i].tensor->shape_p->num_dim != op->tens_slice[i].shape.num_dim) return TALSH_OBJECT_BROKEN;
    void *dst = arg_body(op->tens_arg[i], g);
    if (!dst) return TALSH_OBJECT_BROKEN;
    if (int e = load_box(op->tens_slice[i], op->data_kind, dst, g)) return e;
  }
  op->stage = TALSH_OP_LOADED;
  return TALSH_SUCCESS;
}

// LOADED -> SCHEDULED: the contraction of the actual operands, asynchronous.
int talshTensorOpExecute(talsh_tens_op_t *op, int dev_id, int dev_kind) {
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_LOADED) return TALSH_NOT_ALLOWED;
  if (dev_id != DEV_DEFAULT || dev_kind != DEV_DEFAULT) {
    const int want = dev_kind == DEV_DEFAULT ? dev_id : talshFlatDevId(dev_kind, dev_id);
    if (canonical_dev(want) != canonical_dev(op->exec_dev_id)) return TALSH_INVALID_ARGS;  // operands live there
  }
  if (op->opkind != TALSH_TENSOR_CONTRACT) return TALSH_NOT_IMPLEMENTED;
  if (op->num_args != 3 || !op->symb_pattern) return TALSH_INVALID_ARGS;
  op->time_scheduled = wall_time();
  const int e = talshTensorContract(op->symb_pattern, &op->tens_arg[0], &op->tens_arg[1], &op->tens_arg[2],
                                    op->alpha_real, op->alpha_imag, op->exec_dev_id, DEV_DEFAULT, COPY_MMM, NOPE,
                                    &op->task_handle);
  if (e == TALSH_SUCCESS) op->stage = TALSH_OP_SCHEDULED;
  return e;
}

int talshTensorOpTest(talsh_tens_op_t *op, int *completed, int wait) {
  if (!op || !completed) return TALSH_INVALID_ARGS;
  *completed = NOPE;
  if (op->stage != TALSH_OP_SCHEDULED) return TALSH_NOT_ALLOWED;
  int sts = TALSH_TASK_EMPTY, errc = TALSH_SUCCESS;
  if (wait == YEP) {
    errc = talshTaskWait(&op->task_handle, &sts);
    if (errc == TALSH_SUCCESS && sts == TALSH_TASK_COMPLETED) *completed = YEP;
  } else {
    const int ans = talshTaskComplete(&op->task_handle, &sts, &errc);
    if (errc == TALSH_SUCCESS && ans == YEP && sts == TALSH_TASK_COMPLETED) *completed = YEP;
  }
  if (errc == TALSH_SUCCESS && *completed == YEP) {
    op->time_completed = wall_time();
    op->stage = TALSH_OP_COMPLETED;
  }
  return errc;
}

// COMPLETED -> STORED: the result box accumulated into the destination (blocking).
int talshTensorOpStoreOutput(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_COMPLETED) return TALSH_NOT_ALLOWED;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (op->num_args > 0) {
    const int g = gpu_index(op->exec_dev_id);
    if (g < 0) return TALSH_NOT_AVAILABLE;
    if (op->tens_slice[0].tensor->shape_p->num_dim != op->tens_slice[0].shape.num_dim) return TALSH_OBJECT_BROKEN;
    const void *src = arg_body(op->tens_arg[0], g);
    if (!src) return TALSH_OBJECT_BROKEN;
    if (int e = store_box(op->tens_slice[0], op->data_kind, src, g)) return e;
  }
  op->stage = TALSH_OP_STORED;
  return TALSH_SUCCESS;
}

int talshTensorOpDeactivate(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  if (op->stage != TALSH_OP_RESOURCED && op->stage != TALSH_OP_STORED) return TALSH_NOT_ALLOWED;
  int errc = TALSH_SUCCESS;
  if (op->stage == TALSH_OP_STORED && talshTaskDestruct(&op->task_handle) != TALSH_SUCCESS) errc = NOT_CLEAN;
  for (int i = (int)op->num_args - 1; i >= 0; --i)
    if (talshTensorIsEmpty(&op->tens_arg[i]) == NOPE && talshTensorDestruct(&op->tens_arg[i]) != TALSH_SUCCESS)
      return TALSH_FAILURE;
  op->stage = op->stage == TALSH_OP_STORED ? TALSH_OP_RETIRED : TALSH_OP_DEFINED;
  op->time_finished = wall_time();
  return errc;
}

int talshTensorOpDestruct(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  int stat = talshTaskStatus(&op->task_handle);
  if (stat == TALSH_TASK_COMPLETED || stat == TALSH_TASK_ERROR) {
    talshTaskDestruct(&op->task_handle);
    stat = talshTaskStatus(&op->task_handle);
  }
  if (stat != TALSH_TASK_EMPTY) {
    op->stage = TALSH_OP_UNDEFINED;
    return TALSH_IN_PROGRESS;  // still executing
  }
  for (unsigned i = 0; i < op->num_args; ++i) {
    if (talshTensorIsEmpty(&op->tens_arg[i]) == NOPE) talshTensorDestruct(&op->tens_arg[i]);
    talshTensorSliceDestruct(&op->tens_slice[i]);
  }
  return talshTensorOpClean(op);
}

int talshTensorOpDestroy(talsh_tens_op_t *op) {
  if (!op) return TALSH_INVALID_ARGS;
  const int e = talshTensorOpDestruct(op);
  std::free(op);
  return e;
}

// Advances the operation as far as it can go without waiting (talshc.cpp:2929-3050).
int talshTensorOpProgress(talsh_tens_op_t *op, int *done) {
  if (!done) return TALSH_INVALID_ARGS;
  *done = NOPE;
  if (!op) return TALSH_INVALID_ARGS;
  while (true) {
    int e = TALSH_SUCCESS;
    switch (op->stage) {
      case TALSH_OP_DEFINED: e = talshTensorOpActivate(op); break;
      case TALSH_OP_RESOURCED: e = talshTensorOpLoadInput(op); break;
      case TALSH_OP_LOADED: return talshTensorOpExecute(op);  // yields after scheduling
      case TALSH_OP_SCHEDULED: {
        int completed = NOPE;
        e = talshTensorOpTest(op, &completed, NOPE);
        if (e == TALSH_SUCCESS && completed != YEP) return TALSH_SUCCESS;
        break;
      }
      case TALSH_OP_COMPLETED: e = talshTensorOpStoreOutput(op); break;
      case TALSH_OP_STORED: e = talshTensorOpDeactivate(op); break;
      case TALSH_OP_RETIRED: *done = YEP; return TALSH_SUCCESS;
      default: return TALSH_NOT_ALLOWED;
    }
    if (e != TALSH_SUCCESS) return e;
  }
}

size_t talshTensorOpGetArgVolume(const talsh_tens_op_t *op, unsigned int arg_num) {
  return (op && arg_num < op->num_args) ? talshTensorSliceVolume(&op->tens_slice[arg_num]) : 0;
}

size_t talshTensorOpGetArgSize(const talsh_tens_op_t *op, unsigned int arg_num) {
  int dks = 0;
  if (!op || arg_num >= op->num_args || op->data_kind == NO_TYPE || talshValidDataKind(op->data_kind, &dks) != YEP)
    return 0;
  return talshTensorOpGetArgVolume(op, arg_num) * (size_t)dks;
}

double talshTensorOpGetByteCount(const talsh_tens_op_t *op, unsigned int element_size) {
  if (!op || op->opkind == TALSH_TENSOR_NOOP) return 0.0;
  int dks = (int)element_size;
  if (op->data_kind != NO_TYPE && talshValidDataKind(op->data_kind, &dks) != YEP) return 0.0;
  double bytes = 0.0;
  for (unsigned i = 0; i < op->num_args; ++i) bytes += (double)talshTensorSliceVolume(&op->tens_slice[i]) * dks;
  return bytes;
}

// FLOP of a contraction: each distinct uncontracted and contracted index once,
// times 2 (real) or 8 (complex) (talshc.cpp:3092-3140).
double talshTensorOpGetFlopCount(const talsh_tens_op_t *op) {
  if (!op || op->opkind != TALSH_TENSOR_CONTRACT || op->num_args != 3 || !op->symb_pattern) return 0.0;
  double fma;
  if (op->data_kind == R4 || op->data_kind == R8) fma = 2.0;
  else if (op->data_kind == C4 || op->data_kind == C8) fma = 8.0;
  else return 0.0;
  DigitalPattern p;
  if (parse_pattern(op->symb_pattern, &p) != 0) return 0.0;
  if (p.lrank != op->tens_slice[1].shape.num_dim || p.rrank != op->tens_slice[2].shape.num_dim) return 0.0;
  std::vector<bool> seen(p.drank > 0 ? p.drank : 0, false);
  double flops = 1.0;
  for (int i = 0; i < p.lrank; ++i) {
    const int c = p.dig[i];
    const double e = op->tens_slice[1].shape.dims[i];
    if (c < 0) flops *= e;
    else if (c > 0 && !seen[c - 1]) seen[c - 1] = true, flops *= e;
  }
  for (int j = 0; j < p.rrank; ++j) {
    const int c = p.dig[p.lrank + j];
    if (c > 0 && !seen[c - 1]) seen[c - 1] = true, flops *= op->tens_slice[2].shape.dims[j];
  }
  return flops * fma;
}

double talshTensorOpGetIntensity(const talsh_tens_op_t *op) {
  const double f = talshTensorOpGetFlopCount(op), b = talshTensorOpGetByteCount(op);
  return (b <= 0.0 || f < 0.0) ? -1.0 : f / b;
}

// Splits a contraction in two along one index (talshc.cpp:3150-3317): a batch
// (hyper) index if there is one, else an index of the largest of the right,
// left and contracted groups (by volume; ties prefer right, then left); the
// largest extent of that group is halved, the first child taking the larger half.
int talshTensorOpDecompose2(const talsh_tens_op_t *op, talsh_tens_op_t *c1, talsh_tens_op_t *c2) {
  if (!op || !c1 || !c2) return TALSH_INVALID_ARGS;
  if (op->opkind != TALSH_TENSOR_CONTRACT) return TALSH_NOT_IMPLEMENTED;
  if (op->num_args != 3 || !op->symb_pattern) return TALSH_INVALID_ARGS;
  DigitalPattern p;
  if (parse_pattern(op->symb_pattern, &p) != 0) return TALSH_INVALID_ARGS;
  const int nd = p.drank, nl = p.lrank, nr = p.rrank;
  if (nd <= 0 && nl <= 0 && nr <= 0) return TALSH_NOT_ALLOWED;
  if (nl + nr == 0) return TALSH_NOT_ALLOWED;
  const int *ld = op->tens_slice[1].shape.dims, *rd = op->tens_slice[2].shape.dims;
  // group volumes and the position of each group's largest extent (last one on ties)
  std::vector<int> dpos(nd > 0 ? nd : 0, 0);  // D index -> 1 + its L position
  double lb = 1, lr = 1, ll = 1, lc = 1;
  int mb = -1, mr = -1, ml = -1, mc = -1;
  int sb = 0, sr = 0, sl = 0, sc = 0;
  for (int i = 0; i < nl; ++i) {
    const int c = p.dig[i], e = ld[i];
    if (c < 0) { lc *= e; if (e >= sc) { sc = e; mc = i; } }
    else if (c > 0) { dpos[c - 1] = i + 1; ll *= e; if (e >= sl) { sl = e; ml = i; } }
  }
  for (int j = 0; j < nr; ++j) {
    const int c = p.dig[nl + j], e = rd[j];
    if (c <= 0) continue;
    if (dpos[c - 1] == 0) { lr *= e; if (e >= sr) { sr = e; mr = nl + j; } }
    else { ll /= e; lb *= e; if (e >= sb) { sb = e; mb = nl + j; } }
  }
  if (!(lb > 1 || lr > 1 || ll > 1 || lc > 1)) return TALSH_INVALID_REQUEST;
  int d0 = -1, d1 = -1, d2 = -1;  // dimension of D, L, R to halve (-1: none)
  if (lb > 1) {
    d0 = p.dig[mb] - 1; d1 = dpos[d0] - 1; d2 = mb - nl;
  } else if (lr >= ll && lr >= lc) {
    d0 = p.dig[mr] - 1; d2 = mr - nl;
  } else if (ll >= lr && ll >= lc) {
    d0 = p.dig[ml] - 1; d1 = ml;
  } else {
    d1 = mc; d2 = -p.dig[mc] - 1;
  }
  const int split[3] = {d0, d1, d2};
  for (int a = 0; a < 3; ++a) {
    const talsh_tens_slice_t &s = op->tens_slice[a];
    const int rank = s.shape.num_dim;
    std::vector<size_t> off(s.bases.offsets, s.bases.offsets + (rank > 0 ? rank : 0));
    std::vector<int> dims = ivec(s.shape.dims, rank);
    const int k = split[a];
    if (k >= 0) {
      const int h1 = (dims[k] + 1) / 2, h2 = dims[k] - h1;
      dims[k] = h1;
      if (int e = talshTensorOpSetArgument(c1, s.tensor, off.data(), dims.data())) return e;
      dims[k] = h2;
      off[k] += (size_t)h1;
      if (int e = talshTensorOpSetArgument(c2, s.tensor, off.data(), dims.data())) return e;
    } else {
      if (int e = talshTensorOpSetArgument(c1, s.tensor, off.data(), dims.data())) return e;
      if (int e = talshTensorOpSetArgument(c2, s.tensor, off.data(), dims.data())) return e;
    }
  }
  if (int e = talshTensorOpSpecify(c1, op->opkind, op->data_kind, op->symb_pattern, op->alpha_real, op->alpha_imag))
    return e;
  return talshTensorOpSpecify(c2, op->opkind, op->data_kind, op->symb_pattern, op->alpha_real, op->alpha_imag);
}

void talshTensorOpPrint(const talsh_tens_op_t *op) {
  if (!op) return;
  std::printf("#MSG(TAL-SH::tensor_op): kind %d, data kind %d, stage %d, exec device %d, pattern %s\n", op->opkind,
              op->data_kind, op->stage, op->exec_dev_id, op->symb_pattern ? op->symb_pattern : "(none)");
  for (unsigned i = 0; i < op->num_args; ++i) {
    const talsh_tens_slice_t &s = op->tens_slice[i];
    std::printf("  arg %u: tensor %p, box", i, (void *)s.tensor);
    for (int k = 0; k < s.shape.num_dim; ++k)
      std::printf(" [%zu:%zu)", s.bases.offsets[k], s.bases.offsets[k] + (size_t)s.shape.dims[k]);
    std::printf("\n");
  }
  std::fflush(stdout);
}
