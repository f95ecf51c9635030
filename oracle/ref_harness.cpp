// Harness over TAL-SH's own CPU path (CP-TAL), built from the reference
// sources by oracle/Makefile into oracle/_ref/libtalsh_ref.so.
//
// Test infrastructure only: tests/ and bench.py's CPU-baseline leg load it
// through ctypes to (1) pin the C restatement in talsh_oracle.c against the
// reference's own routines and (2) time the reference CPU path.  Every entry
// point wraps caller-owned host buffers as external-memory TAL-SH tensors on
// DEV_HOST and calls the reference's public C API:
//   talshTensorContract  talshc.cpp:5073 (Host branch :5212-5289)
//   talshTensorCopy      talshc.cpp:4530
//   talshTensorAdd       talshc.cpp:4795
//   talshTensorSlice     talshc.cpp:4008 (Host: tensor_block_slice, tensor_algebra_cpu.F90:2315)
//   talshTensorInsert    talshc.cpp:4269 (Host: tensor_block_insert, tensor_algebra_cpu.F90:2466)
//   get_contr_pattern_sym         tensor_algebra_cpu.F90:5533-5611
//   talsh_get_contr_ptrn_str2dig  talshf.F90:619-667
//   get_contr_permutations        tensor_algebra_cpu.F90:5613-5755
//   tensBlck_norm2_host           tensor_algebra_gpu_nvidia.cu:3269-3299
#include "talsh.h"
#include "device_algebra.h"

#include <cstdio>
#include <cstring>

extern "C" int talsh_get_contr_ptrn_str2dig(const char *c_str, int *dig_ptrn, int *drank, int *lrank,
                                            int *rrank, int *conj_bits);

namespace {

int g_ref_on = 0;

// Wraps an external host buffer as a TAL-SH tensor (no copy, no init).
int wrap(talsh_tens_t *t, int dk, int rank, const int *dims, void *body) {
  int err = talshTensorClean(t);
  if (err) return err;
  return talshTensorConstruct(t, dk, rank, dims, talshFlatDevId(DEV_HOST, 0), body, -1, NULL, 0.0, 0.0);
}

}  // namespace

extern "C" {

int ref_init(void) {
  if (g_ref_on) return 0;
  size_t host_buf = TALSH_NO_HOST_BUFFER;
  int host_arg_max = 0;
  int err = talshInit(&host_buf, &host_arg_max, 0, NULL, 0, NULL, 0, NULL);
  if (err == 0) g_ref_on = 1;
  return err;
}

int ref_shutdown(void) {
  if (!g_ref_on) return 0;
  g_ref_on = 0;
  return talshShutdown();
}

int ref_str2dig(const char *ptrn, int *dig, int *drank, int *lrank, int *rrank, int *conj_bits) {
  return talsh_get_contr_ptrn_str2dig(ptrn, dig, drank, lrank, rrank, conj_bits);
}

// TN matricization used by the reference GPU path (gemm_tl=1, gemm_tr=0).
int ref_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn, int conj_bits,
                           int *dprm, int *lprm, int *rprm, int *ncd, int *nlu, int *nru) {
  int ierr = 0;
  get_contr_permutations(gemm_tl, gemm_tr, lrank, rrank, cptrn, conj_bits, dprm, lprm, rprm, ncd, nlu, nru,
                         &ierr);
  return ierr;
}

int ref_copy(const char *ptrn, int dk, int drank, const int *ddims, void *d, int lrank, const int *ldims,
             void *l) {
  talsh_tens_t td, tl;
  int err = wrap(&td, dk, drank, ddims, d);
  if (err) return err;
  err = wrap(&tl, dk, lrank, ldims, l);
  if (err) { talshTensorDestruct(&td); return err; }
  err = talshTensorCopy(ptrn, &td, &tl, 0, DEV_HOST, COPY_MT, NULL);
  talshTensorDestruct(&tl);
  talshTensorDestruct(&td);
  return err;
}

int ref_add(const char *ptrn, int dk, int drank, const int *ddims, void *d, int lrank, const int *ldims, void *l,
            double sr, double si) {
  talsh_tens_t td, tl;
  int err = wrap(&td, dk, drank, ddims, d);
  if (err) return err;
  err = wrap(&tl, dk, lrank, ldims, l);
  if (err) { talshTensorDestruct(&td); return err; }
  err = talshTensorAdd(ptrn, &td, &tl, sr, si, 0, DEV_HOST, COPY_MT, NULL);
  talshTensorDestruct(&tl);
  talshTensorDestruct(&td);
  return err;
}

int ref_contract(const char *ptrn, int dk, int drank, const int *ddims, void *d, int lrank, const int *ldims,
                 void *l, int rrank, const int *rdims, void *r, double sr, double si, int accumulative) {
  talsh_tens_t td, tl, tr;
  int err = wrap(&td, dk, drank, ddims, d);
  if (err) return err;
  err = wrap(&tl, dk, lrank, ldims, l);
  if (err) { talshTensorDestruct(&td); return err; }
  err = wrap(&tr, dk, rrank, rdims, r);
  if (err) { talshTensorDestruct(&tl); talshTensorDestruct(&td); return err; }
  err = talshTensorContract(ptrn, &td, &tl, &tr, sr, si, 0, DEV_HOST, COPY_MTT, accumulative, NULL);
  talshTensorDestruct(&tr);
  talshTensorDestruct(&tl);
  talshTensorDestruct(&td);
  return err;
}

int ref_init_val(int dk, int rank, const int *dims, void *d, double vr, double vi) {
  talsh_tens_t td;
  int err = wrap(&td, dk, rank, dims, d);
  if (err) return err;
  err = talshTensorInit(&td, vr, vi, 0, DEV_HOST, COPY_M, NULL);
  talshTensorDestruct(&td);
  return err;
}

int ref_scale(int dk, int rank, const int *dims, void *d, double vr, double vi) {
  talsh_tens_t td;
  int err = wrap(&td, dk, rank, dims, d);
  if (err) return err;
  err = talshTensorScale(&td, vr, vi, 0, DEV_HOST, COPY_M, NULL);
  talshTensorDestruct(&td);
  return err;
}

// s (+)= b[offsets + i] (slice) or b[offsets + i] (+)= s (insert).
static int slice_or_insert(int insert, int dk, int rank, const int *sdims, void *s, const int *bdims, void *b,
                           const int *offsets, int accumulative) {
  talsh_tens_t ts, tb;
  int err = wrap(&ts, dk, rank, sdims, s);
  if (err) return err;
  err = wrap(&tb, dk, rank, bdims, b);
  if (err) { talshTensorDestruct(&ts); return err; }
  if (insert)
    err = talshTensorInsert(&tb, &ts, offsets, 0, DEV_HOST, COPY_MT, accumulative ? YEP : NOPE, NULL);
  else
    err = talshTensorSlice(&ts, &tb, offsets, 0, DEV_HOST, COPY_MT, accumulative ? YEP : NOPE, NULL);
  talshTensorDestruct(&tb);
  talshTensorDestruct(&ts);
  return err;
}
int ref_slice(int dk, int rank, const int *sdims, void *s, const int *bdims, void *b, const int *offsets,
              int accumulative) {
  return slice_or_insert(0, dk, rank, sdims, s, bdims, b, offsets, accumulative);
}
int ref_insert(int dk, int rank, const int *sdims, void *s, const int *bdims, void *b, const int *offsets,
               int accumulative) {
  return slice_or_insert(1, dk, rank, sdims, s, bdims, b, offsets, accumulative);
}
// Digital -> symbolic contraction pattern; returns ierr, the string in buf.
int ref_pattern_sym(int lrank, int rrank, int conj_bits, const int *dig, char *buf, int *cpl) {
  int ierr = 0;
  get_contr_pattern_sym(&lrank, &rrank, &conj_bits, dig, buf, cpl, &ierr);
  if (ierr == 0) buf[*cpl] = 0;
  return ierr;
}

// Squared 2-norm via the reference's host reduction (R4/R8 only).
double ref_norm2(int dk, int rank, const int *dims, void *d) {
  tensBlck_t *t = NULL;
  if (tensBlck_create(&t) != 0) return -100.0;
  if (tensBlck_construct(t, NOPE, rank, dims) != 0) { tensBlck_destroy(t); return -101.0; }
  if (tensBlck_attach_body(t, dk, talshFlatDevId(DEV_HOST, 0), d) != 0) { tensBlck_destroy(t); return -102.0; }
  double n2 = tensBlck_norm2_host(t);
  tensBlck_destroy(t);
  return n2;
}

}  // extern "C"
