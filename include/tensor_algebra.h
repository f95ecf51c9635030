/* TAL-SH lower-level constants and interoperable types (MI355X build).
 *
 * Written for this library; the values and struct layouts are the ones the
 * reference defines in tensor_algebra.h:58-317 because callers compiled
 * against TAL-SH pass these structs and codes across the C ABI:
 *   data kinds            tensor_algebra.h:64-73
 *   device kinds / ids    tensor_algebra.h:75-85
 *   aliases, TRY_LATER..  tensor_algebra.h:62, 92-117
 *   coherence codes       tensor_algebra.h:119-203 (2 bits per operand,
 *                         D most significant: D=0 discard, M=1 move,
 *                         T=2 temporary, K=3 keep)
 *   talsh_tens_shape_t    tensor_algebra.h:221-227 (32 B on x86-64)
 *   talsh_dev_rsc_t       tensor_algebra.h:247-252 (24 B)
 */
#ifndef TENSOR_ALGEBRA_H_
#define TENSOR_ALGEBRA_H_

#include <stddef.h>
#include <stdint.h>

#define MAX_TENSOR_RANK 56
#define MAX_TENSOR_OPERANDS 4
#define MAX_CONTRACTION_PATTERN_LEN 1024
#define MAX_MLNDS_PER_TENS 4 /* multi-indices per tensor block: dims, divs, grps, prmn */

/* data kinds */
#define NO_TYPE 0
#define R4 4
#define R8 8
#define C4 14
#define C8 18

/* device kinds and flat device numbering:
   Null -1, Host 0, kind-1 GPUs 1..8, MICs 9..16, kind-3 GPUs 17..24 */
#define MAX_GPUS_PER_NODE 8
#define MAX_MICS_PER_NODE 8
#define MAX_AMDS_PER_NODE 8
#define DEV_NULL -1
#define DEV_DEFAULT DEV_NULL
#define DEV_HOST 0
#define DEV_NVIDIA_GPU 1
#define DEV_INTEL_MIC 2
#define DEV_AMD_GPU 3
#define DEV_MAX (1 + MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE + MAX_AMDS_PER_NODE)

#define MEM_ALLOC_REGULAR 0
#define MEM_ALLOC_TMP_BUF 1
#define MEM_ALLOC_ALL_BUF 2

#define NOPE 0
#define YEP 1
#define BLAS_ON 0
#define BLAS_OFF 1
#define EFF_TRN_OFF 0
#define EFF_TRN_ON 1
#define EFF_TRN_ON_CUTT 2
#define DEVICE_UNABLE -546372819
#define TRY_LATER -918273645
#define NOT_CLEAN -192837465
#define NUM_INSTABILITY -374628195

#define EVERYTHING 0
#define SOURCE 1
#define DESTINATION 2
#define TEMPORARY 3

#define DEV_OFF 0
#define DEV_ON 1
#define DEV_ON_BLAS 2
#define GPU_OFF 0
#define GPU_MINE 1
#define GPU_MINE_CUBLAS 2
#define NO_COPY_BACK 0
#define COPY_BACK 1

/* coherence control: one letter per operand, 2 bits each */
#define COPY_D 0
#define COPY_M 1
#define COPY_T 2
#define COPY_K 3
#define TALSH_COH2(a, b) (((a) << 2) | (b))
#define TALSH_COH3(a, b, c) (((a) << 4) | ((b) << 2) | (c))
#define COPY_DD TALSH_COH2(COPY_D, COPY_D)
#define COPY_DM TALSH_COH2(COPY_D, COPY_M)
#define COPY_DT TALSH_COH2(COPY_D, COPY_T)
#define COPY_DK TALSH_COH2(COPY_D, COPY_K)
#define COPY_MD TALSH_COH2(COPY_M, COPY_D)
#define COPY_MM TALSH_COH2(COPY_M, COPY_M)
#define COPY_MT TALSH_COH2(COPY_M, COPY_T)
#define COPY_MK TALSH_COH2(COPY_M, COPY_K)
#define COPY_TD TALSH_COH2(COPY_T, COPY_D)
#define COPY_TM TALSH_COH2(COPY_T, COPY_M)
#define COPY_TT TALSH_COH2(COPY_T, COPY_T)
#define COPY_TK TALSH_COH2(COPY_T, COPY_K)
#define COPY_KD TALSH_COH2(COPY_K, COPY_D)
#define COPY_KM TALSH_COH2(COPY_K, COPY_M)
#define COPY_KT TALSH_COH2(COPY_K, COPY_T)
#define COPY_KK TALSH_COH2(COPY_K, COPY_K)
#define COPY_DDD TALSH_COH3(COPY_D, COPY_D, COPY_D)
#define COPY_DDM TALSH_COH3(COPY_D, COPY_D, COPY_M)
#define COPY_DDT TALSH_COH3(COPY_D, COPY_D, COPY_T)
#define COPY_DDK TALSH_COH3(COPY_D, COPY_D, COPY_K)
#define COPY_DMD TALSH_COH3(COPY_D, COPY_M, COPY_D)
#define COPY_DMM TALSH_COH3(COPY_D, COPY_M, COPY_M)
#define COPY_DMT TALSH_COH3(COPY_D, COPY_M, COPY_T)
#define COPY_DMK TALSH_COH3(COPY_D, COPY_M, COPY_K)
#define COPY_DTD TALSH_COH3(COPY_D, COPY_T, COPY_D)
#define COPY_DTM TALSH_COH3(COPY_D, COPY_T, COPY_M)
#define COPY_DTT TALSH_COH3(COPY_D, COPY_T, COPY_T)
#define COPY_DTK TALSH_COH3(COPY_D, COPY_T, COPY_K)
#define COPY_DKD TALSH_COH3(COPY_D, COPY_K, COPY_D)
#define COPY_DKM TALSH_COH3(COPY_D, COPY_K, COPY_M)
#define COPY_DKT TALSH_COH3(COPY_D, COPY_K, COPY_T)
#define COPY_DKK TALSH_COH3(COPY_D, COPY_K, COPY_K)
#define COPY_MDD TALSH_COH3(COPY_M, COPY_D, COPY_D)
#define COPY_MDM TALSH_COH3(COPY_M, COPY_D, COPY_M)
#define COPY_MDT TALSH_COH3(COPY_M, COPY_D, COPY_T)
#define COPY_MDK TALSH_COH3(COPY_M, COPY_D, COPY_K)
#define COPY_MMD TALSH_COH3(COPY_M, COPY_M, COPY_D)
#define COPY_MMM TALSH_COH3(COPY_M, COPY_M, COPY_M)
#define COPY_MMT TALSH_COH3(COPY_M, COPY_M, COPY_T)
#define COPY_MMK TALSH_COH3(COPY_M, COPY_M, COPY_K)
#define COPY_MTD TALSH_COH3(COPY_M, COPY_T, COPY_D)
#define COPY_MTM TALSH_COH3(COPY_M, COPY_T, COPY_M)
#define COPY_MTT TALSH_COH3(COPY_M, COPY_T, COPY_T)
#define COPY_MTK TALSH_COH3(COPY_M, COPY_T, COPY_K)
#define COPY_MKD TALSH_COH3(COPY_M, COPY_K, COPY_D)
#define COPY_MKM TALSH_COH3(COPY_M, COPY_K, COPY_M)
#define COPY_MKT TALSH_COH3(COPY_M, COPY_K, COPY_T)
#define COPY_MKK TALSH_COH3(COPY_M, COPY_K, COPY_K)
#define COPY_TDD TALSH_COH3(COPY_T, COPY_D, COPY_D)
#define COPY_TDM TALSH_COH3(COPY_T, COPY_D, COPY_M)
#define COPY_TDT TALSH_COH3(COPY_T, COPY_D, COPY_T)
#define COPY_TDK TALSH_COH3(COPY_T, COPY_D, COPY_K)
#define COPY_TMD TALSH_COH3(COPY_T, COPY_M, COPY_D)
#define COPY_TMM TALSH_COH3(COPY_T, COPY_M, COPY_M)
#define COPY_TMT TALSH_COH3(COPY_T, COPY_M, COPY_T)
#define COPY_TMK TALSH_COH3(COPY_T, COPY_M, COPY_K)
#define COPY_TTD TALSH_COH3(COPY_T, COPY_T, COPY_D)
#define COPY_TTM TALSH_COH3(COPY_T, COPY_T, COPY_M)
#define COPY_TTT TALSH_COH3(COPY_T, COPY_T, COPY_T)
#define COPY_TTK TALSH_COH3(COPY_T, COPY_T, COPY_K)
#define COPY_TKD TALSH_COH3(COPY_T, COPY_K, COPY_D)
#define COPY_TKM TALSH_COH3(COPY_T, COPY_K, COPY_M)
#define COPY_TKT TALSH_COH3(COPY_T, COPY_K, COPY_T)
#define COPY_TKK TALSH_COH3(COPY_T, COPY_K, COPY_K)
#define COPY_KDD TALSH_COH3(COPY_K, COPY_D, COPY_D)
#define COPY_KDM TALSH_COH3(COPY_K, COPY_D, COPY_M)
#define COPY_KDT TALSH_COH3(COPY_K, COPY_D, COPY_T)
#define COPY_KDK TALSH_COH3(COPY_K, COPY_D, COPY_K)
#define COPY_KMD TALSH_COH3(COPY_K, COPY_M, COPY_D)
#define COPY_KMM TALSH_COH3(COPY_K, COPY_M, COPY_M)
#define COPY_KMT TALSH_COH3(COPY_K, COPY_M, COPY_T)
#define COPY_KMK TALSH_COH3(COPY_K, COPY_M, COPY_K)
#define COPY_KTD TALSH_COH3(COPY_K, COPY_T, COPY_D)
#define COPY_KTM TALSH_COH3(COPY_K, COPY_T, COPY_M)
#define COPY_KTT TALSH_COH3(COPY_K, COPY_T, COPY_T)
#define COPY_KTK TALSH_COH3(COPY_K, COPY_T, COPY_K)
#define COPY_KKD TALSH_COH3(COPY_K, COPY_K, COPY_D)
#define COPY_KKM TALSH_COH3(COPY_K, COPY_K, COPY_M)
#define COPY_KKT TALSH_COH3(COPY_K, COPY_K, COPY_T)
#define COPY_KKK TALSH_COH3(COPY_K, COPY_K, COPY_K)

/* Tensor signature (base offsets of a slice). */
typedef struct {
  int num_dim;
  size_t *offsets;
} talsh_tens_signature_t;

/* Tensor shape: rank (-1 = empty) and dimension extents. */
typedef struct {
  int num_dim;
  int *dims;
  int *divs;
  int *grps;
} talsh_tens_shape_t;

/* Tensor data descriptor handed to user initialisers. */
typedef struct {
  void *base;
  size_t volume;
  int data_kind;
} talsh_tens_data_t;

/* One body image of a tensor on one device. */
typedef struct {
  int dev_id;       /* flat device id, -1 = none */
  void *gmem_p;     /* body pointer in that device's memory */
  int buf_entry;    /* buffer-pool handle, -1 = not from a pool */
  int mem_attached; /* 1 = external memory attached by the caller */
} talsh_dev_rsc_t;

typedef int (*talsh_tens_init_i)(const talsh_tens_data_t *tens_data, const talsh_tens_shape_t *tens_shape,
                                 const talsh_tens_signature_t *tens_signature);

#ifdef __cplusplus
extern "C" {
#endif
/* Digital -> symbolic contraction pattern (tensor_algebra.h:265-266, the Fortran routine
 * tensor_algebra_cpu.F90:5533-5611): "D(a,b,..)+=L(..)*R(..)" with uncontracted indices named a, b, ... in
 * D order and contracted ones after them; conj_bits 1/2/4 put '+' after D/L/R.  cptrn_sym receives cpl
 * characters followed by a NUL; ierr = 1 for negative ranks. */
void get_contr_pattern_sym(const int *rank_left, const int *rank_right, const int *conj_bits, const int *cptrn_dig,
                           char *cptrn_sym, int *cpl, int *ierr);
/* Initialiser that leaves a new body as allocated (tensor_algebra.h:256): with it talshTensorConstruct skips
 * the initialisation, on the Host and on GPUs. */
int talsh_tens_no_init(const talsh_tens_data_t *tens_data, const talsh_tens_shape_t *tens_shape,
                       const talsh_tens_signature_t *tens_signature);
#ifdef __cplusplus
}
#endif

#endif /* TENSOR_ALGEBRA_H_ */
