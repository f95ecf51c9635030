/* CPU oracle for the TAL-SH tensor-contraction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of TAL-SH's CPU
 * algorithms (CP-TAL, tensor_algebra_cpu.F90 + talshf.F90), used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
 * The product library (tal_sh_amd/libtalsh_amd.so) never links or calls it.
 *
 * Conventions (as in the reference): dlf storage (first index fastest),
 * 1-based digital patterns, O2N permutations, data kinds R4=4 R8=8 C4=14
 * C8=18, complex values stored as interleaved (re,im) pairs.
 *
 * Parity pinning of each function (tests/test_oracle.py):
 *   orc_parse_pattern      golden vectors from the reference routine
 *                          (tests/golden/patterns.json) + invariants
 *   orc_contr_permutations golden vectors from the reference routine
 *                          (tests/golden/contr_perms.json)
 *   orc_copy               closed-form index map + reference copies
 *   orc_add                reference adds, bit-exact
 *   orc_contract           closed forms on constant tensors, brute-force
 *                          loop nests, and the reference contraction
 *   orc_contract_element / numpy einsum (hyper indices, conjugation, all
 *   orc_contract_direct    kinds), orc_contract on ordinary patterns, and the
 *                          hyper-sum invariant (summing a hyper result over
 *                          its hyper index = contracting that index)
 *   orc_norm2              exact sums of small integers + reference norm2
 *   orc_init / orc_scale   reference init/scale, bit-exact
 *   orc_slice / orc_insert closed-form index map (values that encode their
 *                          own position) + reference slice/insert, bit-exact
 */
#ifndef TALSH_ORACLE_H_
#define TALSH_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Symbolic pattern -> digital pattern (+ ranks, conjugation bits). */
int orc_parse_pattern(const char *ptrn, int *dig, int *drank, int *lrank, int *rrank, int *conj_bits);

/* TN matricization permutations (get_contr_permutations, gemm_tl/gemm_tr). */
int orc_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn, int conj_bits,
                           int *dprm, int *lprm, int *rprm, int *ncd, int *nlu, int *nru);

/* D = perm(L) with an O2N permutation (1-based); conj applies to complex kinds. */
int orc_copy(int dk, int rank, const int *ldims, const int *o2n, const void *l, void *d, int conj);

/* D += alpha * perm(L), reference arithmetic (separate multiply and add). */
int orc_add(int dk, int rank, const int *ldims, const int *o2n, const void *l, void *d, double ar, double ai,
            int conj);

/* D (+)= alpha * L * R via TTGT, accumulating each dot product in long double. */
int orc_contract(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, void *d, int lrank,
                 const int *ldims, const void *l, int rrank, const int *rdims, const void *r, double ar, double ai,
                 int accumulative);

/* One element of the contraction (the K-term dot product, long double). */
int orc_contract_element(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, size_t doff,
                         int lrank, const int *ldims, const void *l, int rrank, const int *rdims, const void *r,
                         double *out_re, double *out_im);
/* Whole contraction by the same direct loop nest (hyper indices allowed);
 * D = [D] + alpha * sum, long-double accumulation, one rounding. */
int orc_contract_direct(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, void *d, int lrank,
                        const int *ldims, const void *l, int rrank, const int *rdims, const void *r, double ar,
                        double ai, int accumulative);

/* Squared Frobenius norm, long double accumulation. */
double orc_norm2(int dk, size_t vol, const void *p);

void orc_init(int dk, size_t vol, void *d, double vr, double vi);
void orc_scale(int dk, size_t vol, void *d, double vr, double vi);

/* Slice (tensor_block_slice_dlf_*, tensor_algebra_cpu.F90:6821-7053):
 *   s[i] = b[offsets + i]  (s[i] = s[i] + b[offsets + i] when accumulative);
 * insert (tensor_block_insert_dlf_*): b[offsets + i] = s[i] (+=).  s has
 * extents sdims, b has extents bdims, both dlf, 0-based offsets.  Returns
 * nonzero for a block that does not fit. */
int orc_slice(int dk, int rank, const int *sdims, void *s, const int *bdims, const void *b, const int *offsets,
              int accumulative);
int orc_insert(int dk, int rank, const int *sdims, const void *s, const int *bdims, void *b, const int *offsets,
               int accumulative);

/* Seeded uniform[-1,1] stream (splitmix64), shared with the Python tests. */
void orc_fill_uniform(uint64_t seed, size_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
