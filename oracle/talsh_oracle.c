/* CPU oracle for the TAL-SH tensor-contraction hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of TAL-SH's CPU algorithms
 * used as the checker by tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py.  The product library never links or calls this file.
 * Each function cites the reference routine it restates; how each one is
 * pinned is listed in talsh_oracle.h.
 */
#include "talsh_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { ORC_R4 = 4, ORC_R8 = 8, ORC_C4 = 14, ORC_C8 = 18, ORC_MAX_RANK = 56 };

static int dk_size(int dk) {
  switch (dk) {
    case ORC_R4: return 4;
    case ORC_R8: return 8;
    case ORC_C4: return 8;
    case ORC_C8: return 16;
    default: return 0;
  }
}

static size_t volume(int rank, const int *dims) {
  size_t v = 1;
  for (int i = 0; i < rank; ++i) v *= (size_t)dims[i];
  return v;
}

/* ------------------------------------------------------------------------- */
/* Pattern parsing.                                                          */
/* Restates talsh_get_contr_ptrn_str2dig (talshf.F90:619-667) and            */
/* get_contr_pattern_dig (tensor_algebra_cpu.F90:5307-5531):                 */
/*  - up to three tensor arguments D, L, R written as NAME[+](i1,i2,...);    */
/*    a '+' right after an alphanumeric name conjugates that argument        */
/*    (bit 0 D, bit 1 L, bit 2 R);                                           */
/*  - indices are alphanumeric labels separated by ',' or '|'; an optional   */
/*    trailing '+' marks a contravariant index and is not part of the label; */
/*  - a unary pattern (no '*' followed by a parenthesised argument) gets a   */
/*    rank-0 right argument "*R()";                                          */
/*  - digit for an L/R dimension: +k if the label is dimension k of D,       */
/*    -k if it is contracted with dimension k of the other input.            */
/* ------------------------------------------------------------------------- */

#define ORC_MAX_LABEL 64
typedef struct {
  int nargs;
  int rank[3];
  char label[3][ORC_MAX_RANK][ORC_MAX_LABEL];
  int conj;
} parsed_t;

static int is_alnum(char c) {
  return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9');
}

/* Tokenises the arguments; returns 0 or a nonzero error code. */
static int tokenize(const char *s, size_t len, parsed_t *p) {
  size_t i = 0;
  memset(p, 0, sizeof(*p));
  while (i < len) {
    while (i < len && s[i] != '(') ++i;
    if (i >= len) break;
    if (p->nargs >= 3) return 1; /* more than two input arguments */
    int a = p->nargs;
    if (i >= 2 && s[i - 1] == '+' && (is_alnum(s[i - 2]) || s[i - 2] == '_')) p->conj |= (1 << a);
    ++i;
    size_t start = i;
    int closed = 0, any_sep = 0;
    while (i < len) {
      char c = s[i];
      if (c == ',' || c == '|' || c == ')') {
        size_t l = i - start;
        if (l == 0) {
          if (c != ')' || any_sep) return (c == ')') ? 3 : 2; /* empty label */
        } else {
          /* label: alnum chars, optional final '+' (not the first char) */
          size_t eff = l;
          for (size_t q = 0; q < l; ++q) {
            char ch = s[start + q];
            if (!is_alnum(ch)) {
              if (q + 1 == l && q > 0 && ch == '+') {
                eff = l - 1;
              } else {
                return (c == ')') ? 4 : 2;
              }
            }
          }
          if (eff >= ORC_MAX_LABEL || p->rank[a] >= ORC_MAX_RANK) return 17;
          memcpy(p->label[a][p->rank[a]], s + start, eff);
          p->label[a][p->rank[a]][eff] = 0;
          p->rank[a]++;
        }
        if (c == ')') { closed = 1; ++i; break; }
        any_sep = 1;
        start = i + 1;
      }
      ++i;
    }
    if (!closed) return 5;
    p->nargs++;
  }
  return 0;
}

int orc_parse_pattern(const char *ptrn, int *dig, int *drank, int *lrank, int *rrank, int *conj_bits) {
  char buf[1100];
  size_t len = strlen(ptrn);
  *drank = -1; *lrank = -1; *rrank = -1; *conj_bits = 0;
  if (len == 0) return 16;
  if (len > 1024) return -1;
  memcpy(buf, ptrn, len);
  buf[len] = 0;
  /* scalar multiplication after '*' (no parenthesis follows) is dropped */
  const char *star = strchr(buf, '*');
  size_t star_pos = star ? (size_t)(star - buf) : len;
  if (star && star_pos < len) {
    int has_paren = 0;
    for (size_t q = star_pos; q < len; ++q)
      if (buf[q] == '(') has_paren = 1;
    if (!has_paren) {
      len = star_pos;
      if (len == 0) return -2;
      if (buf[len - 1] != ')') return -3;
      star = NULL;
    }
  }
  if (!star) { memcpy(buf + len, "*R()", 4); len += 4; }
  buf[len] = 0;

  parsed_t p;
  int err = tokenize(buf, len, &p);
  if (err) return err;
  if (p.nargs != 3) return 15;
  int nd = p.rank[0], nl = p.rank[1], nr = p.rank[2];
  /* mark which occurrences were consumed */
  int used[3][ORC_MAX_RANK];
  memset(used, 0, sizeof(used));
  for (int k = 0; k < nl + nr; ++k) dig[k] = 0;
  for (int a = 0; a < 3; ++a) {
    for (int x = 0; x < p.rank[a]; ++x) {
      if (used[a][x]) continue;
      used[a][x] = 1;
      /* collect later occurrences of this label */
      int occ_a[8], occ_x[8], nocc = 0;
      for (int b = a; b < 3; ++b)
        for (int y = (b == a ? x + 1 : 0); y < p.rank[b]; ++y)
          if (!used[b][y] && strcmp(p.label[a][x], p.label[b][y]) == 0) {
            if (nocc < 8) { occ_a[nocc] = b; occ_x[nocc] = y; }
            nocc++;
          }
      if (nocc == 0) return 11; /* index appears only once */
      if (nocc > 2) return 10;
      for (int t = 0; t < nocc; ++t) {
        int b = occ_a[t], y = occ_x[t];
        used[b][y] = 1;
        if (a == 0 && b == 1) {
          if (t != 0) return 8;
          dig[y] = x + 1;
        } else if (a == 0 && b == 2) {
          dig[nl + y] = x + 1; /* uncontracted (or hyper) index of R */
        } else if (a == 1 && b == 2) {
          if (t != 0) return 9;
          dig[x] = -(y + 1);
          dig[nl + y] = -(x + 1);
        } else {
          return 10; /* repeated within one argument or D-only */
        }
      }
    }
  }
  *drank = nd; *lrank = nl; *rrank = nr; *conj_bits = p.conj;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* TN matricization.  Restates get_contr_permutations                        */
/* (tensor_algebra_cpu.F90:5613-5755): dprm is the N2O permutation of D,     */
/* lprm/rprm are O2N permutations of L/R (all 1-based, element 0 = sign).    */
/* Contracted L dims are ordered like their partners in R; if D's leading    */
/* dims are exactly L's uncontracted dims (and there are >= 2 of them), D     */
/* stays in place and L/R uncontracted dims follow D's order.                */
/* ------------------------------------------------------------------------- */
static int contr_pattern_ok(int lrank, int rrank, const int *c) {
  int jl = lrank + rrank, seen[2 * ORC_MAX_RANK + 1], nunc = 0;
  memset(seen, 0, sizeof(seen));
  for (int j0 = 1; j0 <= jl; ++j0) {
    int j1 = c[j0 - 1];
    if (j1 < 0) {
      if (j0 <= lrank) {
        if (-j1 > rrank) return 0;
        if (c[lrank + (-j1) - 1] != -j0) return 0;
      } else {
        if (-j1 > lrank) return 0;
        if (c[(-j1) - 1] != -(j0 - lrank)) return 0;
      }
    } else if (j1 > 0 && j1 <= jl) {
      nunc++;
      if (seen[j1]) return 0;
      seen[j1] = 1;
    } else {
      return 0;
    }
  }
  for (int j = 1; j <= nunc; ++j)
    if (!seen[j]) return 0;
  return 1;
}

int orc_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *c, int conj_bits,
                           int *dprm, int *lprm, int *rprm, int *ncd_o, int *nlu_o, int *nru_o) {
  if (lrank < 0 || rrank < 0 || !contr_pattern_ok(lrank, rrank, c)) return 1;
  int lconj = (conj_bits >> 1) & 1, rconj = (conj_bits >> 2) & 1;
  if (conj_bits & 1) { lconj = !lconj; rconj = !rconj; }
  if (gemm_tl == 0 && lconj) return 2;
  int drank = 0, simple, nlu, nru = 0, ncd = 0, j;
  dprm[0] = 1;
  for (int i = 1; i <= lrank; ++i)
    if (c[i - 1] > 0) dprm[++drank] = c[i - 1];
  nlu = drank;
  simple = (drank >= 2);
  for (int i = lrank + 1; i <= lrank + rrank; ++i)
    if (c[i - 1] > 0) {
      dprm[++drank] = c[i - 1];
      if (c[i - 1] <= nlu) simple = 0;
    }
  if (simple)
    for (int i = 1; i <= drank; ++i) dprm[i] = i;
  rprm[0] = 1;
  if (rrank > 0) {
    j = 0;
    for (int i = 1; i <= rrank; ++i)
      if (c[lrank + i - 1] < 0) rprm[i] = ++j;
    ncd = j;
    nru = rrank - ncd;
    if (simple && nru >= 2) {
      for (int i = 1; i <= rrank; ++i)
        if (c[lrank + i - 1] > 0) rprm[i] = ncd + (c[lrank + i - 1] - nlu);
    } else {
      for (int i = 1; i <= rrank; ++i)
        if (c[lrank + i - 1] > 0) rprm[i] = ++j;
    }
  }
  lprm[0] = 1;
  if (lrank > 0) {
    /* contracted dims of L, ordered by the position of their partner in R */
    int pos[ORC_MAX_RANK], key[ORC_MAX_RANK];
    j = 0;
    for (int i = 1; i <= lrank; ++i)
      if (c[i - 1] < 0) { pos[j] = i; key[j] = -c[i - 1]; ++j; }
    for (int rank_ = 0; rank_ < j; ++rank_) {
      int cnt = 0;
      for (int q = 0; q < j; ++q)
        if (key[q] < key[rank_]) cnt++;
      lprm[pos[rank_]] = cnt + 1;
    }
    if (simple && nlu >= 2) {
      for (int i = 1; i <= lrank; ++i)
        if (c[i - 1] > 0) lprm[i] = ncd + c[i - 1];
    } else {
      for (int i = 1; i <= lrank; ++i)
        if (c[i - 1] > 0) lprm[i] = ++j;
    }
  }
  if (gemm_tl == 0)
    for (int i = 1; i <= lrank; ++i) lprm[i] = (lprm[i] <= ncd) ? lprm[i] + nlu : lprm[i] - ncd;
  if (gemm_tr != 0 || rconj)
    for (int i = 1; i <= rrank; ++i) rprm[i] = (rprm[i] <= ncd) ? rprm[i] + nru : rprm[i] - ncd;
  *ncd_o = ncd; *nlu_o = nlu; *nru_o = nru;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Copy / add with an O2N permutation.                                       */
/* Restates tensor_block_copy_dlf / _scatter_dlf                             */
/* (tensor_algebra_cpu.F90:7699-7940, :8611-8614): the element at input      */
/* multi-index (i1..in) lands at output offset sum_k i_k * base_out[o2n[k]]. */
/* ------------------------------------------------------------------------- */
typedef struct {
  int rank;
  size_t ext[ORC_MAX_RANK];
  size_t ostr[ORC_MAX_RANK]; /* output stride of input dim k */
} perm_map_t;

static int make_perm_map(int rank, const int *ldims, const int *o2n, perm_map_t *m) {
  int ddims[ORC_MAX_RANK], seen[ORC_MAX_RANK];
  size_t base[ORC_MAX_RANK];
  if (rank < 0 || rank > ORC_MAX_RANK) return 1;
  memset(seen, 0, sizeof(seen));
  for (int k = 0; k < rank; ++k) {
    int t = o2n[k];
    if (t < 1 || t > rank || seen[t - 1]) return 2;
    seen[t - 1] = 1;
    ddims[t - 1] = ldims[k];
  }
  size_t s = 1;
  for (int k = 0; k < rank; ++k) { base[k] = s; s *= (size_t)ddims[k]; }
  m->rank = rank;
  for (int k = 0; k < rank; ++k) { m->ext[k] = (size_t)ldims[k]; m->ostr[k] = base[o2n[k] - 1]; }
  return 0;
}

/* Calls fn(in_offset, out_offset) for every element, input order. */
#define FOR_EACH_PERM(m, vol, ...)                                          \
  do {                                                                       \
    size_t idx_[ORC_MAX_RANK];                                               \
    memset(idx_, 0, sizeof(idx_));                                           \
    size_t out_ = 0;                                                         \
    for (size_t in_ = 0; in_ < (vol); ++in_) {                               \
      __VA_ARGS__;                                                           \
      for (int k_ = 0; k_ < (m).rank; ++k_) {                                \
        if (++idx_[k_] < (m).ext[k_]) { out_ += (m).ostr[k_]; break; }       \
        out_ -= (m).ostr[k_] * ((m).ext[k_] - 1);                            \
        idx_[k_] = 0;                                                        \
      }                                                                      \
    }                                                                        \
  } while (0)

int orc_copy(int dk, int rank, const int *ldims, const int *o2n, const void *l, void *d, int conj) {
  perm_map_t m;
  if (dk_size(dk) == 0) return 1;
  if (make_perm_map(rank, ldims, o2n, &m)) return 2;
  size_t vol = volume(rank, ldims);
  switch (dk) {
    case ORC_R4: { const float *a = l; float *b = d; FOR_EACH_PERM(m, vol, b[out_] = a[in_]); } break;
    case ORC_R8: { const double *a = l; double *b = d; FOR_EACH_PERM(m, vol, b[out_] = a[in_]); } break;
    case ORC_C4: {
      const float *a = l; float *b = d; float sg = conj ? -1.0f : 1.0f;
      FOR_EACH_PERM(m, vol, { b[2 * out_] = a[2 * in_]; b[2 * out_ + 1] = sg * a[2 * in_ + 1]; });
    } break;
    case ORC_C8: {
      const double *a = l; double *b = d; double sg = conj ? -1.0 : 1.0;
      FOR_EACH_PERM(m, vol, { b[2 * out_] = a[2 * in_]; b[2 * out_ + 1] = sg * a[2 * in_ + 1]; });
    } break;
  }
  return 0;
}

/* Restates cpu_tensor_block_add (talshf.F90:1862-1903) + tensor_block_add
 * (tensor_algebra_cpu.F90:3312-3673): d = d + l*alpha with a separate
 * multiply and add; R4 uses alpha rounded to float; complex products are
 * (a+bi)(c+di) = (ac-bd) + (ad+bc)i. */
int orc_add(int dk, int rank, const int *ldims, const int *o2n, const void *l, void *d, double ar, double ai,
            int conj) {
  perm_map_t m;
  if (dk_size(dk) == 0) return 1;
  if (make_perm_map(rank, ldims, o2n, &m)) return 2;
  size_t vol = volume(rank, ldims);
  switch (dk) {
    case ORC_R4: {
      const float *a = l; float *b = d; volatile float al = (float)ar;
      FOR_EACH_PERM(m, vol, { volatile float p = a[in_] * al; b[out_] = b[out_] + p; });
    } break;
    case ORC_R8: {
      const double *a = l; double *b = d;
      FOR_EACH_PERM(m, vol, { volatile double p = a[in_] * ar; b[out_] = b[out_] + p; });
    } break;
    case ORC_C4: {
      const float *a = l; float *b = d; float sg = conj ? -1.0f : 1.0f;
      float cr = (float)ar, ci = (float)ai;
      FOR_EACH_PERM(m, vol, {
        float xr = a[2 * in_], xi = sg * a[2 * in_ + 1];
        volatile float t1 = xr * cr, t2 = xi * ci, t3 = xr * ci, t4 = xi * cr;
        volatile float pr = t1 - t2, pi = t3 + t4;
        b[2 * out_] = b[2 * out_] + pr;
        b[2 * out_ + 1] = b[2 * out_ + 1] + pi;
      });
    } break;
    case ORC_C8: {
      const double *a = l; double *b = d; double sg = conj ? -1.0 : 1.0;
      FOR_EACH_PERM(m, vol, {
        double xr = a[2 * in_], xi = sg * a[2 * in_ + 1];
        volatile double t1 = xr * ar, t2 = xi * ai, t3 = xr * ai, t4 = xi * ar;
        volatile double pr = t1 - t2, pi = t3 + t4;
        b[2 * out_] = b[2 * out_] + pr;
        b[2 * out_ + 1] = b[2 * out_ + 1] + pi;
      });
    } break;
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Contraction.  Restates tensor_block_contract                              */
/* (tensor_algebra_cpu.F90:3675-4358) as TTGT: L and R are permuted into     */
/* K x M and K x N column-major matrices with the TN permutations above, D   */
/* into M x N; D' = beta*D' + alpha * L'^T R' (beta = 0 if not accumulative) */
/* and D' is permuted back.  A conjugation bit on D flips the L and R bits.  */
/* Dot products accumulate in long double (this is the FP64 error reference).*/
/* ------------------------------------------------------------------------- */
static void conj_flags(int conj_bits, int *lc, int *rc) {
  *lc = (conj_bits >> 1) & 1;
  *rc = (conj_bits >> 2) & 1;
  if (conj_bits & 1) { *lc = !*lc; *rc = !*rc; }
}

int orc_contract(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, void *d, int lrank,
                 const int *ldims, const void *l, int rrank, const int *rdims, const void *r, double ar, double ai,
                 int accumulative) {
  int dprm[ORC_MAX_RANK + 1], lprm[ORC_MAX_RANK + 1], rprm[ORC_MAX_RANK + 1], ncd, nlu, nru;
  int es = dk_size(dk);
  if (es == 0) return 1;
  int cplx = (dk == ORC_C4 || dk == ORC_C8);
  if (orc_contr_permutations(1, 0, lrank, rrank, cptrn, 0, dprm, lprm, rprm, &ncd, &nlu, &nru)) return 2;
  if (nlu + nru != drank) return 3;
  int lc, rc;
  conj_flags(cplx ? conj_bits : 0, &lc, &rc);
  size_t vl = volume(lrank, ldims), vr = volume(rrank, rdims), vd = volume(drank, ddims);
  size_t K = 1, M = 1, N = 1;
  for (int i = 0; i < lrank; ++i) if (cptrn[i] < 0) K *= (size_t)ldims[i]; else M *= (size_t)ldims[i];
  for (int i = 0; i < rrank; ++i) if (cptrn[lrank + i] > 0) N *= (size_t)rdims[i];
  if (M * N != vd || K * M != vl || K * N != vr) return 4;
  /* widen operands to long double (re,im) in matricized order */
  int ncomp = cplx ? 2 : 1;
  long double *Lm = malloc(sizeof(long double) * vl * ncomp);
  long double *Rm = malloc(sizeof(long double) * vr * ncomp);
  double *Lt = malloc((size_t)es * vl), *Rt = malloc((size_t)es * vr), *Dt = malloc((size_t)es * vd);
  if (!Lm || !Rm || !Lt || !Rt || !Dt) { free(Lm); free(Rm); free(Lt); free(Rt); free(Dt); return 5; }
  orc_copy(dk, lrank, ldims, lprm + 1, l, Lt, 0);
  orc_copy(dk, rrank, rdims, rprm + 1, r, Rt, 0);
  /* D' = permuted D: D'(new k) = D(dprm[k]) (N2O) -> O2N for the copy */
  int do2n[ORC_MAX_RANK], dn_dims[ORC_MAX_RANK];
  for (int k = 1; k <= drank; ++k) do2n[dprm[k] - 1] = k;
  orc_copy(dk, drank, ddims, do2n, d, Dt, 0);
  for (int k = 1; k <= drank; ++k) dn_dims[k - 1] = ddims[dprm[k] - 1];
  for (size_t q = 0; q < vl * ncomp; ++q)
    Lm[q] = (dk == ORC_R4 || dk == ORC_C4) ? (long double)((float *)Lt)[q] : (long double)Lt[q];
  for (size_t q = 0; q < vr * ncomp; ++q)
    Rm[q] = (dk == ORC_R4 || dk == ORC_C4) ? (long double)((float *)Rt)[q] : (long double)Rt[q];
  long double sl = lc ? -1.0L : 1.0L, sr = rc ? -1.0L : 1.0L;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t n = 0; n < N; ++n) {
    for (size_t m = 0; m < M; ++m) {
      long double sre = 0.0L, sim = 0.0L;
      if (!cplx) {
        const long double *a = Lm + m * K, *b = Rm + n * K;
        for (size_t k = 0; k < K; ++k) sre += a[k] * b[k];
      } else {
        const long double *a = Lm + 2 * m * K, *b = Rm + 2 * n * K;
        for (size_t k = 0; k < K; ++k) {
          long double xr = a[2 * k], xi = sl * a[2 * k + 1], yr = b[2 * k], yi = sr * b[2 * k + 1];
          sre += xr * yr - xi * yi;
          sim += xr * yi + xi * yr;
        }
      }
      size_t o = m + n * M;
      long double vre = (long double)ar * sre - (long double)ai * sim;
      long double vim = (long double)ar * sim + (long double)ai * sre;
      switch (dk) {
        case ORC_R4: { float *D = (float *)Dt; D[o] = (float)((accumulative ? (long double)D[o] : 0.0L) + vre); } break;
        case ORC_R8: { double *D = Dt; D[o] = (double)((accumulative ? (long double)D[o] : 0.0L) + vre); } break;
        case ORC_C4: {
          float *D = (float *)Dt;
          D[2 * o] = (float)((accumulative ? (long double)D[2 * o] : 0.0L) + vre);
          D[2 * o + 1] = (float)((accumulative ? (long double)D[2 * o + 1] : 0.0L) + vim);
        } break;
        case ORC_C8: {
          double *D = Dt;
          D[2 * o] = (double)((accumulative ? (long double)D[2 * o] : 0.0L) + vre);
          D[2 * o + 1] = (double)((accumulative ? (long double)D[2 * o + 1] : 0.0L) + vim);
        } break;
      }
    }
  }
  /* permute back: D'(dims dn_dims) -> D with O2N = dprm */
  orc_copy(dk, drank, dn_dims, dprm + 1, Dt, d, 0);
  free(Lm); free(Rm); free(Lt); free(Rt); free(Dt);
  return 0;
}

/* One element of D = sum over the contracted multi-index, as a direct loop
 * nest over the digital pattern (independent of the matricization above).
 * An index that D shares with both L and R (hyper index) simply fixes that
 * position in both inputs: this is the definition the reference's CPU path
 * evaluates with a batched GEMM (tensor_algebra_cpu.F90:5757-5956,
 * :3983-3985). */
static void element_sum(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, size_t doff, int lrank,
                        const int *ldims, const void *l, int rrank, const int *rdims, const void *r,
                        long double *out_re, long double *out_im) {
  size_t dix[ORC_MAX_RANK];
  int cplx = (dk == ORC_C4 || dk == ORC_C8);
  int lc, rc;
  conj_flags(cplx ? conj_bits : 0, &lc, &rc);
  for (int k = 0; k < drank; ++k) { dix[k] = doff % (size_t)ddims[k]; doff /= (size_t)ddims[k]; }
  size_t lstr[ORC_MAX_RANK], rstr[ORC_MAX_RANK], s = 1;
  for (int k = 0; k < lrank; ++k) { lstr[k] = s; s *= (size_t)ldims[k]; }
  s = 1;
  for (int k = 0; k < rrank; ++k) { rstr[k] = s; s *= (size_t)rdims[k]; }
  size_t lbase = 0, rbase = 0;
  int cl[ORC_MAX_RANK], nc = 0; /* contracted dims of L */
  for (int k = 0; k < lrank; ++k) {
    if (cptrn[k] > 0) lbase += dix[cptrn[k] - 1] * lstr[k];
    else cl[nc++] = k;
  }
  for (int k = 0; k < rrank; ++k)
    if (cptrn[lrank + k] > 0) rbase += dix[cptrn[lrank + k] - 1] * rstr[k];
  size_t cnt[ORC_MAX_RANK];
  memset(cnt, 0, sizeof(cnt));
  size_t KK = 1;
  for (int q = 0; q < nc; ++q) KK *= (size_t)ldims[cl[q]];
  long double sre = 0.0L, sim = 0.0L;
  for (size_t t = 0; t < KK; ++t) {
    size_t lo = lbase, ro = rbase;
    for (int q = 0; q < nc; ++q) {
      int kl = cl[q], kr = -cptrn[kl] - 1;
      lo += cnt[q] * lstr[kl];
      ro += cnt[q] * rstr[kr];
    }
    switch (dk) {
      case ORC_R4: sre += (long double)((const float *)l)[lo] * (long double)((const float *)r)[ro]; break;
      case ORC_R8: sre += (long double)((const double *)l)[lo] * (long double)((const double *)r)[ro]; break;
      case ORC_C4:
      case ORC_C8: {
        long double xr, xi, yr, yi;
        if (dk == ORC_C4) {
          const float *a = l, *b = r;
          xr = a[2 * lo]; xi = a[2 * lo + 1]; yr = b[2 * ro]; yi = b[2 * ro + 1];
        } else {
          const double *a = l, *b = r;
          xr = a[2 * lo]; xi = a[2 * lo + 1]; yr = b[2 * ro]; yi = b[2 * ro + 1];
        }
        if (lc) xi = -xi;
        if (rc) yi = -yi;
        sre += xr * yr - xi * yi;
        sim += xr * yi + xi * yr;
      } break;
    }
    for (int q = 0; q < nc; ++q) {
      if (++cnt[q] < (size_t)ldims[cl[q]]) break;
      cnt[q] = 0;
    }
  }
  *out_re = sre;
  *out_im = sim;
}

int orc_contract_element(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, size_t doff,
                         int lrank, const int *ldims, const void *l, int rrank, const int *rdims, const void *r,
                         double *out_re, double *out_im) {
  long double sre, sim;
  element_sum(dk, cptrn, conj_bits, drank, ddims, doff, lrank, ldims, l, rrank, rdims, r, &sre, &sim);
  *out_re = (double)sre;
  *out_im = (double)sim;
  return 0;
}

/* Whole contraction by the direct loop nest, hyper indices included:
 * D = [D] + alpha * sum, each element accumulated in long double and rounded
 * once, as orc_contract does. */
int orc_contract_direct(int dk, const int *cptrn, int conj_bits, int drank, const int *ddims, void *d, int lrank,
                        const int *ldims, const void *l, int rrank, const int *rdims, const void *r, double ar,
                        double ai, int accumulative) {
  if (dk_size(dk) == 0) return 1;
  size_t vd = volume(drank, ddims);
#pragma omp parallel for schedule(static)
  for (size_t o = 0; o < vd; ++o) {
    long double sre, sim;
    element_sum(dk, cptrn, conj_bits, drank, ddims, o, lrank, ldims, l, rrank, rdims, r, &sre, &sim);
    long double vre = (long double)ar * sre - (long double)ai * sim;
    long double vim = (long double)ar * sim + (long double)ai * sre;
    switch (dk) {
      case ORC_R4: { float *D = d; D[o] = (float)((accumulative ? (long double)D[o] : 0.0L) + vre); } break;
      case ORC_R8: { double *D = d; D[o] = (double)((accumulative ? (long double)D[o] : 0.0L) + vre); } break;
      case ORC_C4: {
        float *D = d;
        D[2 * o] = (float)((accumulative ? (long double)D[2 * o] : 0.0L) + vre);
        D[2 * o + 1] = (float)((accumulative ? (long double)D[2 * o + 1] : 0.0L) + vim);
      } break;
      case ORC_C8: {
        double *D = d;
        D[2 * o] = (double)((accumulative ? (long double)D[2 * o] : 0.0L) + vre);
        D[2 * o + 1] = (double)((accumulative ? (long double)D[2 * o + 1] : 0.0L) + vim);
      } break;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Squared 2-norm.  Restates tensBlck_norm2_host                             */
/* (tensor_algebra_gpu_nvidia.cu:3269-3299): R4 squares in float and         */
/* accumulates in a wider type; R8 squares and accumulates in double; the    */
/* complex kinds add re^2 + im^2 (tensor_block_norm2,                        */
/* tensor_algebra_cpu.F90:2032-2129).  Accumulation here is long double.     */
/* ------------------------------------------------------------------------- */
double orc_norm2(int dk, size_t vol, const void *p) {
  long double s = 0.0L;
  switch (dk) {
    case ORC_R4: { const float *a = p; for (size_t i = 0; i < vol; ++i) { float q = a[i] * a[i]; s += q; } } break;
    case ORC_R8: { const double *a = p; for (size_t i = 0; i < vol; ++i) s += (long double)a[i] * a[i]; } break;
    case ORC_C4: {
      const float *a = p;
      for (size_t i = 0; i < 2 * vol; ++i) { float q = a[i] * a[i]; s += q; }
    } break;
    case ORC_C8: { const double *a = p; for (size_t i = 0; i < 2 * vol; ++i) s += (long double)a[i] * a[i]; } break;
    default: return -1.0;
  }
  return (double)s;
}

/* Restates talshTensorInit / cpu_tensor_block_init (talshc.cpp:1247-1283). */
void orc_init(int dk, size_t vol, void *d, double vr, double vi) {
  for (size_t i = 0; i < vol; ++i) {
    switch (dk) {
      case ORC_R4: ((float *)d)[i] = (float)vr; break;
      case ORC_R8: ((double *)d)[i] = vr; break;
      case ORC_C4: ((float *)d)[2 * i] = (float)vr; ((float *)d)[2 * i + 1] = (float)vi; break;
      case ORC_C8: ((double *)d)[2 * i] = vr; ((double *)d)[2 * i + 1] = vi; break;
    }
  }
}

/* Restates tensor_block_scale (tensor_algebra_cpu.F90): d = d*alpha. */
void orc_scale(int dk, size_t vol, void *d, double vr, double vi) {
  for (size_t i = 0; i < vol; ++i) {
    switch (dk) {
      case ORC_R4: { float *a = d; a[i] = a[i] * (float)vr; } break;
      case ORC_R8: { double *a = d; a[i] = a[i] * vr; } break;
      case ORC_C4: {
        float *a = d, xr = a[2 * i], xi = a[2 * i + 1], cr = (float)vr, ci = (float)vi;
        volatile float t1 = xr * cr, t2 = xi * ci, t3 = xr * ci, t4 = xi * cr;
        a[2 * i] = t1 - t2; a[2 * i + 1] = t3 + t4;
      } break;
      case ORC_C8: {
        double *a = d, xr = a[2 * i], xi = a[2 * i + 1];
        volatile double t1 = xr * vr, t2 = xi * vi, t3 = xr * vi, t4 = xi * vr;
        a[2 * i] = t1 - t2; a[2 * i + 1] = t3 + t4;
      } break;
    }
  }
}

/* Counter-based splitmix64: element i of stream `seed` is mix(seed + (i+1)*gamma),
 * mapped to uniform[-1,1) as 2*(z>>11)*2^-53 - 1. */
void orc_fill_uniform(uint64_t seed, size_t n, double *out) {
  const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
  for (size_t i = 0; i < n; ++i) {
    uint64_t z = seed + (uint64_t)(i + 1) * gamma;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    out[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Slice / insert: walk the block's elements in dlf order with a multi-index
 * odometer and map each to its position in the big tensor,
 *   pos = sum_k (offsets[k] + i_k) * prod_{m<k} bdims[m]
 * (tensor_block_slice_dlf_r8 / tensor_block_insert_dlf_r8,
 * tensor_algebra_cpu.F90:6900-6974 and the insert counterparts). */
static int block_positions(int rank, const int *sdims, const int *bdims, const int *offsets, size_t *vol) {
  *vol = 1;
  for (int k = 0; k < rank; ++k) {
    if (sdims[k] < 1 || bdims[k] < 1 || offsets[k] < 0 || offsets[k] + sdims[k] > bdims[k]) return 1;
    *vol *= (size_t)sdims[k];
  }
  return 0;
}

static int slice_insert(int insert, int dk, int rank, const int *sdims, void *s, const int *bdims, void *b,
                        const int *offsets, int accumulative) {
  size_t vol;
  if (rank < 0 || rank > 56 || block_positions(rank, sdims, bdims, offsets, &vol)) return 1;
  const int esz = (dk == 4) ? 4 : (dk == 8 || dk == 14) ? 8 : (dk == 18) ? 16 : 0;
  const int ncomp = (dk == 14 || dk == 18) ? 2 : 1;
  if (!esz) return 2;
  int idx[56] = {0};
  size_t bstride[56];
  size_t st = 1;
  for (int k = 0; k < rank; ++k) {
    bstride[k] = st;
    st *= (size_t)bdims[k];
  }
  for (size_t l = 0; l < vol; ++l) {
    size_t pos = 0;
    for (int k = 0; k < rank; ++k) pos += (size_t)(offsets[k] + idx[k]) * bstride[k];
    for (int c = 0; c < ncomp; ++c) {
      if (dk == 4 || dk == 14) {
        float *dst = insert ? (float *)b + pos * ncomp + c : (float *)s + l * ncomp + c;
        const float src = insert ? ((const float *)s)[l * ncomp + c] : ((const float *)b)[pos * ncomp + c];
        *dst = accumulative ? *dst + src : src;
      } else {
        double *dst = insert ? (double *)b + pos * ncomp + c : (double *)s + l * ncomp + c;
        const double src = insert ? ((const double *)s)[l * ncomp + c] : ((const double *)b)[pos * ncomp + c];
        *dst = accumulative ? *dst + src : src;
      }
    }
    for (int k = 0; k < rank; ++k) { /* odometer, first index fastest */
      if (++idx[k] < sdims[k]) break;
      idx[k] = 0;
    }
  }
  return 0;
}

int orc_slice(int dk, int rank, const int *sdims, void *s, const int *bdims, const void *b, const int *offsets,
              int accumulative) {
  return slice_insert(0, dk, rank, sdims, s, bdims, (void *)b, offsets, accumulative);
}

int orc_insert(int dk, int rank, const int *sdims, const void *s, const int *bdims, void *b, const int *offsets,
               int accumulative) {
  return slice_insert(1, dk, rank, sdims, (void *)s, bdims, b, offsets, accumulative);
}
