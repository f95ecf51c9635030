// Host-side planning: symbolic pattern parsing and TTGT matricization.
#pragma once

#include <cstdint>
#include <vector>

namespace tal {

// Digital form of a symbolic pattern (TAL-SH convention, 1-based):
// dig[0..lrank) for L dims, dig[lrank..lrank+rrank) for R dims; +k = goes to
// D dim k, -k = contracted with dim k of the other input.  conj bit 0 = D,
// bit 1 = L, bit 2 = R.
struct DigitalPattern {
  int drank = -1, lrank = -1, rrank = -1, conj = 0;
  std::vector<int> dig;
};

// Parses "D(a,b)+=L(a,i)*R(i,b)" style patterns, also unary "D(a,b)=L(b,a)"
// (treated as having a rank-0 right operand).  Returns 0 or a positive
// parse-error code (never throws).
int parse_pattern(const char *pattern, DigitalPattern *out);

// One GEMM operand of the plan.
struct OperandPlan {
  bool transpose = false;   // permute into a temporary before the GEMM
  std::vector<int> o2n;     // 0-based O2N permutation used for that copy
  char op = 'N';            // matrix form fed to the GEMM: 'N' or 'T'
  int64_t ld = 1;
};

// TTGT plan: D'(MxN) (+)= alpha * op(A) * op(B), A/B = L/R or R/L.
struct ContractPlan {
  bool swap = false;        // A = R and B = L
  int64_t M = 1, N = 1, K = 1;
  int64_t batch = 1;        // product of the hyper extents: GEMMs of M x N x K, strided by M*K, K*N, M*N
  OperandPlan a, b;
  bool d_transpose = false;
  // D' is written in place by the GEMM through per-row / per-column offsets
  // instead of through a transposed temporary: D' starts with D's stride-1
  // index, of extent >= 16 (stores run along m), and has no batch index
  bool d_in_place = false;
  std::vector<int> d_o2n;        // D -> D' (0-based O2N)
  std::vector<int> d_back_o2n;   // D' -> D
  std::vector<int> dprime_dims;  // extents of D' in its own order
  double moved_elems = 0;        // elements copied by the transposes
};

// Validates the pattern against the extents and picks the matricization that
// moves the fewest elements.  Returns 0, or 1 for an invalid pattern/shape.
// Hyper indices (present in D, L and R) give plan->batch > 1.
int plan_contraction(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims,
                     int accumulative, ContractPlan *plan);

}  // namespace tal
