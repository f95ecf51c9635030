"""Writes the golden fixtures in tests/golden/ from TAL-SH's own CPU path.

Run in the development container after `make -C oracle ref` (the reference
build needs /root/reference and amdflang):

    python tests/golden/make_golden.py

Outputs (all data; no reference source is stored):
  patterns.json     symbolic pattern -> (error, ranks, conj bits, digits) from
                    talsh_get_contr_ptrn_str2dig (talshf.F90:619-667), over the
                    config patterns, the reference tests' patterns
                    (test.cpp, main.F90), the 48 shapes of
                    tensor_contractions.txt, edge cases and random patterns.
  contr_perms.json  get_contr_permutations (tensor_algebra_cpu.F90:5613-5755)
                    outputs for those digital patterns, all four GEMM forms.
                    Each parsed pattern also carries "sym", the reference's
                    digital -> symbolic form (get_contr_pattern_sym,
                    tensor_algebra_cpu.F90:5533-5611).
  numeric.npz       small seeded copies/adds/contractions/norms/scales and
                    slices/inserts computed by the reference CPU path (inputs
                    are regenerated from the seeds recorded in numeric.json).
"""
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from tal_common import (C4, C8, R4, R8, Reference, random_contraction,  # noqa: E402
                        random_permutation_pattern, tensor)

REF_CONTRACTIONS_TXT = "/root/reference/tensor_contractions.txt"

FIXED_PATTERNS = [
    # configs
    "D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)",
    "D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)",
    "D(a,b,c,d,e,f)=L(f,e,d,c,b,a)",
    "D(a,b,c,d,e,f)+=L(c,a,f,b,e,d)",
    "D(a,b,c,d,e,f)=L(a,b,c,d,e,f)",
    # reference tests (test.cpp:103-147, :193-226, :495, :504-532; main.F90)
    "D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)",
    "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)",
    "D(a,b,c,d,e)=B(e,d,c,b,a)",
    "D(a,b,c,d,e)=L(c,i,e,j,a)*R(j,d,i,b)",
    "B(a,b,c,d,e)+=D(e,d,c,b,a)",
    "D(a,b)+=L(a,i)*R(i,b)",
    "D(a,b)+=L+(a,i)*R(i,b)",
    "D(a,b)+=L(a,i)*R+(i,b)",
    "D+(a,b)+=L(a,i)*R(i,b)",
    "D(a1,b2)+=L(b2,k2+,c3)*R(c3+,a1,k2)",
    "D(a|b)+=L(a|i)*R(i|b)",
    "D(a,b)+=L(a,i)*R(i,b)*2",
    "D()+=L(a,b)*R(b,a)",
    "D(a,b)+=L(a)*R(b)",
    "D(a)+=L(a)*R()",
    "D()+=L()*R()",
    "D()=L()",
    "D(a,b,c)+=L(a,i,c)*R(i,b,c)",
    # errors
    "D(a,b)+=L(a,a)*R(b)",
    "D(a,b)+=L(a)*R(c)",
    "D(a,b)+=L(a,c)*R(c,b)*Z(d)",
    "D(a,b",
    "D(a,b)+=L(a,,i)*R(i,b)",
    "D(a-b)+=L(a)*R(b)",
    "D(a,b)+=L(a,i)*R(i)",
]


# (pattern, D dims, L dims, R dims) with a rank-0 operand, or no contracted index
RANK0_CASES = [
    ("D(a)+=L(a)*R()", [7], [7], []),
    ("D(a,b,c)+=L(c,a,b)*R()", [3, 4, 5], [5, 3, 4], []),
    ("D(a,b)+=L()*R(b,a)", [4, 6], [], [6, 4]),
    ("D()+=L()*R()", [], [], []),
    ("D(a)+=L+(a)*R()", [9], [9], []),
    ("D()+=L()*R+()", [], [], []),
    ("D(a,b)+=L(a)*R(b)", [5, 6], [5], [6]),
    ("D(a,b)+=L(b)*R+(a)", [4, 3], [3], [4]),
    ("D()+=L(a,b)*R(b,a)", [], [4, 5], [5, 4]),
]


def ref_contractions_txt():
    pats = []
    if os.path.exists(REF_CONTRACTIONS_TXT):
        lines = [l.rstrip("\n") for l in open(REF_CONTRACTIONS_TXT) if l.strip()]
        for i in range(0, len(lines) - 1, 2):
            if lines[i].startswith("D("):
                nums = [int(x) for x in lines[i + 1].split()]
                rd = nums[0]
                dd = nums[1:1 + rd]
                rl = nums[1 + rd]
                ld = nums[2 + rd:2 + rd + rl]
                rr = nums[2 + rd + rl]
                rdims = nums[3 + rd + rl:3 + rd + rl + rr]
                pats.append({"pattern": lines[i].strip(), "ddims": dd, "ldims": ld, "rdims": rdims})
    return pats


def main():
    ref = Reference()
    rng = random.Random(20261202)
    shapes = ref_contractions_txt()
    patterns = list(FIXED_PATTERNS) + [s["pattern"] for s in shapes]
    for _ in range(400):
        patterns.append(random_contraction(rng, max_rank=6, allow_scalar=True)[0])
    for rank in range(1, 9):
        for _ in range(4):
            patterns.append(random_permutation_pattern(rng, rank)[0])
    seen = set()
    pat_rows = []
    for p in patterns:
        if p in seen:
            continue
        seen.add(p)
        err, dr, lr, rr, cj, dig = ref.parse(p)
        row = {"pattern": p, "err": err}
        if err == 0:
            row.update({"drank": dr, "lrank": lr, "rrank": rr, "conj": cj, "dig": dig})
            serr, sym = ref.pattern_sym(lr, rr, cj, dig)
            if serr == 0:
                row["sym"] = sym
        pat_rows.append(row)
    with open(os.path.join(HERE, "patterns.json"), "w") as f:
        json.dump({"patterns": pat_rows, "tensor_contractions_txt": shapes}, f, separators=(",", ":"))

    perm_rows = []
    for row in [r for r in pat_rows if r["err"] == 0][:240]:
        for tl, tr in ((1, 0), (0, 0), (1, 1), (0, 1)):
            for cj in (0, 1, 2, 4, 6):
                err, dp, lp, rp, ncd, nlu, nru = ref.contr_permutations(tl, tr, row["lrank"], row["rrank"],
                                                                          row["dig"], cj)
                r = {"lrank": row["lrank"], "rrank": row["rrank"], "dig": row["dig"], "conj": cj,
                     "gemm_tl": tl, "gemm_tr": tr, "err": err}
                if err == 0:
                    r.update({"dprm": dp, "lprm": lp, "rprm": rp, "ncd": ncd, "nlu": nlu, "nru": nru})
                perm_rows.append(r)
    with open(os.path.join(HERE, "contr_perms.json"), "w") as f:
        json.dump(perm_rows, f, separators=(",", ":"))

    # numeric cases: small seeded tensors, outputs from the reference CPU path
    cases, arrays = [], {}
    nrng = random.Random(7)
    for dk in (R4, R8, C4, C8):
        for t in range(6):
            rank = nrng.randint(1, 6)
            pat, perm, labels = random_permutation_pattern(nrng, rank)
            ldims = [nrng.randint(1, 4) for _ in range(rank)]
            err, dr, lr, rr, cj, dig = ref.parse(pat)
            ddims = [0] * rank
            for k, tt in enumerate(dig):
                ddims[tt - 1] = ldims[k]
            seed_l, seed_d = 1000 + 10 * t + dk, 5000 + 10 * t + dk
            L = tensor(dk, ldims, seed_l)
            key = "c%d" % len(cases)
            arrays[key] = ref.copy(dk, pat, ddims, ldims, L)
            cases.append({"key": key, "op": "copy", "dk": dk, "pattern": pat, "ddims": ddims, "ldims": ldims,
                          "seed_l": seed_l})
            alpha = (0.37, -0.21) if dk in (C4, C8) else (0.37, 0.0)
            D0 = tensor(dk, ddims, seed_d)
            apat = pat.replace(")=L", ")+=L")
            key = "c%d" % len(cases)
            arrays[key] = ref.add(dk, apat, ddims, D0, ldims, L, alpha)
            cases.append({"key": key, "op": "add", "dk": dk, "pattern": apat, "ddims": ddims, "ldims": ldims,
                          "seed_l": seed_l, "seed_d": seed_d, "alpha": list(alpha)})
        for t in range(8):
            pat, dd, ld, rd = random_contraction(nrng, max_rank=4, ext_lo=1, ext_hi=5, allow_scalar=(t > 5))
            sl, sr, sd = 11 * t + dk, 13 * t + dk + 1, 17 * t + dk + 2
            alpha = (0.5, 0.25) if dk in (C4, C8) else (0.75, 0.0)
            acc = t % 2
            out = ref.contract(dk, pat, dd, tensor(dk, dd, sd), ld, tensor(dk, ld, sl), rd, tensor(dk, rd, sr),
                               alpha, acc)
            key = "c%d" % len(cases)
            arrays[key] = out
            cases.append({"key": key, "op": "contract", "dk": dk, "pattern": pat, "ddims": dd, "ldims": ld,
                          "rdims": rd, "seed_l": sl, "seed_r": sr, "seed_d": sd, "alpha": list(alpha),
                          "accumulative": acc})
        if dk in (R4, R8):
            dims = [5, 7, 3]
            key = "c%d" % len(cases)
            arrays[key] = np.array([ref.norm2(dk, dims, tensor(dk, dims, 77 + dk))])
            cases.append({"key": key, "op": "norm2", "dk": dk, "ldims": dims, "seed_l": 77 + dk})
        dims = [4, 3, 5]
        alpha = (-1.25, 0.5) if dk in (C4, C8) else (-1.25, 0.0)
        key = "c%d" % len(cases)
        arrays[key] = ref.scale(dk, dims, tensor(dk, dims, 88 + dk), alpha)
        cases.append({"key": key, "op": "scale", "dk": dk, "ldims": dims, "seed_l": 88 + dk, "alpha": list(alpha)})
    # slices and inserts (talshTensorSlice / talshTensorInsert on the Host)
    srng = random.Random(41)
    for dk in (R4, R8, C4, C8):
        for t in range(6):
            rank = srng.randint(0, 5)
            bdims = [srng.randint(1, 6) for _ in range(rank)]
            sdims = [srng.randint(1, b) for b in bdims]
            offsets = [srng.randint(0, b - x) for b, x in zip(bdims, sdims)]
            seed_b, seed_s = 3000 + 10 * t + dk, 4000 + 10 * t + dk
            B, S = tensor(dk, bdims, seed_b), tensor(dk, sdims, seed_s)
            for op in ("slice", "insert"):
                acc = t % 2
                key = "c%d" % len(cases)
                fn = ref.slice if op == "slice" else ref.insert
                arrays[key] = fn(dk, sdims, S, bdims, B, offsets, acc)
                cases.append({"key": key, "op": op, "dk": dk, "sdims": sdims, "bdims": bdims, "offsets": offsets,
                              "seed_b": seed_b, "seed_s": seed_s, "accumulative": acc})
    # contractions with a rank-0 operand and outer products (the reference GPU path
    # has dedicated kernels for these, tensor_algebra_gpu_nvidia.cu:6732-6849)
    for dk in (R4, R8, C4, C8):
        for t, (pat, dd, ld, rd) in enumerate(RANK0_CASES):
            sl, sr, sd = 2000 + 7 * t + dk, 2100 + 7 * t + dk, 2200 + 7 * t + dk
            alpha = (0.5, -0.75) if dk in (C4, C8) else (-1.5, 0.0)
            acc = (t + dk) % 2
            key = "c%d" % len(cases)
            arrays[key] = ref.contract(dk, pat, dd, tensor(dk, dd, sd), ld, tensor(dk, ld, sl), rd,
                                       tensor(dk, rd, sr), alpha, acc)
            cases.append({"key": key, "op": "contract", "dk": dk, "pattern": pat, "ddims": dd, "ldims": ld,
                          "rdims": rd, "seed_l": sl, "seed_r": sr, "seed_d": sd, "alpha": list(alpha),
                          "accumulative": acc})
    np.savez_compressed(os.path.join(HERE, "numeric.npz"), **arrays)
    with open(os.path.join(HERE, "numeric.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print("patterns", len(pat_rows), "perms", len(perm_rows), "numeric", len(cases))


if __name__ == "__main__":
    main()
