"""Pins the CPU oracle (oracle/talsh_oracle.c) before it is trusted as the checker.

Every oracle function is checked against (a) golden vectors written by
TAL-SH's own CPU path (tests/golden/, see make_golden.py), and (b) facts the
mathematics fixes: closed forms on constant tensors, an independent loop-nest
evaluation, numpy einsum in float64, exact sums of small integers.  The last
test plants mistakes in oracle inputs to show that these checks detect them.
"""
import json
import os
import random

import numpy as np
import pytest

from tal_common import C4, C8, R4, R8, NP_DTYPE, einsum_contract, random_contraction, random_hyper_contraction, tensor

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_parse_matches_reference_golden(oracle):
    rows = _load("patterns.json")["patterns"]
    assert len(rows) > 400
    for row in rows:
        err, dr, lr, rr, cj, dig = oracle.parse(row["pattern"])
        if row["err"] != 0:
            assert err != 0, row["pattern"]
        else:
            assert err == 0, row["pattern"]
            assert (dr, lr, rr, cj, dig) == (row["drank"], row["lrank"], row["rrank"], row["conj"], row["dig"])


def test_parse_invariants(oracle):
    """Each label occurs exactly twice; contracted pairs point at each other."""
    rng = random.Random(3)
    for _ in range(300):
        pat, dd, ld, rd = random_contraction(rng, max_rank=6, allow_scalar=True)
        err, dr, lr, rr, cj, dig = oracle.parse(pat)
        assert err == 0
        assert (dr, lr, rr) == (len(dd), len(ld), len(rd))
        pos = sorted(x for x in dig if x > 0)
        assert pos == list(range(1, dr + 1))
        for i in range(lr):
            if dig[i] < 0:
                assert dig[lr - dig[i] - 1] == -(i + 1)


def test_contr_permutations_match_reference_golden(oracle):
    rows = _load("contr_perms.json")
    assert len(rows) > 1000
    for r in rows:
        got = oracle.contr_permutations(r["gemm_tl"], r["gemm_tr"], r["lrank"], r["rrank"], r["dig"], r["conj"])
        if r["err"] != 0:
            assert got[0] != 0
        else:
            assert got == (0, r["dprm"], r["lprm"], r["rprm"], r["ncd"], r["nlu"], r["nru"]), r


def _inputs(case):
    dk = case["dk"]
    L = tensor(dk, case["ldims"], case["seed_l"])
    D = tensor(dk, case["ddims"], case["seed_d"]) if "seed_d" in case else None
    R = tensor(dk, case["rdims"], case["seed_r"]) if "seed_r" in case else None
    return L, R, D


def test_numeric_golden(oracle):
    cases = _load("numeric.json")
    arrays = np.load(os.path.join(GOLDEN, "numeric.npz"))
    seen = set()
    for c in cases:
        want = arrays[c["key"]]
        dk = c["dk"]
        seen.add((c["op"], dk))
        if c["op"] in ("slice", "insert"):
            B, S = tensor(dk, c["bdims"], c["seed_b"]), tensor(dk, c["sdims"], c["seed_s"])
            fn = oracle.slice if c["op"] == "slice" else oracle.insert
            got = fn(dk, c["sdims"], S, c["bdims"], B, c["offsets"], c["accumulative"])
            assert got.tobytes() == want.astype(NP_DTYPE[dk]).tobytes(), c
            continue
        L, R, D = _inputs(c)
        if c["op"] in ("copy", "add"):
            err, dr, lr, rr, cj, dig = oracle.parse(c["pattern"])
            o2n = dig
            if c["op"] == "copy":
                got = oracle.copy(dk, c["ldims"], o2n, L)
            else:
                got = oracle.add(dk, c["ldims"], o2n, L, D, tuple(c["alpha"]))
            assert got.tobytes() == want.astype(NP_DTYPE[dk]).tobytes(), c
        elif c["op"] == "contract":
            got = oracle.contract(dk, c["pattern"], c["ddims"], D, c["ldims"], L, c["rdims"], R,
                                  tuple(c["alpha"]), c["accumulative"])
            tol = 1e-13 if dk in (R8, C8) else 2e-6
            scale = max(1.0, np.max(np.abs(want)))
            assert np.max(np.abs(got - want)) <= tol * scale, c
        elif c["op"] == "norm2":
            got = oracle.norm2(dk, L)
            assert abs(got - want[0]) <= 1e-13 * abs(want[0]), c
        elif c["op"] == "scale":
            got = oracle.scale(dk, L, tuple(c["alpha"]))
            assert got.tobytes() == want.astype(NP_DTYPE[dk]).tobytes(), c
    for op in ("copy", "add", "contract", "scale", "slice", "insert"):
        for dk in (R4, R8, C4, C8):
            assert (op, dk) in seen
    assert ("norm2", R8) in seen and ("norm2", R4) in seen


def test_copy_closed_form(oracle):
    """D[sum_k i_k*base_out[o2n[k]]] == L[sum_k i_k*base_in[k]] on every element."""
    rng = np.random.default_rng(5)
    for rank in range(1, 7):
        dims = list(rng.integers(1, 5, size=rank))
        o2n = list(rng.permutation(rank) + 1)
        L = np.arange(int(np.prod(dims)), dtype=np.float64)
        got = oracle.copy(R8, dims, o2n, L)
        # numpy: dlf = Fortran order
        A = L.reshape(dims, order="F")
        axes = [0] * rank
        for k, t in enumerate(o2n):
            axes[t - 1] = k
        want = np.transpose(A, axes).reshape(-1, order="F")
        assert np.array_equal(got, want)


def test_slice_insert_closed_form(oracle):
    """Big-tensor values encode their own dlf position, so a slice must hold
    sum_k (offset_k + i_k) * stride_k at slice position i; an insert followed by
    the same slice returns the inserted block; an accumulative slice adds."""
    rng = random.Random(9)
    for trial in range(60):
        rank = rng.randint(0, 5)
        bd = [rng.randint(1, 6) for _ in range(rank)]
        sd = [rng.randint(1, b) for b in bd]
        off = [rng.randint(0, b - x) for b, x in zip(bd, sd)]
        big = np.arange(int(np.prod(bd)) if bd else 1, dtype=np.float64)
        sl = oracle.slice(R8, sd, np.zeros(int(np.prod(sd)) if sd else 1), bd, big, off)
        strides = np.cumprod([1] + bd[:-1]) if bd else []
        want = np.zeros(sl.size)
        for lin in range(sl.size):
            q, pos = lin, 0
            for k in range(rank):
                pos += (off[k] + q % sd[k]) * strides[k]
                q //= sd[k]
            want[lin] = pos
        assert np.array_equal(sl, want), (sd, bd, off)
        blk = -1.0 - np.arange(sl.size, dtype=np.float64)
        big2 = oracle.insert(R8, sd, blk, bd, big, off)
        assert np.array_equal(oracle.slice(R8, sd, np.zeros(sl.size), bd, big2, off), blk)
        assert int(np.sum(big2 != big)) == sl.size  # nothing outside the block moved
        acc = oracle.slice(R8, sd, np.full(sl.size, 0.5), bd, big, off, accumulative=1)
        assert np.array_equal(acc, want + 0.5)


def test_contract_closed_form_constant_tensors(oracle):
    """test.cpp:193-226 style: D = d0 + alpha * K * l * r for constant tensors."""
    pat = "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)"
    V, O = 6, 4
    dd, ld, rd = [V, V, O, O], [O, V, O, V], [V, V, V, V]
    D = np.full(int(np.prod(dd)), 0.25)
    L = np.full(int(np.prod(ld)), 0.01)
    R = np.full(int(np.prod(rd)), 0.001)
    got = oracle.contract(R8, pat, dd, D, ld, L, rd, R, (0.5, 0.0), 1)
    want = 0.25 + 0.5 * V * V * 0.01 * 0.001
    assert np.allclose(got, want, rtol=1e-15, atol=0)


def test_contract_matches_einsum_and_loop_nest(oracle):
    rng = random.Random(11)
    for t in range(60):
        dk = (R8, C8)[t % 2]
        pat, dd, ld, rd = random_contraction(rng, max_rank=4, ext_lo=1, ext_hi=5, allow_scalar=True)
        L, R, D = tensor(dk, ld, t), tensor(dk, rd, t + 100), tensor(dk, dd, t + 200)
        got = oracle.contract(dk, pat, dd, D, ld, L, rd, R, (0.75, 0.0), 1)
        # einsum over Fortran-ordered views
        lhs, rhs = pat.split("+=")
        dl = lhs[2:-1].replace(",", "")
        ll, rl = rhs.split("*")
        ll, rl = ll[2:-1].replace(",", ""), rl[2:-1].replace(",", "")
        E = np.einsum("%s,%s->%s" % (ll, rl, dl), L.reshape(ld, order="F") if ld else L.reshape(()),
                      R.reshape(rd, order="F") if rd else R.reshape(()))
        want = D + 0.75 * np.asarray(E).reshape(-1, order="F")
        assert np.max(np.abs(got - want)) <= 1e-14 * max(1.0, np.max(np.abs(want))), pat
        # independent loop nest on a few elements
        for off in (0, len(D) // 2, len(D) - 1):
            el = oracle.contract_element(dk, pat, dd, off, ld, L, rd, R)
            assert abs(D[off] + 0.75 * el - want[off]) <= 1e-13 * max(1.0, abs(want[off]))


def test_contract_direct_matches_einsum_incl_hyper(oracle):
    """orc_contract_direct (the direct loop nest) against numpy einsum for contractions with hyper
    indices, all data kinds, conjugated operands and both accumulate modes; and against the
    matricized orc_contract for ordinary patterns."""
    rng = random.Random(23)
    for t in range(80):
        dk = (R8, C8, R4, C4)[t % 4]
        pat, dd, ld, rd = random_hyper_contraction(rng, max_rank=5, ext_lo=1, ext_hi=5)
        cl = cr = False
        if dk in (C4, C8) and t % 3:
            cl, cr = (t % 3 == 1), (t % 3 == 2)
            pat = pat.replace("L(", "L+(") if cl else pat.replace("R(", "R+(")
        acc = t % 2
        alpha = complex(0.5, -0.75) if dk in (C4, C8) else 0.75
        L, R, D = tensor(dk, ld, t), tensor(dk, rd, t + 100), tensor(dk, dd, t + 200)
        a = (alpha.real, alpha.imag) if dk in (C4, C8) else (alpha, 0.0)
        got = oracle.contract_direct(dk, pat, dd, D, ld, L, rd, R, a, acc)
        want = einsum_contract(pat, dd, D, ld, L, rd, R, alpha, acc, cl, cr)
        tol = 1e-14 if dk in (R8, C8) else 2e-6
        assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want))), (pat, dk)
    for t in range(20):
        dk = (R8, C8)[t % 2]
        pat, dd, ld, rd = random_contraction(rng, max_rank=4, ext_lo=1, ext_hi=5, allow_scalar=True)
        L, R, D = tensor(dk, ld, t), tensor(dk, rd, t + 100), tensor(dk, dd, t + 200)
        got = oracle.contract_direct(dk, pat, dd, D, ld, L, rd, R, (0.75, 0.0), 1)
        ref = oracle.contract(dk, pat, dd, D, ld, L, rd, R, (0.75, 0.0), 1)
        assert np.max(np.abs(got - ref)) <= 1e-15 * max(1.0, np.max(np.abs(ref))), pat
    # invariant: summing a hyper result over its hyper index gives the contraction over that index
    L, R = tensor(R8, [3, 4, 5], 1), tensor(R8, [5, 4, 2], 2)
    hyp = oracle.contract_direct(R8, "D(a,h,b)+=L(a,h,k)*R(k,h,b)", [3, 4, 2], np.zeros(24), [3, 4, 5], L,
                                 [5, 4, 2], R, (1.0, 0.0), 0)
    full = oracle.contract(R8, "D(a,b)+=L(a,h,k)*R(k,h,b)", [3, 2], np.zeros(6), [3, 4, 5], L, [5, 4, 2], R,
                           (1.0, 0.0), 0)
    summed = hyp.reshape(3, 4, 2, order="F").sum(axis=1).reshape(-1, order="F")
    assert np.max(np.abs(summed - full)) <= 1e-14


def test_norm2_exact_small_integers(oracle):
    for dk in (R4, R8, C4, C8):
        n = 1000
        x = (np.arange(n) % 7 - 3).astype(np.float64)
        if dk in (C4, C8):
            a = (x + 1j * (np.arange(n) % 5 - 2)).astype(NP_DTYPE[dk])
            want = float(np.sum(x * x) + np.sum((np.arange(n) % 5 - 2.0) ** 2))
        else:
            a = x.astype(NP_DTYPE[dk])
            want = float(np.sum(x * x))
        assert oracle.norm2(dk, a) == want


def test_planted_mistakes_are_detected(oracle):
    """Mistakes planted in the oracle's inputs must fail the checks above."""
    rng = np.random.default_rng(1)
    dims = [3, 4, 5]
    L = rng.standard_normal(60)
    good = oracle.copy(R8, dims, [2, 3, 1], L)
    wrong = oracle.copy(R8, dims, [3, 2, 1], L)  # wrong permutation
    assert not np.array_equal(good, wrong)
    pat = "D(a,b)+=L(a,i)*R(i,b)"
    A, B, D = tensor(R8, [4, 5], 1), tensor(R8, [5, 3], 2), np.zeros(12)
    ok = oracle.contract(R8, pat, [4, 3], D, [4, 5], A, [5, 3], B)
    bad = oracle.contract(R8, pat, [4, 3], D, [4, 5], A, [5, 3], tensor(R8, [5, 3], 3))
    assert np.max(np.abs(ok - bad)) > 1e-3
    # a transposed operand (pattern mismatch) changes the result
    sq = tensor(R8, [4, 4], 7)
    r1 = oracle.contract(R8, "D(a,b)+=L(a,i)*R(i,b)", [4, 4], np.zeros(16), [4, 4], sq, [4, 4], sq)
    r2 = oracle.contract(R8, "D(a,b)+=L(i,a)*R(i,b)", [4, 4], np.zeros(16), [4, 4], sq, [4, 4], sq)
    assert np.max(np.abs(r1 - r2)) > 1e-3
