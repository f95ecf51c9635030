"""CPU tests of the C-ABI boundary and host logic of libtalsh_amd.so.

No kernel runs here: the library is loaded, its exported symbols are checked
against include/*.h, struct layouts against the reference ABI sizes, the
product's own pattern parser against the reference's golden vectors, and the
contraction planner by executing its plans with numpy and comparing with the
oracle.  Host-side error paths run with TAL-SH initialised on zero GPUs.
"""
import ctypes
import json
import os
import random
import re

import numpy as np
import pytest

import tal_sh_amd as ts
from tal_common import C8, R8, random_contraction, tensor

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def header_functions():
    names = set()
    for h in ("talsh.h", "talsh_amd.h"):
        text = open(os.path.join(REPO, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof", "typedef"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ts.load()
    names = header_functions()
    assert "talshTensorContract" in names and "talshTensorNorm2" in names and len(names) > 70
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference_abi():
    # sizes/offsets of the reference structs on x86-64 (talsh.h:87-121, tensor_algebra.h:221-252)
    assert ctypes.sizeof(ts.TalshTens) == 40
    assert ctypes.sizeof(ts.TalshTask) == 120
    assert ctypes.sizeof(ts.TalshTensArg) == 16
    assert ctypes.sizeof(ts.DevRsc) == 24
    assert ts.TalshTask.data_vol.offset == 96 and ts.TalshTask.exec_time.offset == 112
    assert ts.TalshTens.ndev.offset == 36 and ts.DevRsc.mem_attached.offset == 20
    # slice view and tensor operation (talsh.h:96-101, :124-140)
    assert ctypes.sizeof(ts.TalshTensSlice) == 56 and ctypes.sizeof(ts.TalshTensOp) == 584
    assert ts.TalshTensOp.task_handle.offset == 424 and ts.TalshTensOp.stage.offset == 548


def test_header_struct_layouts_compiled(tmp_path):
    """The C headers give the structs the layout of the ctypes mirrors above (compiled with gcc)."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text("""#include <stdio.h>
#include <stddef.h>
#include "talsh.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(talsh_tens_t), sizeof(talsh_task_t), sizeof(talsh_tens_slice_t),
         sizeof(talsh_tens_op_t), offsetof(talsh_tens_op_t, task_handle), offsetof(talsh_tens_op_t, stage),
         offsetof(talsh_tens_op_t, time_finished));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(ts.TalshTens), ctypes.sizeof(ts.TalshTask), ctypes.sizeof(ts.TalshTensSlice),
            ctypes.sizeof(ts.TalshTensOp), ts.TalshTensOp.task_handle.offset, ts.TalshTensOp.stage.offset,
            ts.TalshTensOp.time_finished.offset]
    assert got == want, (got, want)


def test_header_constants():
    text = open(os.path.join(REPO, "include", "tensor_algebra.h")).read()
    assert "#define TRY_LATER -918273645" in text and "#define DEVICE_UNABLE -546372819" in text
    assert ts.coh("M", "T", "T") == 26 and ts.coh("M", "T") == 6 and ts.coh("T", "T") == 10


def test_product_parser_matches_reference_golden():
    lib = ts.load()
    rows = json.load(open(os.path.join(GOLDEN, "patterns.json")))["patterns"]
    dig = (ctypes.c_int * 128)()
    dr, lr, rr, cj = (ctypes.c_int() for _ in range(4))
    for row in rows:
        err = lib.talsh_get_contr_ptrn_str2dig(row["pattern"].encode(), dig, ctypes.byref(dr), ctypes.byref(lr),
                                               ctypes.byref(rr), ctypes.byref(cj))
        if row["err"] != 0:
            assert err != 0, row["pattern"]
            continue
        assert err == 0, row["pattern"]
        n = lr.value + rr.value
        assert (dr.value, lr.value, rr.value, cj.value, list(dig[:n])) == (
            row["drank"], row["lrank"], row["rrank"], row["conj"], row["dig"]), row["pattern"]


def test_product_pattern_sym_matches_reference_golden():
    """get_contr_pattern_sym (digital -> symbolic) against the reference routine's
    outputs, and (for ordinary contractions) parsing its output gives back the
    same digital pattern."""
    rows = [r for r in json.load(open(os.path.join(GOLDEN, "patterns.json")))["patterns"] if "sym" in r]
    assert len(rows) > 400
    lib = ts.load()
    dig = (ctypes.c_int * 128)()
    dr, lr, rr, cj = (ctypes.c_int() for _ in range(4))
    for row in rows:
        err, sym = ts.pattern_sym(row["lrank"], row["rrank"], row["conj"], row["dig"])
        assert err == 0 and sym == row["sym"], (row["pattern"], sym, row["sym"])
        pos = [d for d in row["dig"] if d > 0]
        if len(pos) != len(set(pos)):
            continue  # hyper-contraction: the routine names D's indices per occurrence, as the reference does
        assert lib.talsh_get_contr_ptrn_str2dig(sym.encode(), dig, ctypes.byref(dr), ctypes.byref(lr),
                                                ctypes.byref(rr), ctypes.byref(cj)) == 0
        assert list(dig[:lr.value + rr.value]) == row["dig"] and cj.value == row["conj"]
    assert ts.pattern_sym(-1, 2, 0, [1, 2])[0] == 1


def plan(pattern, dd, ld, rd, acc=1):
    lib = ts.load()
    mnk = (ctypes.c_longlong * 3)()
    flags = (ctypes.c_int * 4)()
    ops = (ctypes.c_char * 2)()
    perms = (ctypes.c_int * 200)()
    moved = ctypes.c_double()
    ia = lambda a: (ctypes.c_int * max(1, len(a)))(*a)
    err = lib.talshAmdPlanContraction(pattern.encode(), ia(dd), ia(ld), ia(rd), acc, mnk, flags, ops, perms,
                                      ctypes.byref(moved))
    return err, list(mnk), list(flags), ops.raw.decode(), list(perms), moved.value


def execute_plan_numpy(oracle, dk, pattern, dd, D, ld, L, rd, R, alpha):
    """Runs the planner's TTGT recipe with numpy matmul (as the GPU path does)."""
    err, (M, N, K), (swap, at, bt, dt), ops, perms, moved = plan(pattern, dd, ld, rd)
    assert err == 0
    X, xd, Y, yd = (R, rd, L, ld) if swap else (L, ld, R, rd)
    q = 0
    pa = perms[q:q + len(xd)]; q += len(xd)
    pb = perms[q:q + len(yd)]; q += len(yd)
    pd = perms[q:q + len(dd)]
    if at:
        X = oracle.copy(dk, xd, [p + 1 for p in pa], X)
    if bt:
        Y = oracle.copy(dk, yd, [p + 1 for p in pb], Y)
    A = X.reshape((K, M) if ops[0] == "T" else (M, K), order="F")
    B = Y.reshape((K, N) if ops[1] == "N" else (N, K), order="F")
    opA = A.T if ops[0] == "T" else A
    opB = B if ops[1] == "N" else B.T
    if dt:
        Dp = oracle.copy(dk, dd, [p + 1 for p in pd], D)
        dpd = [0] * len(dd)
        for i, p in enumerate(pd):
            dpd[p] = dd[i]
    else:
        Dp, dpd = D.copy(), dd
    C = Dp.reshape((M, N), order="F") + alpha * (opA @ opB)
    C = C.reshape(-1, order="F")
    if dt:
        inv = [0] * len(dd)
        for i, p in enumerate(pd):
            inv[p] = i
        return oracle.copy(dk, dpd, [i + 1 for i in inv], C), moved
    return C, moved


def test_planner_recipes_reproduce_oracle(oracle):
    rng = random.Random(17)
    for t in range(150):
        dk = R8
        pat, dd, ld, rd = random_contraction(rng, max_rank=5, ext_lo=1, ext_hi=4, allow_scalar=True)
        L, R, D = tensor(dk, ld, t), tensor(dk, rd, t + 1), tensor(dk, dd, t + 2)
        got, moved = execute_plan_numpy(oracle, dk, pat, dd, D, ld, L, rd, R, 0.5)
        want = oracle.contract(dk, pat, dd, D, ld, L, rd, R, (0.5, 0.0), 1)
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want))), pat


def test_planner_avoids_transposes_when_layout_allows():
    # plain matrix products in every storage form need no copies
    for pat in ("D(a,b)+=L(a,i)*R(i,b)", "D(a,b)+=L(i,a)*R(i,b)", "D(a,b)+=L(a,i)*R(b,i)",
                "D(a,b)+=L(i,a)*R(b,i)", "D(b,a)+=L(a,i)*R(i,b)", "D(a,b,c)+=L(a,b,i)*R(i,c)"):
        ld = {"D(a,b,c)+=L(a,b,i)*R(i,c)": [3, 4, 5]}.get(pat, [3, 5])
        err, mnk, flags, ops, perms, moved = plan(pat, [3, 4, 6] if "c" in pat else [3, 4] if pat[2] == "a" else
                                                  [4, 3], ld if "c" in pat else
                                                  ([3, 5] if "L(a" in pat else [5, 3]),
                                                  [5, 6] if "c" in pat else ([5, 4] if "R(i" in pat else [4, 5]))
        assert err == 0, pat
        assert moved == 0 and flags[1:] == [0, 0, 0], (pat, flags)
    # the configs' patterns interleave indices: all three operands are transposed
    err, mnk, flags, ops, perms, moved = plan("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [32] * 4, [32] * 4, [32] * 4)
    assert err == 0 and mnk == [1024, 1024, 1024] and flags[1:] == [1, 1, 1]



def test_plan_cache_returns_each_shape_its_own_plan():
    """Plans are memoised by (ranks, digits, extents, accumulate): interleaved repeats of different shapes,
    patterns and accumulate flags return exactly what their first planning returned, and shapes that differ
    only in one extent or in the accumulate flag are told apart."""
    import random
    from tal_common import random_contraction
    rng = random.Random(31337)
    cases = []
    for t in range(40):
        pat, dd, ld, rd = random_contraction(rng, max_rank=5, ext_lo=1, ext_hi=7)
        cases.append((pat, dd, ld, rd, t % 2))
    # near-duplicates: one extent changed, or the other accumulate flag
    pat, dd, ld, rd, acc = cases[0]
    cases.append((pat, dd, ld, rd, 1 - acc))
    cases.append(("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [32] * 4, [32] * 4, [32] * 4, 1))
    cases.append(("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [32, 32, 32, 16], [32, 32, 32, 32], [16, 32, 32, 32], 1))
    first = [plan(p_, d_, l_, r_, a_) for (p_, d_, l_, r_, a_) in cases]
    assert all(f[0] == 0 for f in first)
    assert sorted(first[-1][1][:2]) == [512, 1024] and first[-1][1][2] == 1024  # (A and B may be swapped)
    assert first[-2][1] == [1024, 1024, 1024]
    order = list(range(len(cases))) * 3
    rng.shuffle(order)
    for i in order:
        p_, d_, l_, r_, a_ = cases[i]
        assert plan(p_, d_, l_, r_, a_) == first[i], cases[i]

def test_host_side_errors_without_gpu():
    lib = ts.load()
    hb, ham = ctypes.c_size_t(0), ctypes.c_int(0)
    assert lib.talshInit(ctypes.byref(hb), ctypes.byref(ham), 0, None, 0, None, 0, None) == 0
    try:
        assert lib.talshInit(ctypes.byref(hb), ctypes.byref(ham), 0, None, 0, None, 0, None) == \
            ts.TALSH_ALREADY_INITIALIZED
        d, l, r = (ts.Tensor(R8, dims, 0) for dims in ([2, 3], [2, 4], [4, 3]))
        # invalid pattern / shape mismatch / D aliasing an input are rejected before any device work
        assert ts.contract("D(a,b)+=L(a,i)*R(i", d, l, r) == ts.TALSH_INVALID_ARGS
        sq = ts.Tensor(R8, [2, 2], 0)
        assert ts.contract("D(a,b)+=L(a,i)*R(i,b)", sq, sq, sq) == ts.TALSH_INVALID_ARGS
        sq.destruct()
        assert ts.contract("D(a,b)+=L(a,i)*R(b,i)", d, l, r) == ts.TALSH_INVALID_ARGS
        # host execution is not provided by this build
        assert ts.contract("D(a,b)+=L(a,i)*R(i,b)", d, l, r, dev_kind=ts.DEV_HOST, dev_id=0) == \
            ts.TALSH_NOT_AVAILABLE
        assert ts.contract("D(a,b)+=L(a,i)*R(i,b)", d, l, r) == ts.TALSH_NOT_AVAILABLE
        # hyper-contractions are accepted (batched GEMM); with no GPU they are not available either
        dh, lh, rh = ts.Tensor(R8, [2, 3], 0), ts.Tensor(R8, [2, 4, 3], 0), ts.Tensor(R8, [4, 3], 0)
        assert ts.contract("D(a,c)+=L(a,i,c)*R(i,c)", dh, lh, rh) == ts.TALSH_NOT_AVAILABLE
        assert plan("D(a,c)+=L(a,i,c)*R(i,c)", [2, 3], [2, 4, 3], [4, 3])[:2] == (0, [2, 1, 4])
        # host images: import/export and scalar access
        a = tensor(R8, [2, 3], 5)
        d.upload(a)
        assert np.array_equal(d.host_array(), a)
        s = ts.Tensor(R8, [], 0, init_val=2.5)
        assert s.scalar() == 2.5
        assert lib.talshTensorRank(d.ptr) == 2 and lib.talshTensorVolume(d.ptr) == 6
        # task handles
        task = ts.Task()
        assert lib.talshTaskIsEmpty(task.ptr) == ts.YEP
        st = ctypes.c_int()
        assert lib.talshTaskWait(task.ptr, ctypes.byref(st)) == ts.TALSH_OBJECT_IS_EMPTY
        for t in (d, l, r, dh, lh, rh, s):
            t.destruct()
    finally:
        assert lib.talshShutdown() == 0


def test_host_memory_policy():
    """talshSetMemAllocPolicyHost (talshf.F90:882-893): Host bodies from the pinned Host buffer or ordinary memory,
    with or without fallback when the buffer cannot hold them.  No GPU needed (without a GPU driver the pinned
    buffer cannot be allocated and has size 0, so every body needs the fallback)."""
    lib = ts.load()
    hb, ham = ctypes.c_size_t(1 << 20), ctypes.c_int(0)
    assert lib.talshInit(ctypes.byref(hb), ctypes.byref(ham), 0, None, 0, None, 0, None) == 0
    err = ctypes.c_int(-1)
    try:
        lo = lib.talshDeviceBufferBasePtr(0, ts.DEV_HOST)
        inside = (lambda t: lo <= t.host_address() < lo + hb.value) if lo else (lambda t: False)
        small = ts.Tensor(R8, [16, 16], 0)                    # default: Host buffer when there is one
        assert inside(small) == bool(lo)
        lib.talshSetMemAllocPolicyHost(ts.MEM_ALLOC_ALL_BUF, 0, ctypes.byref(err))
        assert err.value == 0
        with pytest.raises(RuntimeError):                      # larger than the buffer, no fallback
            ts.Tensor(R8, [1024, 1024], 0)
        lib.talshSetMemAllocPolicyHost(ts.MEM_ALLOC_TMP_BUF, 1, ctypes.byref(err))
        big = ts.Tensor(R8, [1024, 1024], 0)                   # falls back to ordinary memory
        assert not inside(big)
        lib.talshSetMemAllocPolicyHost(ts.MEM_ALLOC_REGULAR, 0, ctypes.byref(err))
        reg = ts.Tensor(R8, [16, 16], 0)
        assert not inside(reg)
        lib.talshSetMemAllocPolicyHost(7, 0, ctypes.byref(err))
        assert err.value != 0
        for t in (small, big, reg):
            t.destruct()
    finally:
        lib.talshSetMemAllocPolicyHost(ts.MEM_ALLOC_TMP_BUF, 1, ctypes.byref(err))
        assert lib.talshShutdown() == 0


def test_fast_math_split_selection():
    """talshAmdSetFastMathSplit / talshAmdGetFastMathSplit (include/talsh_amd.h): bf16x6 by default, fp16x3 on
    request, anything else refused and leaving the choice unchanged; fast math itself (talshEnableFastMath,
    talsh.h:167-174) is queried per device kind.  No GPU needed."""
    lib = ts.load()
    assert lib.talshAmdGetFastMathSplit() == ts.SPLIT_BF16X6
    try:
        assert lib.talshAmdSetFastMathSplit(ts.SPLIT_FP16X3) == 0
        assert lib.talshAmdGetFastMathSplit() == ts.SPLIT_FP16X3
        assert lib.talshAmdSetFastMathSplit(7) == ts.TALSH_INVALID_ARGS
        assert lib.talshAmdSetFastMathSplit(-1) == ts.TALSH_INVALID_ARGS
        assert lib.talshAmdGetFastMathSplit() == ts.SPLIT_FP16X3
        assert ts.fast_math(True, ts.SPLIT_BF16X6) == 0
        assert lib.talshQueryFastMath(ts.DEV_NVIDIA_GPU, 0) == ts.YEP
        assert lib.talshAmdGetFastMathSplit() == ts.SPLIT_BF16X6
        assert ts.fast_math(False) == 0
        assert lib.talshQueryFastMath(ts.DEV_NVIDIA_GPU, 0) == ts.NOPE
    finally:
        ts.fast_math(False)
        lib.talshAmdSetFastMathSplit(ts.SPLIT_BF16X6)
