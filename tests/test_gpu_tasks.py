"""Ordering of dependent asynchronous tasks, and contractions with rank-0 operands.

The reference orders every task on a GPU after the previous task's compute
event (LastTask, tensor_algebra_gpu_nvidia.cu:6287-6295).  This library
instead tracks hazards per byte range across several streams, so these tests
submit chains of dependent tasks without waiting in between and compare every
result with the oracle:
  RAW   a task reads memory an in-flight task is writing,
  WAR   a task overwrites an operand an in-flight task is reading,
  WAW   two tasks write (or accumulate into) the same memory,
  a copy or an add feeding a contraction.
Within the TAL-SH API an in-flight output is unavailable to other operations
(its image is marked in use, as in the reference), so reads and writes of the
same bytes through *different* tensors are made with views: tensors
constructed on external memory inside another tensor's GPU body, at the same
or at an overlapping address range.

Also here: a test_talsh_qc-style pipeline (test.cpp:656-892) that keeps
retrying TRY_LATER while earlier tasks complete, tensors destructed before
their task is waited on, the rank-0 / outer-product contractions that the
reference runs with dedicated kernels (tensor_algebra_gpu_nvidia.cu:6732-6849),
repeated shapes in flight (plans, paired transposes and D offset tables
reused across streams), permutation shapes repeated while their kernel
variants are timed and one is chosen, operands that are views at every
alignment (the scalar kernel paths), GEMMs of at most 32 columns (the
narrow-tile path) and high-rank contractions.

Bars: copies and initialisations bit-exact; R8/C8 contractions normwise
relative error <= 1e-12 against the oracle; R4/C4 contractions elementwise
within 1e-6 * (|D| + |alpha| * sum |l| |r|) of it (dot products cancel).
"""
import ctypes
import gc

import numpy as np
import pytest

from tal_common import C4, C8, R4, R8, NP_DTYPE, tensor

pytestmark = pytest.mark.gpu

CFG1 = "D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)"


def say(*a):
    print("[case]", *a, flush=True)


def relerr(got, want):
    den = np.linalg.norm(want.astype(np.complex128))
    return np.linalg.norm((got - want).astype(np.complex128)) / (den if den > 0 else 1.0)


def labels(n):
    return ",".join("i%d" % k for k in range(n))


def on_gpu(ts, dk, dims, data):
    t = ts.Tensor(dk, dims, ts.gpu_flat(0))
    h = ts.Tensor(dk, dims, 0, data=data)
    lab = labels(len(dims))
    ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), t, h), "upload")
    h.destruct()
    return t


def snapshot(ts, t):
    """Values of t's GPU image, read through a copy into a new Host tensor (t keeps its image)."""
    h = ts.Tensor(t.dk, t.dims, 0)
    lab = labels(len(t.dims))
    ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), h, t, copy_ctrl=ts.coh("T", "T")), "snapshot")
    out = h.host_array()
    h.destruct()
    return out


def gpu_address(ts, t):
    body = ctypes.c_void_p()
    ts.check(ts.load().talshTensorGetBodyAccess(t.ptr, ctypes.byref(body), t.dk, 0, ts.DEV_NVIDIA_GPU),
             "talshTensorGetBodyAccess")
    return body.value


def view(ts, dk, dims, address):
    """Tensor on external GPU memory at `address` (attached: it owns nothing)."""
    return ts.Tensor(dk, dims, ts.gpu_flat(0), ext=address)


def finish(*tasks):
    for tk in tasks:
        assert tk.wait() == 2000005  # TALSH_TASK_COMPLETED
        assert tk.destruct() == 0


def test_raw_view_reads_output_of_inflight_contraction(talsh_gpu, oracle):
    """B reads (through a view) the bytes that in-flight contraction A writes."""
    ts = talsh_gpu
    n = 40
    dims = [n] * 4
    L, R, D0 = tensor(R8, dims, 11), tensor(R8, dims, 12), tensor(R8, dims, 13)
    lt, rt = on_gpu(ts, R8, dims, L), on_gpu(ts, R8, dims, R)
    x = on_gpu(ts, R8, dims, D0)
    addr = gpu_address(ts, x)
    dv, lv = view(ts, R8, dims, addr), view(ts, R8, dims, addr)
    R2 = tensor(R8, [n, n], 14)
    r2 = on_gpu(ts, R8, [n, n], R2)
    d2 = on_gpu(ts, R8, [n, n], np.zeros(n * n))
    say("RAW view", n)
    ta, tb = ts.Task(), ts.Task()
    ts.check(ts.contract(CFG1, dv, lt, rt, alpha=0.5, task=ta), "A: contract into the view")
    ts.check(ts.contract("D(a,b)+=L(a,i,j,b)*R(i,j)", d2, lv, r2, task=tb), "B: contract from the view")
    finish(tb, ta)
    want_a = oracle.contract(R8, CFG1, dims, D0, dims, L, dims, R, (0.5, 0.0), 1)
    want_b = oracle.contract(R8, "D(a,b)+=L(a,i,j,b)*R(i,j)", [n, n], np.zeros(n * n), dims, want_a, [n, n], R2)
    for v in (dv, lv):
        v.destruct()
    assert relerr(snapshot(ts, x), want_a) <= 1e-12
    assert relerr(snapshot(ts, d2), want_b) <= 1e-12


def test_raw_same_tensor_is_unavailable_until_complete(talsh_gpu, oracle):
    """The TAL-SH contract: an in-flight output's image is in use (NOT_AVAILABLE) until its task
    is complete; afterwards the dependent operation runs and sees the result."""
    ts = talsh_gpu
    dims = [16] * 4
    L, R = tensor(R8, dims, 21), tensor(R8, dims, 22)
    lt, rt = on_gpu(ts, R8, dims, L), on_gpu(ts, R8, dims, R)
    d = on_gpu(ts, R8, dims, np.zeros(16 ** 4))
    e = on_gpu(ts, R8, dims, np.zeros(16 ** 4))
    ta = ts.Task()
    ts.check(ts.contract(CFG1, d, lt, rt, task=ta), "A")
    assert ts.copy("D(a,b,c,d)=L(d,c,b,a)", e, d) == ts.TALSH_NOT_AVAILABLE
    finish(ta)
    ts.check(ts.copy("D(a,b,c,d)=L(d,c,b,a)", e, d), "B after A")
    want = oracle.contract(R8, CFG1, dims, np.zeros(16 ** 4), dims, L, dims, R)
    got = snapshot(ts, d)
    assert relerr(got, want) <= 1e-12
    assert snapshot(ts, e).tobytes() == oracle.copy(R8, dims, [4, 3, 2, 1], got).tobytes()


def test_war_overwrite_operand_of_inflight_contraction(talsh_gpu, oracle):
    """B overwrites L (init) and R (scale) while contraction A is reading them."""
    ts = talsh_gpu
    n = 40
    dims = [n] * 4
    L, R = tensor(R8, dims, 31), tensor(R8, dims, 32)
    lt, rt = on_gpu(ts, R8, dims, L), on_gpu(ts, R8, dims, R)
    d = on_gpu(ts, R8, dims, np.zeros(n ** 4))
    say("WAR", n)
    ta, tb, tc = ts.Task(), ts.Task(), ts.Task()
    ts.check(ts.contract(CFG1, d, lt, rt, task=ta), "A reads L and R")
    ts.check(ts.init_value(lt, 0.0, task=tb), "B overwrites L")
    ts.check(ts.scale(rt, -2.0, task=tc), "C scales R")
    finish(tb, tc, ta)
    want = oracle.contract(R8, CFG1, dims, np.zeros(n ** 4), dims, L, dims, R)
    assert relerr(snapshot(ts, d), want) <= 1e-12
    assert not np.any(snapshot(ts, lt))
    assert snapshot(ts, rt).tobytes() == (-2.0 * R).tobytes()


def test_war_host_image_read_by_inflight_upload(talsh_gpu, oracle):
    """A uploads the Host image of L (COPY_TTT); B writes the same Host image back (scale with
    COPY_T on a Host-only tensor: upload, scale, copy back) and must wait for A's read."""
    ts = talsh_gpu
    n = 36
    dims = [n] * 4
    L, R = tensor(R8, dims, 41), tensor(R8, dims, 42)
    lh, rh = ts.Tensor(R8, dims, 0, data=L), ts.Tensor(R8, dims, 0, data=R)
    dh = ts.Tensor(R8, dims, 0, data=np.zeros(n ** 4))
    say("WAR host", n)
    ta, tb = ts.Task(), ts.Task()
    ts.check(ts.contract(CFG1, dh, lh, rh, copy_ctrl=ts.coh("T", "T", "T"), task=ta), "A")
    ts.check(ts.scale(rh, 3.0, copy_ctrl=ts.COPY_T, task=tb), "B")
    finish(tb, ta)
    want = oracle.contract(R8, CFG1, dims, np.zeros(n ** 4), dims, L, dims, R)
    assert relerr(dh.host_array(), want) <= 1e-12
    assert rh.host_array().tobytes() == (3.0 * R).tobytes()


def test_waw_accumulate_twice_into_one_body(talsh_gpu, oracle):
    """A and B accumulate into two views of the same bytes: both contributions land."""
    ts = talsh_gpu
    n = 40
    dims = [n] * 4
    L, R, D0 = tensor(R8, dims, 51), tensor(R8, dims, 52), tensor(R8, dims, 53)
    lt, rt = on_gpu(ts, R8, dims, L), on_gpu(ts, R8, dims, R)
    x = on_gpu(ts, R8, dims, D0)
    addr = gpu_address(ts, x)
    v1, v2 = view(ts, R8, dims, addr), view(ts, R8, dims, addr)
    say("WAW accumulate", n)
    ta, tb = ts.Task(), ts.Task()
    ts.check(ts.contract(CFG1, v1, lt, rt, alpha=1.0, task=ta), "A")
    ts.check(ts.contract(CFG1, v2, lt, rt, alpha=-0.25, task=tb), "B")
    finish(tb, ta)
    mid = oracle.contract(R8, CFG1, dims, D0, dims, L, dims, R, (1.0, 0.0), 1)
    want = oracle.contract(R8, CFG1, dims, mid, dims, L, dims, R, (-0.25, 0.0), 1)
    v1.destruct()
    v2.destruct()
    assert relerr(snapshot(ts, x), want) <= 1e-12


def test_waw_overlapping_views_keep_program_order(talsh_gpu, oracle):
    """A overwrites view 1 (first half of a body) with a contraction; B then initialises view 2,
    which overlaps the second half of view 1 and starts at another address.  The overlap must hold
    B's value, the rest A's result, the untouched tail its initial value (bit-exact)."""
    ts = talsh_gpu
    n = 40
    dims = [n] * 4
    v = n ** 4
    L, R, X0 = tensor(R8, dims, 61), tensor(R8, dims, 62), tensor(R8, [2 * v], 63)
    lt, rt = on_gpu(ts, R8, dims, L), on_gpu(ts, R8, dims, R)
    x = on_gpu(ts, R8, [2 * v], X0)
    addr = gpu_address(ts, x)
    v1 = view(ts, R8, dims, addr)
    v2 = view(ts, R8, [v], addr + 8 * (v // 2))
    say("WAW overlapping views", n)
    ta, tb = ts.Task(), ts.Task()
    ts.check(ts.contract(CFG1, v1, lt, rt, accumulative=ts.NOPE, task=ta), "A overwrites view 1")
    ts.check(ts.init_value(v2, 7.0, task=tb), "B initialises view 2")
    finish(tb, ta)
    v1.destruct()
    v2.destruct()
    got = snapshot(ts, x)
    want_a = oracle.contract(R8, CFG1, dims, np.zeros(v), dims, L, dims, R, (1.0, 0.0), 0)
    h = v // 2
    assert relerr(got[:h], want_a[:h]) <= 1e-12
    assert np.all(got[h:h + v] == 7.0)
    assert got[h + v:].tobytes() == X0[h + v:].tobytes()


def test_copy_and_add_feed_contraction(talsh_gpu, oracle):
    """A copy (permutation) and then an add write L through one view; a contraction reads L through
    another view of the same bytes.  No waits between the three tasks."""
    ts = talsh_gpu
    n = 40
    dims = [n] * 4
    S, S2, R = tensor(R8, dims, 71), tensor(R8, dims, 72), tensor(R8, dims, 73)
    st, s2t, rt = on_gpu(ts, R8, dims, S), on_gpu(ts, R8, dims, S2), on_gpu(ts, R8, dims, R)
    x = on_gpu(ts, R8, dims, np.zeros(n ** 4))
    addr = gpu_address(ts, x)
    lw, lr = view(ts, R8, dims, addr), view(ts, R8, dims, addr)
    d = on_gpu(ts, R8, dims, np.zeros(n ** 4))
    say("copy/add -> contract", n)
    ta, tb, tc = ts.Task(), ts.Task(), ts.Task()
    ts.check(ts.copy("D(a,b,c,d)=L(c,d,a,b)", lw, st, task=ta), "A copy")
    # lw is the output of A (in use): the add goes through the second view, the contraction
    # through a third one at the same address
    ts.check(ts.add("D(a,b,c,d)+=L(d,c,b,a)", lr, s2t, alpha=0.5, task=tb), "B add")
    l3 = view(ts, R8, dims, addr)
    ts.check(ts.contract(CFG1, d, l3, rt, task=tc), "C contract")
    finish(tc, tb, ta)
    lval = oracle.add(R8, dims, [4, 3, 2, 1], S2, oracle.copy(R8, dims, [3, 4, 1, 2], S), (0.5, 0.0))
    want = oracle.contract(R8, CFG1, dims, np.zeros(n ** 4), dims, lval, dims, R)
    for t in (lw, lr, l3):
        t.destruct()
    assert snapshot(ts, x).tobytes() == lval.tobytes()
    assert relerr(snapshot(ts, d), want) <= 1e-12


def test_qc_pipeline_retries_try_later(talsh_gpu, oracle):
    """test_talsh_qc-style loop (test.cpp:656-892): 40 asynchronous config-1 contractions (32^4 R8)
    on Host operands through a nearly full pool; TRY_LATER is retried while earlier tasks complete.
    Every result is checked; the pool returns to its free size."""
    ts = talsh_gpu
    lib = ts.load()
    gc.collect()
    dims = [32] * 4
    L, R = tensor(R8, dims, 81), tensor(R8, dims, 82)
    lh, rh = ts.Tensor(R8, dims, 0, data=L), ts.Tensor(R8, dims, 0, data=R)
    free0 = lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU)
    # ballast: leave about 120 MB of the pool free (room for ~3 tasks of ~40 MB)
    ballast, chunk = [], 1 << 33  # 2^30 R8 elements: within a 32-bit extent
    while lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) > (120 << 20) + (64 << 20) and chunk >= (16 << 20):
        room = lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) - (120 << 20)
        want = min(chunk, room) // 8
        t = ts.TalshTens()
        lib.talshTensorClean(ctypes.byref(t))
        dd = (ctypes.c_int * 1)(int(want))
        e = lib.talshTensorConstruct(ctypes.byref(t), R8, 1, dd, ts.gpu_flat(0), None, -1, None, 0.0, 0.0)
        if e == 0:
            ballast.append(t)
        else:
            assert e == ts.TRY_LATER, e
            chunk //= 2
    say("qc pipeline: pool free", lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU))
    try:
        base = oracle.contract(R8, CFG1, dims, np.zeros(32 ** 4), dims, L, dims, R)
        d0 = [tensor(R8, dims, 100 + k) for k in range(40)]
        dh = [ts.Tensor(R8, dims, 0, data=d0[k]) for k in range(40)]
        pending, inflight, try_later, done = list(range(40)), {}, 0, []
        while pending or inflight:
            if pending:
                k = pending[0]
                tk = ts.Task()
                e = ts.contract(CFG1, dh[k], lh, rh, alpha=0.5 + k, copy_ctrl=ts.coh("T", "T", "T"), task=tk)
                if e == 0:
                    inflight[k] = tk
                    pending.pop(0)
                else:
                    assert e == ts.TRY_LATER, e
                    assert lib.talshTaskIsEmpty(tk.ptr) == ts.YEP
                    assert inflight, "TRY_LATER with nothing in flight"
                    try_later += 1
            for k, tk in list(inflight.items()):
                fin, st, ierr = tk.complete()
                if fin:
                    assert st == ts.TALSH_TASK_COMPLETED and ierr == 0
                    assert tk.destruct() == 0
                    done.append(k)
                    del inflight[k]
        say("qc pipeline: TRY_LATER returned", try_later, "times")
        assert try_later > 0 and sorted(done) == list(range(40))
        for k in range(40):
            assert relerr(dh[k].host_array(), d0[k] + (0.5 + k) * base) <= 1e-12, k
        for t in dh:
            t.destruct()
    finally:
        for t in ballast:
            lib.talshTensorDestruct(ctypes.byref(t))
    lh.destruct()
    rh.destruct()
    assert lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) == free0


def test_destruct_before_wait_releases_task_buffers(talsh_gpu, oracle):
    """Tensors destructed while their task is in flight: the task completes first, every pool buffer
    it owns is released, and the handle then reports COMPLETED."""
    ts = talsh_gpu
    lib = ts.load()
    gc.collect()
    dims = [24] * 4
    L, R = tensor(R8, dims, 91), tensor(R8, dims, 92)
    free0 = lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU)
    for c in (ts.coh("T", "T", "T"), ts.coh("M", "T", "T"), ts.coh("K", "K", "D")):
        say("destruct before wait", c)
        lh, rh = ts.Tensor(R8, dims, 0, data=L), ts.Tensor(R8, dims, 0, data=R)
        dh = ts.Tensor(R8, dims, 0, data=np.zeros(24 ** 4))
        tk = ts.Task()
        ts.check(ts.contract(CFG1, dh, lh, rh, copy_ctrl=c, task=tk), "contract")
        for t in (dh, lh, rh):
            t.destruct()
        assert tk.wait() == ts.TALSH_TASK_COMPLETED
        assert tk.destruct() == 0
        assert lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) == free0
    # a Place in flight whose tensor goes away
    src = ts.Tensor(R8, dims, 0, data=L)
    tk = ts.Task()
    assert lib.talshTensorPlace(src.ptr, ts.gpu_flat(0), ts.DEV_NULL, None, ts.COPY_K, tk.ptr) == 0
    src.destruct()
    assert tk.wait() == ts.TALSH_TASK_COMPLETED
    assert tk.destruct() == 0
    assert lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) == free0


RANK0 = [
    ("D(a)+=L(a)*R()", [1000], [1000], []),
    ("D(a,b,c)+=L(c,a,b)*R()", [20, 30, 40], [40, 20, 30], []),
    ("D(a,b)+=L()*R(b,a)", [64, 48], [], [48, 64]),
    ("D()+=L()*R()", [], [], []),
    ("D(a)+=L+(a)*R()", [777], [777], []),
    ("D(a)+=L(a)*R+()", [333], [333], []),
    ("D()+=L+()*R()", [], [], []),
    ("D(a,b)+=L(a)*R(b)", [300, 200], [300], [200]),
    ("D(a,b)+=L(b)*R+(a)", [130, 70], [70], [130]),
    ("D(a,b,c)+=L(c,a)*R(b)", [17, 19, 23], [23, 17], [19]),
    ("D()+=L(a,b)*R(b,a)", [], [40, 50], [50, 40]),
    ("D()+=L+(a)*R(a)", [], [3000], [3000]),
]


def _conj_flags(pat):
    lhs, rhs = pat.split("+=")
    lp, rp = rhs.split("*")
    return "+" in lp, "+" in rp


@pytest.mark.parametrize("dk", (R4, R8, C4, C8))
def test_rank0_operands_and_outer_products(talsh_gpu, oracle, dk):
    """D(a)+=L(a)*R(), D()+=L()*R(), outer products and dot products, conjugated where complex,
    accumulating and overwriting, against the oracle (reference kernels .cu:6732-6849)."""
    ts = talsh_gpu
    cplx = dk in (C4, C8)
    for t, (pat, dd, ld, rd) in enumerate(RANK0):
        if not cplx and "+(" in pat:
            continue
        for acc in (ts.YEP, ts.NOPE):
            say("rank-0", dk, pat, dd, ld, rd, acc)
            L, R, D = tensor(dk, ld, 300 + t), tensor(dk, rd, 400 + t), tensor(dk, dd, 500 + t)
            lt, rt, dt = on_gpu(ts, dk, ld, L), on_gpu(ts, dk, rd, R), on_gpu(ts, dk, dd, D)
            alpha = (0.75, -0.5) if cplx else (-1.25, 0.0)
            ts.check(ts.contract(pat, dt, lt, rt, alpha=complex(*alpha), accumulative=acc), "contract")
            want = oracle.contract(dk, pat, dd, D, ld, L, rd, R, alpha, acc)
            got = snapshot(ts, dt)
            if dk in (R8, C8):
                assert relerr(got, want) <= 1e-12, (pat, acc)
            else:
                # FP32 arithmetic: elementwise against |D| + |alpha| * sum |l| |r| (a dot product
                # can cancel far below its terms, so a normwise bound would not hold)
                plain = pat.replace("+(", "(")
                mag = oracle.contract(R8, plain, dd, np.zeros(max(1, int(np.prod(dd)))), ld,
                                      np.abs(L).astype(np.float64), rd, np.abs(R).astype(np.float64))
                bound = abs(complex(*alpha)) * mag + (np.abs(D) if acc else 0.0)
                assert np.all(np.abs(got - want) <= 1e-6 * bound), (pat, acc)
            for x in (lt, rt, dt):
                x.destruct()


@pytest.mark.parametrize("dk", (R8, R4, C8))
def test_repeated_shapes_in_flight(talsh_gpu, oracle, dk):
    """The same contraction shapes submitted several times without waiting, each into its own D: the plan
    is reused, both operand transposes share one launch where they take the same kernel, and the D offset
    table of an in-place GEMM is kept on the device from a shape's second use on, read by the later tasks
    (on whichever stream each lands) once the task that wrote it has.  Shapes with unusual extents so the
    first use here is the first anywhere in the process.  Every result against the oracle."""
    ts = talsh_gpu
    shapes = [("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [17, 19, 18, 21], [18, 5, 19, 7], [21, 7, 17, 5]),
              ("D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)", [19, 17, 21, 18], [19, 6, 21, 9], [9, 17, 6, 18]),
              ("D(a,b,c)+=L(c,i,a)*R(i,b)", [23, 17, 19], [19, 33, 23], [33, 17]),
              ("D(a,b,c,d)+=L(d,b,i)*R(a,i,c)", [17, 18, 19, 20], [20, 18, 11], [17, 11, 19])]
    copies = 4
    tol = 1e-12 if dk in (R8, C8) else 1e-5
    for s, (pat, dd, ld, rd) in enumerate(shapes):
        say("repeated shape", dk, pat, dd, ld, rd, "x", copies)
        L, R = tensor(dk, ld, 700 + s), tensor(dk, rd, 710 + s)
        lt, rt = on_gpu(ts, dk, ld, L), on_gpu(ts, dk, rd, R)
        ds, d0s, tasks = [], [], []
        for c in range(copies):
            D0 = tensor(dk, dd, 720 + 10 * s + c)
            ds.append(on_gpu(ts, dk, dd, D0))
            d0s.append(D0)
        alpha = complex(0.5, -0.25) if dk == C8 else 0.5
        for c in range(copies):
            tk = ts.Task()
            ts.check(ts.contract(pat, ds[c], lt, rt, alpha=alpha, task=tk), "contract %d" % c)
            tasks.append(tk)
        finish(*tasks)
        for c in range(copies):
            want = oracle.contract(dk, pat, dd, d0s[c], ld, L, rd, R, (alpha.real, alpha.imag) if dk == C8
                                   else (alpha, 0.0), 1)
            assert relerr(snapshot(ts, ds[c]), want) <= tol, (pat, c)
        for t in ds + [lt, rt]:
            t.destruct()


@pytest.mark.parametrize("dk", (R8, R4, C8))
def test_norm2_of_views_at_every_alignment(talsh_gpu, oracle, dk):
    """norm2 of views that start 0..3 elements into a GPU body (so 16-byte vector loads are and are not
    possible) and end on odd lengths, against the oracle on the same elements."""
    ts = talsh_gpu
    n = 100003
    data = tensor(dk, [n], 77)
    base = on_gpu(ts, dk, [n], data)
    addr = gpu_address(ts, base)
    esz = {R8: 8, R4: 4, C8: 16}[dk]
    tol = 1e-12 if dk in (R8, C8) else 1e-6
    for start in range(4):
        for length in (1, 5, 4099, n - start):
            say("norm2 view", dk, start, length)
            v = view(ts, dk, [length], addr + start * esz)
            d = ts.Tensor(dk, [], ts.gpu_flat(0))
            ts.check(ts.norm2(d, v), "norm2")
            got = d.scalar().real
            want = oracle.norm2(dk, data[start:start + length])
            assert abs(got - want) <= tol * want, (start, length, got, want)
            d.destruct()
            v.destruct()
    base.destruct()


@pytest.mark.parametrize("dk", (R8, R4, C4))
def test_permute_misaligned_views(talsh_gpu, oracle, dk):
    """Permuted copies and adds between views that start 1 and 3 elements into GPU bodies (so the 16-byte
    vector paths are not taken), random rank-4 permutations with odd and even extents: bit-exact against the
    oracle."""
    import random
    ts = talsh_gpu
    rng = random.Random(606 + dk)
    esz = {R8: 8, R4: 4, C4: 8}[dk]
    cap = 9 ** 4 + 8
    lbase = on_gpu(ts, dk, [cap], tensor(dk, [cap], 61))
    dbase = on_gpu(ts, dk, [cap], tensor(dk, [cap], 62))
    la, da = gpu_address(ts, lbase), gpu_address(ts, dbase)
    labels_ = list("abcd")
    for t in range(12):
        ld = [rng.randint(2, 9) for _ in range(4)]
        perm = labels_[:]
        rng.shuffle(perm)
        o2n = [labels_.index(x) + 1 for x in perm]  # L dim k (label perm[k]) goes to D position o2n[k]
        dd = [0] * 4
        for k in range(4):
            dd[o2n[k] - 1] = ld[k]
        pat_copy = "D(a,b,c,d)=L(%s)" % ",".join(perm)
        pat_add = "D(a,b,c,d)+=L(%s)" % ",".join(perm)
        vol = int(np.prod(ld))
        L = snapshot(ts, lbase)[1:1 + vol]
        lv = view(ts, dk, ld, la + 1 * esz)
        dv = view(ts, dk, dd, da + 3 * esz)
        say("misaligned", dk, pat_copy, ld)
        ts.check(ts.copy(pat_copy, dv, lv), "copy")
        got = snapshot(ts, dv)
        assert got.tobytes() == oracle.copy(dk, ld, o2n, L).tobytes(), (pat_copy, ld)
        alpha = complex(0.5, -0.25) if dk == C4 else 0.5
        ts.check(ts.add(pat_add, dv, lv, alpha=alpha), "add")
        want = oracle.add(dk, ld, o2n, L, got, (alpha.real, alpha.imag) if dk == C4 else (alpha, 0.0))
        assert snapshot(ts, dv).tobytes() == want.tobytes(), (pat_add, ld)
        lv.destruct()
        dv.destruct()
    lbase.destruct()
    dbase.destruct()


@pytest.mark.parametrize("dk,n", ((R8, 12), (R4, 12), (C8, 10)))
def test_permute_repeated_shapes(talsh_gpu, oracle, dk, n):
    """The first launches of a permutation shape of 8 MB or more time its kernel variants (tile dimensions,
    block order) and later ones run the fastest: five copies of each rank-6 permutation (12-16 MB) into a
    destination refilled with a sentinel before each, then four accumulating adds, every result bit-exact
    against the oracle."""
    import random
    ts = talsh_gpu
    rng = random.Random(707 + dk)
    labels_ = list("abcdef")
    ld = [n] * 6
    L = tensor(dk, ld, 71)
    lt = on_gpu(ts, dk, ld, L)
    dt = on_gpu(ts, dk, ld, tensor(dk, ld, 72))
    perms = [list("fcbead"), list("bafedc"), list("dbfcae")]
    for _ in range(2):
        p = labels_[:]
        rng.shuffle(p)
        perms.append(p)
    for perm in perms:
        o2n = [labels_.index(x) + 1 for x in perm]
        pat_copy = "D(a,b,c,d,e,f)=L(%s)" % ",".join(perm)
        want = oracle.copy(dk, ld, o2n, L)
        for call in range(5):
            ts.check(ts.init_value(dt, -7.0), "init")
            say("repeated", dk, pat_copy, ld, "call", call)
            ts.check(ts.copy(pat_copy, dt, lt), "copy")
            assert snapshot(ts, dt).tobytes() == want.tobytes(), (pat_copy, call)
        got = want
        pat_add = "D(a,b,c,d,e,f)+=L(%s)" % ",".join(perm)
        alpha = complex(0.5, -0.25) if dk == C8 else 0.75
        for call in range(4):
            ts.check(ts.add(pat_add, dt, lt, alpha=alpha), "add")
            got = oracle.add(dk, ld, o2n, L, got, (alpha.real, alpha.imag) if dk == C8 else (alpha, 0.0))
            assert snapshot(ts, dt).tobytes() == got.tobytes(), (pat_add, call)
    lt.destruct()
    dt.destruct()


@pytest.mark.parametrize("dk", (R8, R4, C8))
def test_contract_misaligned_views(talsh_gpu, oracle, dk):
    """Contractions whose L, R and D are views 1, 2 and 3 elements into GPU bodies (no 16-byte vector
    accesses anywhere), the configs' interleaved patterns and random ones, against the oracle."""
    import random
    from tal_common import random_contraction
    ts = talsh_gpu
    rng = random.Random(707 + dk)
    esz = {R8: 8, R4: 4, C8: 16}[dk]
    cap = 40 ** 4 + 8
    bases = [on_gpu(ts, dk, [cap], tensor(dk, [cap], 81 + i)) for i in range(3)]
    addrs = [gpu_address(ts, b) for b in bases]
    cases = [("D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)", [20, 18, 22, 17], [20, 9, 22, 7], [7, 18, 9, 17]),
             ("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [17, 19, 18, 21], [18, 5, 19, 7], [21, 7, 17, 5])]
    for _ in range(6):
        cases.append(random_contraction(rng, max_rank=4, ext_lo=2, ext_hi=17))
    tol = 1e-12 if dk in (R8, C8) else 1e-5
    for pat, dd, ld, rd in cases:
        say("misaligned contract", dk, pat, dd, ld, rd)
        host = [snapshot(ts, b) for b in bases]
        L = host[0][1:1 + int(np.prod(ld))]
        R = host[1][2:2 + int(np.prod(rd))]
        D0 = host[2][3:3 + int(np.prod(dd))]
        lv, rv = view(ts, dk, ld, addrs[0] + esz), view(ts, dk, rd, addrs[1] + 2 * esz)
        dv = view(ts, dk, dd, addrs[2] + 3 * esz)
        alpha = complex(0.75, 0.5) if dk == C8 else 0.75
        ts.check(ts.contract(pat, dv, lv, rv, alpha=alpha), "contract")
        want = oracle.contract(dk, pat, dd, D0, ld, L, rd, R, (alpha.real, alpha.imag) if dk == C8 else
                               (alpha, 0.0), 1)
        assert relerr(snapshot(ts, dv), want) <= tol, (pat, dd)
        for v in (lv, rv, dv):
            v.destruct()
    for b in bases:
        b.destruct()


@pytest.mark.parametrize("dk,c", ((R8, 24), (R8, 32), (R8, 7), (R8, 72), (R4, 24), (R4, 9), (R4, 96)))
def test_contract_skinny(talsh_gpu, oracle, dk, c):
    """Real contractions whose GEMM has at most 96 columns and more than 65,536 rows (the narrow-tile path, as
    the reference list's N = 24 and N = 72 shapes take): D overwritten and then accumulated into, with D written in place
    and through the plain layout, against the oracle (R8: normwise relative error <= 1e-12; R4: elementwise
    within 1e-6 * (|D| + |alpha| * sum |l| |r|))."""
    ts = talsh_gpu
    a, b, d = 256, 300, 40
    cases = [("D(a,b,c)+=L(b,d,a)*R(d,c)", [a, b, c], [b, d, a], [d, c]),
             ("D(a,b,c)+=L(a,b,d)*R(d,c)", [a, b, c], [a, b, d], [d, c]),
             ("D(a,c,b)+=L(d,a,b)*R(c,d)", [a, c, b], [d, a, b], [c, d])]

    def check(got, want, prev, alpha, pat):
        if dk == R8:
            assert relerr(got, want) <= 1e-12, (pat, c)
        else:
            plain = pat.replace("+(", "(")
            mag = oracle.contract(R8, plain, dd, np.zeros(int(np.prod(dd))), ld, np.abs(L).astype(np.float64), rd,
                                  np.abs(R).astype(np.float64))
            assert np.all(np.abs(got - want) <= 1e-6 * (abs(alpha) * mag + np.abs(prev))), (pat, c)

    for k, (pat, dd, ld, rd) in enumerate(cases):
        L, R, D0 = tensor(dk, ld, 81 + k), tensor(dk, rd, 82 + k), tensor(dk, dd, 83 + k)
        lt, rt, dt = on_gpu(ts, dk, ld, L), on_gpu(ts, dk, rd, R), on_gpu(ts, dk, dd, D0)
        say("skinny", dk, pat, dd, ld, rd)
        ts.check(ts.contract(pat, dt, lt, rt, alpha=0.5, accumulative=ts.NOPE), "overwrite")
        zero = np.zeros_like(D0)
        want = oracle.contract(dk, pat, dd, zero, ld, L, rd, R, (0.5, 0.0), 0)
        check(snapshot(ts, dt), want, zero, 0.5, pat)
        ts.check(ts.contract(pat, dt, lt, rt, alpha=-1.5), "accumulate")
        prev = want
        want = oracle.contract(dk, pat, dd, prev, ld, L, rd, R, (-1.5, 0.0), 1)
        check(snapshot(ts, dt), want, prev, -1.5, pat)
        for t in (lt, rt, dt):
            t.destruct()


@pytest.mark.parametrize("dk", (R8, C4))
def test_contract_high_rank(talsh_gpu, oracle, dk):
    """Contractions of rank-6 to rank-9 operands with extents 2-3 (after fusion up to eight outer dimensions
    per transpose, past what the paired transpose launch carries, so both the paired and the separate launches
    run), against the oracle."""
    import random
    ts = talsh_gpu
    rng = random.Random(808 + dk)
    tol = 1e-12 if dk == R8 else 1e-5
    for t in range(10):
        nl, nr, nc = rng.randint(2, 5), rng.randint(2, 5), rng.randint(2, 4)
        labels_ = [chr(ord("a") + i) for i in range(nl + nr + nc)]
        ext = {x: rng.randint(2, 3) for x in labels_}
        lu, ru, cc = labels_[:nl], labels_[nl:nl + nr], labels_[nl + nr:]
        d, l, r = lu + ru, lu + cc, ru + cc
        rng.shuffle(d)
        rng.shuffle(l)
        rng.shuffle(r)
        pat = "D(%s)+=L(%s)*R(%s)" % (",".join(d), ",".join(l), ",".join(r))
        dd, ld, rd = [ext[x] for x in d], [ext[x] for x in l], [ext[x] for x in r]
        say("high rank", dk, pat)
        L, R, D = tensor(dk, ld, 900 + t), tensor(dk, rd, 920 + t), tensor(dk, dd, 940 + t)
        lt, rt, dt = on_gpu(ts, dk, ld, L), on_gpu(ts, dk, rd, R), on_gpu(ts, dk, dd, D)
        ts.check(ts.contract(pat, dt, lt, rt, alpha=0.5), "contract")
        want = oracle.contract(dk, pat, dd, D, ld, L, rd, R, (0.5, 0.0), 1)
        assert relerr(snapshot(ts, dt), want) <= tol, pat
        for x in (lt, rt, dt):
            x.destruct()
