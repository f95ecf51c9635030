"""Benchmark of the TAL-SH contraction hot path on one MI355X.

Headline (`value`): R8 contraction TFLOP/s of the large dense contraction
    D(a,b,c,d) += L(a,i,c,j) * R(j,b,i,d), all extents 96
(84.9 M elements per tensor, TTGT GEMM M=N=K=9216, 1.5655e12 FLOP per step),
through talshTensorContract on GPU-resident tensors; one step = one
contraction (both input transposes in one launch + the FP64 MFMA GEMM
stage, which reads and writes D in place through row/column offsets).
Secondary figures on the same line: the same contraction in R4 and C8, the
32^4 contraction, and the rank-6 16^6 transpose / permuted-add / norm2 sweep
(34 permutations) as GB/s.  `roofline` is computed for the FP64 GEMM kernel
from its HIP-event duration on the task stream (talshTaskTime "mmul").
`cpu_baseline` times TAL-SH's own CPU path (oracle/_ref, built from the
reference sources) on a bounded sample of the same contraction.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

# The serving run overlaps requests on 16 streams, one hardware queue each;
# HIP reads this when it starts, so it is set before the library is loaded.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import tal_sh_amd as ts  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0

R4, R8, C4, C8 = ts.R4, ts.R8, ts.C4, ts.C8


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def uniform(seed, n):
    from tal_common import uniform as u
    return u(seed, n)


def gpu_tensor(dk, dims, seed):
    """GPU-resident tensor with seeded uniform[-1,1] data (uploaded once)."""
    n = int(np.prod(dims))
    if dk in (R4, R8):
        data = uniform(seed, n).astype(ts.NP_DTYPE[dk])
    else:
        data = (uniform(seed, n) + 1j * uniform(seed + 1, n)).astype(ts.NP_DTYPE[dk])
    t = ts.Tensor(dk, dims, ts.gpu_flat(0))
    h = ts.Tensor(dk, dims, 0, data=data)
    lab = ",".join("i%d" % k for k in range(len(dims)))
    ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), t, h), "upload")
    h.destruct()
    return t, data


def timed_contractions(pat, d, l, r, steps, warmup, alpha=1.0):
    """Wall time of `steps` contractions, each submitted and waited for before the next (task handles and
    arguments prepared outside the timed loop; the tasks' event times are read after it)."""
    lib = ts.load()
    for _ in range(warmup):
        ts.check(ts.contract(pat, d, l, r, alpha=alpha), "contract")
    lib.talshAmdDeviceSynchronize(0)
    tasks = [ts.Task() for _ in range(steps)]
    cpat, a, cc = pat.encode(), complex(alpha), ts.coh("M", "T", "T")
    dp, lp, rp = d.ptr, l.ptr, r.ptr
    t0 = time.perf_counter()
    for tk in tasks:
        ts.check(lib.talshTensorContract(cpat, dp, lp, rp, a.real, a.imag, ts.DEV_DEFAULT, ts.DEV_DEFAULT, cc,
                                         ts.YEP, tk.ptr), "contract")
        tk.wait()
    lib.talshAmdDeviceSynchronize(0)
    wall = time.perf_counter() - t0
    mm, tot = [], []
    for tk in tasks:
        tm = tk.times()
        mm.append(tm["mmul"])
        tot.append(tm["total"])
        tk.destruct()
    return wall, mm, tot


def pipelined_contractions(pat, dk, dims, l, r, steps, alpha=1.0):
    """Wall time of `steps` contractions submitted back to back as tasks and waited at the end, so that the
    host's launch work overlaps the GPU's.  Each step writes its own D (an output image is unavailable while a
    task writes it); L and R are shared."""
    lib = ts.load()
    outs = [gpu_tensor(dk, dims, 1000 + i)[0] for i in range(steps)]
    lib.talshAmdDeviceSynchronize(0)
    tasks = []
    t0 = time.perf_counter()
    for d in outs:
        tk = ts.Task()
        ts.check(ts.contract(pat, d, l, r, alpha=alpha, task=tk), "contract")
        tasks.append(tk)
    for tk in tasks:
        tk.wait()
    lib.talshAmdDeviceSynchronize(0)
    wall = time.perf_counter() - t0
    for tk in tasks:
        tk.destruct()
    for d in outs:
        d.destruct()
    return wall


def sampled_error(oracle, dk, pat, dims, D0, L, R, Dout, alpha, nsamples=64, seed=5):
    """Errors on sampled elements (each a K-term dot product in long double): normwise relative error
    ||D - D*|| / ||D*|| and max |D - D*| / max |D*| over the samples."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(D0.size, size=nsamples, replace=False)
    num = den = 0.0
    dmax = wmax = 0.0
    for off in idx:
        el = oracle.contract_element(dk, pat, dims, int(off), dims, L, dims, R)
        want = complex(D0[off]) + alpha * el
        diff = abs(complex(Dout[off]) - want)
        num += diff ** 2
        den += abs(want) ** 2
        dmax, wmax = max(dmax, diff), max(wmax, abs(want))
    return float(np.sqrt(num / den)), float(dmax / wmax)


def bench_contraction(oracle, dk, pat, n, steps, warmup, check=True, seed=100, pipelined=False):
    dims = [n] * 4
    log("contraction", dk, pat, dims)
    l, L = gpu_tensor(dk, dims, seed)
    r, R = gpu_tensor(dk, dims, seed + 2)
    d, D0 = gpu_tensor(dk, dims, seed + 4)
    wall, mm, tot = timed_contractions(pat, d, l, r, steps, warmup)
    flop = (8.0 if dk in (C4, C8) else 2.0) * float(n) ** 6
    res = {"tflops": flop * steps / wall / 1e12, "ms_per_step": 1e3 * wall / steps,
           "gemm_ms": 1e3 * float(np.mean(mm)), "gemm_tflops": flop / float(np.mean(mm)) / 1e12,
           "task_ms": 1e3 * float(np.mean(tot)), "flop_per_step": flop}
    reps = warmup + steps
    if pipelined:  # the same contraction submitted back to back into separate outputs
        pwall = pipelined_contractions(pat, dk, dims, l, r, steps)
        res["pipelined_tflops"] = flop * steps / pwall / 1e12
        res["pipelined_ms_per_step"] = 1e3 * pwall / steps
    if check:
        # D now holds D0 + reps * L*R: check against the oracle on samples
        out = d.host_array()
        res["rel_err_vs_fp64"], res["max_abs_err_ratio"] = sampled_error(oracle, dk, pat, dims, D0, L, R, out,
                                                                         float(reps))
    for t in (l, r, d):
        t.destruct()
    return res


def logged_call(fn):
    """Runs fn with the library's op log on; returns (result, log text)."""
    import tempfile
    lib = ts.load()
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tmp:
        os.dup2(tmp.fileno(), 1)
        lib.talshTensorOpLogStart()
        try:
            res = fn()
        finally:
            lib.talshTensorOpLogFinish()
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return res, tmp.read().decode(errors="replace")


def bench_xl(oracle, n=96, grant=2 << 30, reps=3, seed=300):
    """talshTensorContractXL of the config-3 contraction with Host-resident operands (pinned host buffer):
    with the default device-memory grant (everything fits: L crosses the link in 4 boxes that overlap the
    contraction, R once, whole) and with `grant` bytes (cut into pieces that fit).  Each call moves L, R
    and D to the GPU and D back, so the bound is the host link."""
    import re
    pat = "D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)"
    dims = [n] * 4
    size = int(np.prod(dims))
    L, R, D0 = uniform(seed, size), uniform(seed + 1, size), uniform(seed + 2, size)
    l, r, d = (ts.Tensor(R8, dims, 0, data=x) for x in (L, R, D0))
    flop = 2.0 * float(n) ** 6
    out = {"flop_per_call": flop, "operand_bytes": 8 * size}
    for label, g in (("default_grant", 0), ("grant_%dgb" % (grant >> 30), grant)):
        ts.set_xl_memory(g)
        log("xl", label, "grant", g)
        (err, text) = logged_call(lambda: ts.contract_xl(pat, d, l, r, alpha=1.0))
        ts.check(err, "contract_xl")
        m = re.findall(r"(\d+) pieces \((\d+) D boxes x (\d+) contracted tiles\)", text)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ts.check(ts.contract_xl(pat, d, l, r, alpha=1.0), "contract_xl")
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        res = {"ms": 1e3 * t, "tflops": flop / t / 1e12}
        if m:
            pc, nb, nk = (int(x) for x in m[-1])
            res.update({"pieces": pc, "d_boxes": nb, "k_tiles": nk})
        mv = re.findall(r"talshTensorContractXL done: \S+ flop, (\S+) bytes in, (\S+) bytes out", text)
        if mv:
            bi, bo = (float(x) for x in mv[-1])
            res.update({"link_bytes": bi + bo, "link_gbs": (bi + bo) / t / 1e9})
        out[label] = res
        log("xl", label, res)
    ts.set_xl_memory(0)
    # D = D0 + (2 + 2 * reps) * L*R after the calls above
    calls = 2 * (1 + reps)
    out["rel_err_vs_fp64"], out["max_abs_err_ratio"] = sampled_error(oracle, R8, pat, dims, D0, L, R,
                                                                     d.host_array(), float(calls))
    for x in (l, r, d):
        x.destruct()
    return out


def bench_hyper(oracle, steps=5, warmup=1, seed=400):
    """Hyper-contraction D(a,b,h) += L(a,k,h) * R(k,b,h) (h in D, L and R), R8, a = b = k = 1024, h = 16:
    16 independent 1024^3 products as one strided batched FP64 MFMA GEMM, GPU-resident operands."""
    pat = "D(a,b,h)+=L(a,k,h)*R(k,b,h)"
    n, nh = 1024, 16
    dims = [n, n, nh]
    l, L = gpu_tensor(R8, dims, seed)
    r, R = gpu_tensor(R8, dims, seed + 2)
    d, D0 = gpu_tensor(R8, dims, seed + 4)
    wall, mm, tot = timed_contractions(pat, d, l, r, steps, warmup)
    flop = 2.0 * n ** 3 * nh
    res = {"tflops": flop * steps / wall / 1e12, "ms_per_step": 1e3 * wall / steps,
           "gemm_ms": 1e3 * float(np.mean(mm)), "gemm_tflops": flop / float(np.mean(mm)) / 1e12,
           "gemm_frac_of_fp64_peak": flop / float(np.mean(mm)) / 1e12 / FP64_PEAK_TFLOPS, "flop_per_step": flop}
    out = d.host_array()
    reps = warmup + steps
    # sampled elements against the oracle's direct loop nest: D = D0 + reps * sum_k L * R
    rng = np.random.default_rng(9)
    idx = rng.choice(out.size, size=64, replace=False)
    num = den = 0.0
    for off in idx:
        el = oracle.contract_element(R8, pat, dims, int(off), dims, L, dims, R)
        want = D0[off] + reps * el.real
        num += (out[off] - want) ** 2
        den += want ** 2
    res["rel_err_vs_fp64"] = float(np.sqrt(num / den))
    for t in (l, r, d):
        t.destruct()
    return res


def bench_transposes(oracle, steps):
    """Config 2: rank-6 R8 16^6, identity + reversal + 32 seeded permutations; add; norm2; slice and insert of a quarter box."""
    import random

    dims = [16] * 6
    nbytes = 16 ** 6 * 8
    l, L = gpu_tensor(R8, dims, 2026)
    d = ts.Tensor(R8, dims, ts.gpu_flat(0))
    labels = list("abcdef")
    rng = random.Random(1234)
    perms = [labels[:], labels[::-1]]
    while len(perms) < 34:
        p = labels[:]
        rng.shuffle(p)
        if p not in perms:
            perms.append(p)
    lib = ts.load()

    def kernel_time(submit, n):
        """Shortest of n kernel spans (talshTaskTime "comput": dispatch timestamps of the operation's
        first and last kernel, so the host's launch latency is not counted)."""
        tms = []
        for _ in range(n):
            tk = ts.Task()
            ts.check(submit(tk), "task")
            tk.wait()
            tms.append(tk.times()["comput"])
            tk.destruct()
        return float(np.min(tms))

    per = []
    exact = True
    for p in perms:
        pat = "D(a,b,c,d,e,f)=L(%s)" % ",".join(p)
        ts.check(ts.copy(pat, d, l), "copy")  # warm-up
        per.append(kernel_time(lambda tk: ts.copy(pat, d, l, task=tk), steps))
    # bit-exact check of the last permutation
    o2n = [labels.index(x) + 1 for x in perms[-1]]
    exact = d.host_array().tobytes() == oracle.copy(R8, dims, o2n, L).tobytes()
    d.place(ts.gpu_flat(0))
    copy_gbs = [2 * nbytes / t / 1e9 for t in per]
    # permuted add D += 0.5 * perm(L)
    add_s = kernel_time(lambda tk: ts.add("D(a,b,c,d,e,f)+=L(c,a,f,b,e,d)", d, l, alpha=0.5, task=tk), steps + 1)
    add_gbs = 3 * nbytes / add_s / 1e9
    # norm2 (both stages: partial sums, then their fixed-order total)
    s = ts.Tensor(R8, [], ts.gpu_flat(0))
    norm_s = kernel_time(lambda tk: ts.norm2(s, l, copy_ctrl=ts.coh("M", "T"), task=tk), steps + 1)
    norm_gbs = nbytes / norm_s / 1e9
    # slice / insert of an 8 x 16^4 x 8 box (1/4 of the tensor, offsets 4 and 8 on the outer dims)
    sdims, off = [8, 16, 16, 16, 16, 8], [4, 0, 0, 0, 0, 8]
    sbytes = int(np.prod(sdims)) * 8
    sm = ts.Tensor(R8, sdims, ts.gpu_flat(0))
    slice_gbs = 2 * sbytes / kernel_time(lambda tk: ts.slice(sm, l, off, task=tk), steps + 1) / 1e9
    insert_gbs = 3 * sbytes / kernel_time(lambda tk: ts.insert(d, sm, off, accumulative=ts.YEP, task=tk),
                                          steps + 1) / 1e9
    for t in (l, d, s, sm):
        t.destruct()
    lib.talshAmdDeviceSynchronize(0)
    worst = int(np.argmin(copy_gbs))
    return {"timing": "kernel dispatch timestamps (talshTaskTime comput), shortest of %d runs per case" % steps,
            "copy_gbs_mean": float(np.mean(copy_gbs)), "copy_gbs_min": float(np.min(copy_gbs)),
            "copy_gbs_max": float(np.max(copy_gbs)), "copy_worst_permutation": ",".join(perms[worst]),
            "copy_gbs_identity": copy_gbs[0], "copy_gbs_reversal": copy_gbs[1],
            "copy_frac_hbm_mean": float(np.mean(copy_gbs)) / HBM_PEAK_GBS,
            "copy_frac_hbm_min": float(np.min(copy_gbs)) / HBM_PEAK_GBS, "copy_bit_exact": bool(exact),
            "add_gbs": add_gbs, "add_frac_hbm": add_gbs / HBM_PEAK_GBS, "norm2_gbs": norm_gbs,
            "norm2_frac_hbm": norm_gbs / HBM_PEAK_GBS, "permutations": len(perms),
            "slice_gbs": slice_gbs, "insert_accumulate_gbs": insert_gbs,
            "slice_frac_hbm": slice_gbs / HBM_PEAK_GBS, "insert_accumulate_frac_hbm": insert_gbs / HBM_PEAK_GBS,
            "per_permutation_gbs": [round(x, 1) for x in copy_gbs]}


def bench_reference_shapes(steps=3):
    """The 48 shapes of the reference's benchmark list (tensor_contractions.txt, driven by
    main.F90:957-1014: D = 0, L = 0.01, R = 0.001, one contraction per shape), R8, operands resident on
    the GPU, overwriting D.  Shapes come from the committed fixture tests/golden/patterns.json.  Each
    shape: the shortest of `steps` runs by talshTaskTime (compute: transposes + GEMM; mmul: GEMM), and
    every element of D checked against the closed form K * 1e-5 through its squared norm."""
    with open(os.path.join(REPO, "tests", "golden", "patterns.json")) as f:
        shapes = json.load(f)["tensor_contractions_txt"]
    gpu = ts.gpu_flat(0)
    rows, worst_err = [], 0.0
    tot_flop = tot_comp = tot_mm = 0.0
    for sh in shapes:
        dd, ld, rd = sh["ddims"], sh["ldims"], sh["rdims"]
        vd, vl, vr = (float(np.prod(x)) for x in (dd, ld, rd))
        k = float(np.sqrt(vl * vr / vd))
        flop = 2.0 * vd * k
        d = ts.Tensor(R8, dd, gpu, init_val=0.0)
        l = ts.Tensor(R8, ld, gpu, init_val=0.01)
        r = ts.Tensor(R8, rd, gpu, init_val=0.001)
        ts.check(ts.contract(sh["pattern"], d, l, r, accumulative=ts.NOPE), "contract")  # warm-up
        comp, mm = [], []
        for _ in range(steps):
            tk = ts.Task()
            ts.check(ts.contract(sh["pattern"], d, l, r, accumulative=ts.NOPE, task=tk), "contract")
            tk.wait()
            tm = tk.times()
            comp.append(tm["comput"])
            mm.append(tm["mmul"])
            tk.destruct()
        s2 = ts.Tensor(R8, [], gpu)
        ts.check(ts.norm2(s2, d, copy_ctrl=ts.coh("M", "T")), "norm2")
        want = vd * (k * 1e-5) ** 2
        err = abs(s2.scalar().real - want) / want
        worst_err = max(worst_err, err)
        for t in (d, l, r, s2):
            t.destruct()
        c, m = float(np.min(comp)), float(np.min(mm))
        tot_flop += flop
        tot_comp += c
        tot_mm += m
        rows.append({"pattern": sh["pattern"], "gflop": flop / 1e9, "tflops": flop / c / 1e12,
                     "gemm_tflops": flop / m / 1e12})
    tf = [x["tflops"] for x in rows]
    return {"shapes": len(rows), "tflop_total": tot_flop / 1e12, "tflops_aggregate": tot_flop / tot_comp / 1e12,
            "gemm_tflops_aggregate": tot_flop / tot_mm / 1e12,
            "frac_fp64_peak_aggregate": tot_flop / tot_comp / 1e12 / FP64_PEAK_TFLOPS,
            "tflops_min": float(np.min(tf)), "tflops_median": float(np.median(tf)), "tflops_max": float(np.max(tf)),
            "max_rel_err_norm2_vs_closed_form": worst_err,
            "per_shape": [{k2: (round(v, 3) if isinstance(v, float) else v) for k2, v in x.items()} for x in rows]}


def batch_contraction(rng):
    from tal_common import batch_contraction as gen
    return gen(rng)


def bench_batch(oracle, dk, ntask=4096, reps=5, check_sample=128, pertask=True):
    """Config 4: 4096 seeded small contractions; one persistent launch (graph replay)
    versus one talshTensorContract per task (each its own transpose+GEMM launches;
    pertask=False skips that comparison, e.g. under a counter profiler)."""
    import random

    rng = random.Random(4096 + dk)
    tasks = []
    for t in range(ntask):
        pat, dd, ld, rd = batch_contraction(rng)
        lt, L = gpu_tensor(dk, ld, 7 * t + 1)
        rt, R = gpu_tensor(dk, rd, 7 * t + 3)
        dt = ts.Tensor(dk, dd, ts.gpu_flat(0))
        tasks.append((pat, dd, ld, rd, L, R, lt, rt, dt))
    lib = ts.load()
    b = ts.Batch()
    for (pat, dd, ld, rd, L, R, lt, rt, dt) in tasks:
        ts.check(b.add(pat, dt, lt, rt, alpha=1.0, accumulative=0), "batch add")
    t0 = time.perf_counter()
    tk = ts.Task()
    ts.check(b.execute(tk), "batch execute")  # plans, autotunes, captures
    tk.wait()
    tk.destruct()
    first_s = time.perf_counter() - t0
    st = b.stats()
    flop = st["flops"]
    lib.talshAmdDeviceSynchronize(0)
    t0 = time.perf_counter()
    gpu_ms = []
    for _ in range(reps):
        tk = ts.Task()
        ts.check(b.execute(tk), "batch execute")
        tk.wait()
        gpu_ms.append(1e3 * tk.times()["mmul"])
        tk.destruct()
    lib.talshAmdDeviceSynchronize(0)
    batched_s = (time.perf_counter() - t0) / reps
    # correctness on a sample of tasks (each checked against the oracle)
    srng = np.random.default_rng(9)
    worst = 0.0
    for i in srng.choice(ntask, size=min(check_sample, ntask), replace=False):
        pat, dd, ld, rd, L, R, lt, rt, dt = tasks[i]
        got = dt.host_array()
        want = oracle.contract(dk, pat, dd, np.zeros(int(np.prod(dd)) if dd else 1, dtype=ts.NP_DTYPE[dk]),
                               ld, L, rd, R, (1.0, 0.0), 0)
        den = np.linalg.norm(want.astype(np.float64)) or 1.0
        worst = max(worst, float(np.linalg.norm((got - want).astype(np.float64)) / den))
        dt.place(ts.gpu_flat(0))
    # one launch sequence per task: talshTensorContract, all queued, then waited
    single_s = float("nan")
    if pertask:
        queued = []
        t0 = time.perf_counter()
        for (pat, dd, ld, rd, L, R, lt, rt, dt) in tasks:
            tk = ts.Task()
            ts.check(ts.contract(pat, dt, lt, rt, alpha=1.0, accumulative=0, task=tk), "contract")
            queued.append(tk)
        for tk in queued:
            tk.wait()
            tk.destruct()
        lib.talshAmdDeviceSynchronize(0)
        single_s = time.perf_counter() - t0
    b.destroy()
    for (pat, dd, ld, rd, L, R, lt, rt, dt) in tasks:
        for t in (lt, rt, dt):
            t.destruct()
    # per-task roofline: max(FLOP / MFMA peak, bytes / HBM peak), bytes = L + R read + D written
    # (non-accumulative); K = sqrt(vol L * vol R / vol D) for a contraction without hyper-indices
    esz, peak = (8, FP64_PEAK_TFLOPS) if dk == R8 else (4, FP32_PEAK_TFLOPS)
    bound_s, nbytes = 0.0, 0.0
    for (pat, dd, ld, rd, *_rest) in tasks:
        vd, vl, vr = (float(np.prod(x)) if x else 1.0 for x in (dd, ld, rd))
        f = 2.0 * vd * np.sqrt(vl * vr / vd)
        by = esz * (vl + vr + vd)
        nbytes += by
        bound_s += max(f / (peak * 1e12), by / (HBM_PEAK_GBS * 1e9))
    gpu_s = float(np.median(gpu_ms)) * 1e-3
    return {"tasks": ntask, "flop": flop, "bytes": nbytes, "batched_tflops": flop / batched_s / 1e12,
            "batched_gpu_tflops": flop / gpu_s / 1e12, "batched_gpu_gbs": nbytes / gpu_s / 1e9,
            "roofline_bound_ms": 1e3 * bound_s, "roofline_frac": bound_s / gpu_s,
            "batched_ms": 1e3 * batched_s, "batched_gpu_ms": float(np.median(gpu_ms)),
            "per_task_tflops": flop / single_s / 1e12, "per_task_ms": 1e3 * single_s,
            "speedup_vs_per_task": single_s / batched_s, "first_execution_s": first_s,
            "tuned_classes": st["tuned_classes"], "tune_ms": st["tune_ms"], "work_items": st["work_items"],
            "max_rel_err_sampled": worst}


SERVE_MIX = ((R8, 0.6), (R4, 0.3), (C8, 0.1))
SERVE_ELEM = {R4: 4, R8: 8, C8: 16}


class PinnedBuffer:
    """Pinned host memory taken from the library's host arena (body of a rank-2 R8 Host tensor)."""

    def __init__(self, nbytes):
        cols = 1 << 20
        rows = max(1, (nbytes + 8 * cols - 1) // (8 * cols))
        self.t = ts.Tensor(R8, [cols, rows], 0, noinit=True)
        self.addr = self.t.host_address()
        self.nbytes = 8 * cols * rows

    def view(self, dtype, offset_bytes, count):
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(self.addr + offset_bytes)
        return np.frombuffer(buf, dtype=dtype)

    def fill_uniform(self, dtype, seed):
        """uniform[-1,1) values of `dtype` over the whole buffer, written in place."""
        rng = np.random.default_rng(seed)
        n = self.nbytes // np.dtype(dtype).itemsize
        step = 1 << 26
        for s in range(0, n, step):
            v = self.view(dtype, s * np.dtype(dtype).itemsize, min(step, n - s))
            rng.random(dtype=dtype, out=v)
            v *= 2
            v -= 1


class OutputRing:
    """Output bodies for requests in flight, carved from one pinned region (first fit from a moving head)."""

    def __init__(self, nbytes):
        self.size, self.head, self.live = nbytes, 0, {}

    def alloc(self, n):
        n = (n + 255) // 256 * 256
        p = self.head if self.head + n <= self.size else 0
        for _ in range(len(self.live) + 2):
            hit = [o + m for o, m in self.live.items() if p < o + m and o < p + n]
            if not hit:
                self.live[p] = n
                self.head = p + n
                return p
            p = max(hit)
            if p + n > self.size:
                p = 0
        return None

    def free(self, p):
        self.live.pop(p, None)


def serve_stream(ntask, seed=5, ntemplates=48, reuse=0.3, vmin=1e3, vmax=1e8):
    """Seeded request stream (the first n requests are the same for every n).

    Each request: (pattern, ddims, ldims, rdims, dk, flop, template index or -1, L offset, R offset);
    70 % bring both operands from host memory, 30 % contract a resident template L with a new R."""
    import random
    from tal_common import serving_request

    rng = random.Random(seed)
    kinds, weights = [k for k, _ in SERVE_MIX], [w for _, w in SERVE_MIX]

    def fresh():
        pat, dd, ld, rd, flop = serving_request(rng, vmin, vmax)
        dk = rng.choices(kinds, weights)[0]
        return pat, dd, ld, rd, dk, flop * (4.0 if dk == C8 else 1.0)

    templates = [fresh() for _ in range(ntemplates)]
    out = []
    for _ in range(ntask):
        if rng.random() < reuse:
            ti = rng.randrange(ntemplates)
            pat, dd, ld, rd, dk, flop = templates[ti]
        else:
            ti = -1
            pat, dd, ld, rd, dk, flop = fresh()
        out.append((pat, dd, ld, rd, dk, flop, ti, rng.random(), rng.random()))
    return templates, out


def bench_serve(oracle, ntask=4000, rate=None, budget=200 * 10 ** 9, max_inflight=32, ring_bytes=24 << 30,
                check_max=24):
    """Config 5: Poisson stream of contraction requests through talshAmdServer.

    Operands come from pinned host memory (views into two seeded input buffers:
    float64 for R8/C8, float32 for R4) or, for reuse requests, are resident
    template tensors on the GPU; results return to pinned host memory (COPY_TTT).
    Three runs over the same stream: one request at a time (submit, wait), a
    closed loop keeping 2*max_inflight requests outstanding (capacity), and
    Poisson arrivals at `rate` requests/s (default 0.8 x measured capacity).
    Latency = completion observed - scheduled arrival."""
    import random

    templates, stream = serve_stream(ntask)
    lib = ts.load()
    free0 = lib.talshDeviceBufferFreeSize(0, ts.DEV_HOST)
    in64 = PinnedBuffer(5 << 29)  # largest operand: 1.3e8 C8 elements = 2.08 GB
    in32 = PinnedBuffer(5 << 27)  # largest R4 operand: 520 MB
    ring = PinnedBuffer(ring_bytes)
    if free0 - lib.talshDeviceBufferFreeSize(0, ts.DEV_HOST) < in64.nbytes + in32.nbytes + ring.nbytes:
        raise RuntimeError("host buffers did not come from the pinned host arena (talshInit host_buf_size)")
    in64.fill_uniform(np.float64, 1)
    in32.fill_uniform(np.float32, 2)

    def operand_view(dk, dims, frac):
        vol = int(np.prod(dims))
        buf, dt = (in32, np.float32) if dk == R4 else (in64, np.float64)
        n = vol * (2 if dk == C8 else 1)
        slots = (buf.nbytes // np.dtype(dt).itemsize - n) // 64
        assert slots >= 0, (dk, dims)
        off = 64 * int(frac * slots) * np.dtype(dt).itemsize
        return buf.addr + off, (buf, dt, off, n)

    def host_data(dk, info):
        buf, dt, off, n = info
        a = buf.view(dt, off, n).copy()
        return a.view(np.complex128) if dk == C8 else a

    # resident templates (left operands on the GPU)
    tpl = []
    for i, (pat, dd, ld, rd, dk, flop) in enumerate(templates):
        addr, info = operand_view(dk, ld, (i + 0.5) / len(templates))
        h = ts.Tensor(dk, ld, 0, ext=addr)
        g = ts.Tensor(dk, ld, ts.gpu_flat(0))
        lab = ",".join("i%d" % k for k in range(len(ld)))
        ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), g, h), "template upload")
        h.destruct()
        tpl.append((g, info))
    resident_bytes = sum(int(np.prod(t[2])) * SERVE_ELEM[t[4]] for t in templates)
    ts.load().talshAmdDeviceSynchronize(0)

    def make(i):
        pat, dd, ld, rd, dk, flop, ti, fl, fr = stream[i]
        if ti >= 0:
            lt, linfo, lbytes = tpl[ti][0], tpl[ti][1], 0
        else:
            addr, linfo = operand_view(dk, ld, fl)
            lt, lbytes = ts.Tensor(dk, ld, 0, ext=addr), int(np.prod(ld)) * SERVE_ELEM[dk]
        addr, rinfo = operand_view(dk, rd, fr)
        rt = ts.Tensor(dk, rd, 0, ext=addr)
        dbytes = int(np.prod(dd)) * SERVE_ELEM[dk]
        return pat, dk, lt, rt, linfo, rinfo, dbytes, lbytes + int(np.prod(rd)) * SERVE_ELEM[dk] + dbytes

    want_check = set(i for i in range(ntask) if stream[i][5] <= 3e8)
    want_check = set(sorted(want_check, key=lambda i: (i * 2654435761) % 1000003)[:check_max])
    checked = {}

    def run(n, arrivals=None, outstanding=None, collect=False):
        rg = OutputRing(ring.nbytes)
        srv = ts.Server(gpu=0, hbm_budget=budget, max_inflight=max_inflight)
        live = {}
        lat, flop, pcie = [], 0.0, 0
        i = 0
        t0 = time.perf_counter()
        while i < n or live:
            now = time.perf_counter() - t0
            due = i < n and (arrivals[i] <= now if arrivals is not None else len(live) < outstanding)
            off = rg.alloc(int(np.prod(stream[i][1])) * SERVE_ELEM[stream[i][4]]) if due else None
            if off is not None:  # (a full output ring holds the request back: its wait counts as latency)
                pat, dk, lt, rt, linfo, rinfo, dbytes, moved = make(i)
                dt = ts.Tensor(dk, stream[i][1], 0, ext=ring.addr + off)
                sched = arrivals[i] if arrivals is not None else now
                ts.check(srv.submit(i, pat, dt, lt, rt, alpha=1.0, copy_ctrl=ts.coh("T", "T", "T"),
                                    accumulative=0), "server submit")
                live[i] = (sched, dt, lt, rt, off, linfo, rinfo, moved)
                i += 1
                continue
            for tag, _, st in srv.progress():
                if st != ts.TALSH_TASK_COMPLETED:
                    raise RuntimeError("request %d failed with status %d" % (tag, st))
                done = time.perf_counter() - t0
                sched, dt, lt, rt, off, linfo, rinfo, moved = live.pop(tag)
                lat.append(done - sched)
                flop += stream[tag][5]
                pcie += moved
                if collect and tag in want_check:
                    checked[tag] = (dt.host_array(), linfo, rinfo)
                dt.destruct()
                rt.destruct()
                if stream[tag][6] < 0:
                    lt.destruct()
                rg.free(off)
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.destroy()
        lat = np.array(lat)
        return {"requests": n, "wall_s": wall, "requests_per_s": n / wall, "tflops": flop / wall / 1e12,
                "p50_latency_ms": 1e3 * float(np.percentile(lat, 50)),
                "p99_latency_ms": 1e3 * float(np.percentile(lat, 99)),
                "mean_latency_ms": 1e3 * float(lat.mean()), "host_link_gbs": pcie / wall / 1e9,
                "server_p99_queue_wait_ms": 1e3 * st["p99_queue_wait_s"],
                "peak_transient_gb": st["peak_transient_bytes"] / 1e9, "deferred": int(st["deferred"])}

    def serial(n):
        t0 = time.perf_counter()
        flop = 0.0
        for i in range(n):
            pat, dk, lt, rt, linfo, rinfo, dbytes, moved = make(i)
            dt = ts.Tensor(dk, stream[i][1], 0, ext=ring.addr)
            tk = ts.Task()
            ts.check(ts.contract(pat, dt, lt, rt, copy_ctrl=ts.coh("T", "T", "T"), accumulative=0, task=tk),
                     "contract")
            tk.wait()
            tk.destruct()
            flop += stream[i][5]
            for x in (dt, rt) + ((lt,) if stream[i][6] < 0 else ()):
                x.destruct()
        wall = time.perf_counter() - t0
        return {"requests": n, "wall_s": wall, "requests_per_s": n / wall, "tflops": flop / wall / 1e12}

    nser = min(ntask, 300)
    log("serve: one at a time", nser)
    one = serial(nser)
    ncap = min(ntask, 1500)
    log("serve: closed loop", ncap)
    cap = run(ncap, outstanding=2 * max_inflight)
    if rate is None:
        rate = 0.8 * cap["requests_per_s"]
    arng = random.Random(77)
    arrivals, t = [], 0.0
    for _ in range(ntask):
        t += arng.expovariate(rate)
        arrivals.append(t)
    log("serve: poisson", ntask, "at", rate, "req/s")
    pois = run(ntask, arrivals=arrivals, collect=True)
    errs = {}
    for tag, (got, linfo, rinfo) in checked.items():
        pat, dd, ld, rd, dk, flop, ti, fl, fr = stream[tag]
        L, R = host_data(dk, linfo), host_data(dk, rinfo)
        want = oracle.contract(dk, pat, dd, np.zeros(int(np.prod(dd)), dtype=ts.NP_DTYPE[dk]), ld, L, rd, R,
                               (1.0, 0.0), 0)
        den = np.linalg.norm(want.astype(np.complex128)) or 1.0
        e = float(np.linalg.norm((got - want).astype(np.complex128)) / den)
        key = {R4: "r4", R8: "r8", C8: "c8"}[dk]
        errs[key] = max(errs.get(key, 0.0), e)
    for g, _ in tpl:
        g.destruct()
    total_flop = sum(s[5] for s in stream)
    return {"requests": ntask, "rate_req_per_s": rate, "budget_gb": budget / 1e9, "max_inflight": max_inflight,
            "stream_tflop": total_flop / 1e12, "resident_template_gb": resident_bytes / 1e9,
            "poisson": pois, "closed_loop": cap, "one_at_a_time": one,
            "overlap_speedup": cap["requests_per_s"] / one["requests_per_s"],
            "checked_requests": len(checked), "max_rel_err_sampled": errs}


def cpu_baseline():
    """TAL-SH's own CPU path on the headline contraction at its full shape.

    One call of D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d), all extents 96 (M=N=K=9216, 1.5655e12 FLOP; about
    25 s on 16 cores), after an untimed call on a small shape that starts the OpenMP thread pool."""
    from tal_common import REF_SO, Oracle, Reference

    pat = "D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)"
    dims = [96] * 4
    n = 96 ** 4
    L, R, D = uniform(11, n), uniform(12, n), uniform(13, n)
    flop = 2.0 * 9216.0 ** 3
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if os.path.exists(REF_SO):
        ref = Reference()
        kind = "reference"
        run = lambda dd, a, b, c: ref.contract(R8, pat, dd, c, dd, a, dd, b)
    else:
        orc = Oracle()
        kind = "port"
        run = lambda dd, a, b, c: orc.contract(R8, pat, dd, c, dd, a, dd, b)
    small = uniform(14, 16 ** 4)
    run([16] * 4, small, small, small)
    t0 = time.perf_counter()
    run(dims, L, R, D)
    t = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((x.split(":", 1)[1].strip() for x in f if x.startswith("model name")), "")
    except OSError:
        pass
    return {"value": flop / t / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": kind,
            "nproc": os.cpu_count(), "cpu_model": model, "seconds": t,
            "sample": "the headline contraction at full size: D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d) R8, extents 96 "
                      "(M=N=K=9216, 1.5655e12 FLOP), one call of %.1f s after a 16^4 warm-up call, "
                      "OMP threads %d" % (t, cores)}


def gemm_io_lower_bound(m, n, k, esize, xcds=8, l2_bytes=4 << 20):
    """Bytes an m x n x k GEMM must move between L2 and memory with `xcds` L2 caches of l2_bytes: each XCD
    computes mnk / xcds multiply-adds, and a processor with S words of fast memory moves at least
    2 * (its multiply-adds) / sqrt(S) words (classical algorithm); plus C read and written once."""
    s = l2_bytes / esize
    return int(xcds * 2.0 * (m * n * k / xcds) / np.sqrt(s) * esize + 2 * m * n * esize)


def check_init(err):
    if err:
        raise RuntimeError("talshInit failed: %d" % err)


def profiled_traffic(path=os.path.join(REPO, "profiles", "contract_r8_96_traffic.md")):
    """HBM-side bytes per GEMM stage (the whole-tile launch, the K-split launch and its summing pass), from
    the committed counter summary (FETCH_SIZE and WRITE_SIZE passes scaled by a plain copy of known size);
    None when the summary has no line for it."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_gemm_stage"):
                return int(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--only", default="",
                    help="comma list of parts to run "
                         "(head,r4,bf16,fp16x3,c8,config1,transpose,shapes,batch,hyper,xl,serve)")
    ap.add_argument("--batch-no-pertask", action="store_true",
                    help="batched config 4 without the one-launch-per-task comparison (for counter profiles)")
    ap.add_argument("--serve-requests", type=int, default=20000, help="serving stream length")
    ap.add_argument("--serve-rate", type=float, default=None, help="arrivals per second (default 0.8 x capacity)")
    ap.add_argument("--serve-streams", type=int, default=16, help="streams per GPU for the serving run")
    ap.add_argument("--serve-inflight", type=int, default=32, help="requests in flight (server slots)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    from tal_common import Oracle

    oracle = Oracle()
    check_init(ts.init(host_buf_size=4 << 30, dev_buffer_size=48 << 30))
    pat3 = "D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)"
    only = set(x for x in args.only.split(",") if x)
    want = lambda part: (not args.quick and not only) or part in only
    head = bench_contraction(oracle, R8, pat3, 96, args.steps, args.warmup) if (not only or "head" in only) else {}
    log("headline", head)
    secondary = {}
    if want("r4"):
        secondary["r4_fp32_config3"] = bench_contraction(oracle, R4, pat3, 96, max(2, args.steps // 2), 1)
    for name, method in (("bf16", ts.SPLIT_BF16X6), ("fp16x3", ts.SPLIT_FP16X3)):
        if want(name):
            if ts.fast_math(True, method) == 0:
                key = "r4_split_bf16_config3" if name == "bf16" else "r4_split_fp16x3_config3"
                secondary[key] = bench_contraction(oracle, R4, pat3, 96, max(2, args.steps // 2), 1)
            ts.fast_math(False)
            ts.load().talshAmdSetFastMathSplit(ts.SPLIT_BF16X6)
    fp32 = secondary.get("r4_fp32_config3")
    for key in ("r4_split_bf16_config3", "r4_split_fp16x3_config3"):
        if fp32 and key in secondary:  # the fast-math target is stated against the plain FP32 kernel
            secondary[key]["speedup_vs_fp32_gemm"] = secondary[key]["gemm_tflops"] / fp32["gemm_tflops"]
            secondary[key]["speedup_vs_fp32_step"] = secondary[key]["tflops"] / fp32["tflops"]
    if want("c8"):
        secondary["c8_config3"] = bench_contraction(oracle, C8, pat3, 96, max(2, args.steps // 4), 1)
    if want("config1"):
        secondary["r8_config1"] = bench_contraction(oracle, R8, "D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", 32, 50, 5,
                                                    pipelined=True)
    if want("transpose"):
        secondary["transpose_config2"] = bench_transposes(oracle, 5)
    if want("shapes"):
        secondary["reference_shapes_r8"] = bench_reference_shapes()
    if want("batch"):
        secondary["batched_config4_r8"] = bench_batch(oracle, R8, pertask=not args.batch_no_pertask)
        secondary["batched_config4_r4"] = bench_batch(oracle, R4, pertask=not args.batch_no_pertask)
    if want("hyper"):
        secondary["hyper_r8_batched"] = bench_hyper(oracle)
    if want("xl"):
        secondary["xl_config3_host_operands"] = bench_xl(oracle)
    ts.shutdown()
    if want("serve"):
        # 200 GB admission budget: the device pool holds it plus the resident templates
        check_init(ts.init(host_buf_size=29 << 30, dev_buffer_size=224 << 30, nstreams=args.serve_streams))
        secondary["serving_config5"] = bench_serve(oracle, ntask=args.serve_requests, rate=args.serve_rate,
                                                     max_inflight=args.serve_inflight)
        ts.shutdown()
    log("secondary", secondary)
    if not head:
        print(json.dumps({"secondary": secondary}), flush=True)
        return
    cpu = None if args.no_cpu else cpu_baseline()
    gemm_s = head["gemm_ms"] * 1e-3
    achieved = head["flop_per_step"] / gemm_s / 1e12
    line = {
        "metric": "R8/R4 contraction TFLOP/s (%MFMA peak); transpose GB/s (%HBM); p99 task latency",
        "value": round(head["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform[-1,1])",
        "config": {"workload": "large dense contraction D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d), R8, extents 96 "
                               "(GEMM M=N=K=9216), GPU-resident operands, accumulative, alpha 1",
                   "parallelism": "single GPU"},
        "rel_err_vs_fp64": head.get("rel_err_vs_fp64"),
        "max_abs_err_ratio": head.get("max_abs_err_ratio"),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic(),
                     "traffic_source": "profiles/contract_r8_96_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated; bytes per GEMM stage)",
                     # classical-GEMM I/O lower bound with 4 MiB of L2 per XCD (8 XCDs), plus D read and written once
                     "traffic_io_lower_bound": gemm_io_lower_bound(9216, 9216, 9216, 8),
                     "kernel": "FP64 MFMA GEMM stage: whole-tile launch + K-split launch of the last round + "
                               "summing pass (2*M*N*K FLOP / HIP-event duration of the stage)"},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
