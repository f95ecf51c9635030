"""Batched contractions (talshAmdBatch*): one persistent launch replayed as a hipGraph.

Every task of a batch is checked against the oracle; graph replays are
checked one by one with ordinary (eager) library work between them, and the
batched result must equal, to rounding, what talshTensorContract gives for
the same task alone.
"""
import random

import numpy as np
import pytest

from tal_common import R4, R8, batch_contraction, tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def relerr(got, want):
    den = np.linalg.norm(want.astype(np.float64))
    return np.linalg.norm((got - want).astype(np.float64)) / (den if den > 0 else 1.0)


def upload(ts, dk, dims, data):
    t = ts.Tensor(dk, dims, ts.gpu_flat(0), init_val=complex(data[0]) if not len(dims) else 0.0)
    if len(dims):
        h = ts.Tensor(dk, dims, 0, data=data)
        lab = ",".join("i%d" % k for k in range(len(dims)))
        ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), t, h), "upload")
        h.destruct()
    return t


@pytest.mark.parametrize("dk", (R8, R4))
def test_batch_matches_oracle_and_replays(talsh_gpu, oracle, dk):
    ts = talsh_gpu
    rng = random.Random(2024 + dk)
    ntask = 160
    tasks = []
    for t in range(ntask):
        pat, dd, ld, rd = batch_contraction(rng)
        L, R, D = tensor(dk, ld, 3 * t), tensor(dk, rd, 3 * t + 1), tensor(dk, dd, 3 * t + 2)
        acc = t % 2
        tasks.append((pat, dd, ld, rd, L, R, D, acc, upload(ts, dk, ld, L), upload(ts, dk, rd, R),
                      upload(ts, dk, dd, D)))
    b = ts.Batch()
    for (pat, dd, ld, rd, L, R, D, acc, lt, rt, dt) in tasks:
        assert b.add(pat, dt, lt, rt, alpha=0.5, accumulative=acc) == 0
    tol = 1e-12 if dk == R8 else 1e-5
    wants = [oracle.contract(dk, pat, dd, D, ld, L, rd, R, (0.5, 0.0), acc)
             for (pat, dd, ld, rd, L, R, D, acc, lt, rt, dt) in tasks]
    for rep in range(3):
        say("batch execute", dk, "replay", rep)
        tk = ts.Task()
        assert b.execute(tk) == 0
        assert tk.wait() == ts.TALSH_TASK_COMPLETED
        assert tk.t.flops > 0
        tk.destruct()
        for i, (pat, dd, ld, rd, L, R, D, acc, lt, rt, dt) in enumerate(tasks):
            got = dt.host_array()
            # accumulative tasks add alpha*L*R once per execution
            want = wants[i] + rep * (wants[i] - D) if acc else wants[i]
            assert relerr(got, want) <= tol, (rep, i, pat)
            # eager library work between replays: put the output back on the GPU
            dt.place(ts.gpu_flat(0))
    # same results as talshTensorContract on each task alone (non-accumulative ones)
    for (pat, dd, ld, rd, L, R, D, acc, lt, rt, dt) in tasks[:40]:
        if acc:
            continue
        alone = upload(ts, dk, dd, D)
        ts.check(ts.contract(pat, alone, lt, rt, alpha=0.5, accumulative=0), "contract")
        assert relerr(dt.host_array(), alone.host_array()) <= tol
        dt.place(ts.gpu_flat(0))
    st = b.stats()
    assert st["executions"] == 3 and st["work_items"] > 0
    b.destroy()


def test_batch_rejects_aliased_outputs(talsh_gpu):
    ts = talsh_gpu
    d = upload(ts, R8, [4, 4], np.zeros(16))
    l = upload(ts, R8, [4, 3], tensor(R8, [4, 3], 1))
    r = upload(ts, R8, [3, 4], tensor(R8, [3, 4], 2))
    b = ts.Batch()
    for _ in range(2):
        assert b.add("D(a,b)+=L(a,i)*R(i,b)", d, l, r) == 0
    assert b.execute() == ts.TALSH_INVALID_ARGS
    b.destroy()


def test_batch_refers_hyper_contractions_to_ttgt(talsh_gpu):
    """A hyper index (in D, L and R) is not a gather-GEMM task: the batch reports DEVICE_UNABLE and
    talshTensorContract, which batches hyper indices, takes it."""
    ts = talsh_gpu
    d = upload(ts, R8, [4, 3], np.zeros(12))
    l = upload(ts, R8, [4, 5, 3], tensor(R8, [4, 5, 3], 1))
    r = upload(ts, R8, [5, 3], tensor(R8, [5, 3], 2))
    b = ts.Batch()
    assert b.add("D(a,h)+=L(a,k,h)*R(k,h)", d, l, r) == 0
    assert b.execute() == ts.DEVICE_UNABLE
    b.destroy()


@pytest.mark.parametrize("dk", (R8, R4))
def test_batch_back_to_back_replays(talsh_gpu, oracle, dk):
    """Replays with nothing else touching the operands in between take the
    replay fast path (one wait on the batch's own event); accumulative tasks
    must then hold D0 + reps * alpha * L * R, and a Host read in between still
    sees every replay before it."""
    ts = talsh_gpu
    rng = random.Random(77 + dk)
    tasks = []
    for t in range(96):
        pat, dd, ld, rd = batch_contraction(rng)
        L, R, D = tensor(dk, ld, 5 * t), tensor(dk, rd, 5 * t + 1), tensor(dk, dd, 5 * t + 2)
        tasks.append((pat, dd, ld, rd, L, R, D, upload(ts, dk, ld, L), upload(ts, dk, rd, R), upload(ts, dk, dd, D)))
    b = ts.Batch()
    for (pat, dd, ld, rd, L, R, D, lt, rt, dt) in tasks:
        assert b.add(pat, dt, lt, rt, alpha=0.25, accumulative=1) == 0
    one = [oracle.contract(dk, pat, dd, np.zeros_like(D), ld, L, rd, R, (0.25, 0.0), 0)
           for (pat, dd, ld, rd, L, R, D, lt, rt, dt) in tasks]
    tol = 1e-12 if dk == R8 else 1e-5
    reps = 0
    for burst in (3, 2):
        say("batch back-to-back", dk, burst)
        for _ in range(burst):
            tk = ts.Task()
            assert b.execute(tk) == 0
            assert tk.wait() == ts.TALSH_TASK_COMPLETED
            tk.destruct()
            reps += 1
        for i, (pat, dd, ld, rd, L, R, D, lt, rt, dt) in enumerate(tasks):
            want = D + reps * one[i]
            assert relerr(dt.host_array(), want) <= tol, (reps, i, pat)
            dt.place(ts.gpu_flat(0))
    assert b.stats()["executions"] == reps
    b.destroy()
