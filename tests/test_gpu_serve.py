"""Serving engine (talshAmdServer*): continuous admission of contraction requests.

Requests of mixed kinds and sizes are submitted back to back so that many are
in flight together; each result must be bit-identical to the same request run
alone (the kernels are deterministic per element), and a sample is checked
against the oracle.  Budget and slot limits are exercised with a tiny budget.
"""
import random

import numpy as np
import pytest

from tal_common import C8, R4, R8, serving_request, tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def relerr(got, want):
    den = np.linalg.norm(want.astype(np.complex128))
    return np.linalg.norm((got - want).astype(np.complex128)) / (den if den > 0 else 1.0)


def make_requests(ts, n, seed, vmax=3e5):
    rng = random.Random(seed)
    reqs = []
    resident = {}
    for i in range(n):
        pat, dd, ld, rd, flop = serving_request(rng, vmin=1e3, vmax=vmax)
        dk = rng.choices([R8, R4, C8], weights=[6, 3, 1])[0]
        L, R = tensor(dk, ld, 10 * i + 1), tensor(dk, rd, 10 * i + 2)
        reuse = rng.random() < 0.3
        if reuse:  # left operand resident on the GPU
            lt = ts.Tensor(dk, ld, ts.gpu_flat(0))
            h = ts.Tensor(dk, ld, 0, data=L)
            lab = ",".join("i%d" % k for k in range(len(ld)))
            ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), lt, h), "upload")
            h.destruct()
        else:
            lt = ts.Tensor(dk, ld, 0, data=L)
        rt = ts.Tensor(dk, rd, 0, data=R)
        dt = ts.Tensor(dk, dd, 0)
        reqs.append(dict(tag=i, pat=pat, dk=dk, dd=dd, ld=ld, rd=rd, L=L, R=R, lt=lt, rt=rt, dt=dt, reuse=reuse))
    return reqs


def test_results_same_alone_and_in_mix(talsh_gpu, oracle):
    ts = talsh_gpu
    reqs = make_requests(ts, 48, seed=11)
    srv = ts.Server(gpu=0, hbm_budget=2 << 30, max_inflight=12)
    say("serve mix", len(reqs))
    for q in reqs:
        assert srv.submit(q["tag"], q["pat"], q["dt"], q["lt"], q["rt"], alpha=0.5, accumulative=0) == 0
    done = {}
    max_inflight = 0
    while len(done) < len(reqs):
        for tag, lat, st in srv.progress():
            assert st == ts.TALSH_TASK_COMPLETED
            done[tag] = lat
        max_inflight = max(max_inflight, srv.pending()[1])
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_transient_bytes"] <= 2 << 30
    srv.destroy()
    mixed = {q["tag"]: q["dt"].host_array() for q in reqs}
    for q in reqs:
        say("alone", q["tag"], q["pat"], q["dk"])
        alone = ts.Tensor(q["dk"], q["dd"], 0)
        ts.check(ts.contract(q["pat"], alone, q["lt"], q["rt"], alpha=0.5, accumulative=0,
                             copy_ctrl=ts.coh("T", "T", "T")), "contract alone")
        assert alone.host_array().tobytes() == mixed[q["tag"]].tobytes(), q["pat"]
    for q in reqs[:12]:
        want = oracle.contract(q["dk"], q["pat"], q["dd"], np.zeros(int(np.prod(q["dd"])), dtype=q["L"].dtype),
                               q["ld"], q["L"], q["rd"], q["R"], (0.5, 0.0), 0)
        tol = 1e-5 if q["dk"] == R4 else 1e-12
        assert relerr(mixed[q["tag"]], want) <= tol, q["pat"]


def test_admission_respects_budget_and_slots(talsh_gpu):
    ts = talsh_gpu
    reqs = make_requests(ts, 16, seed=5, vmax=2e5)
    srv = ts.Server(gpu=0, hbm_budget=24 << 20, max_inflight=2)
    for q in reqs:
        err = srv.submit(q["tag"], q["pat"], q["dt"], q["lt"], q["rt"], accumulative=0)
        assert err in (0, ts.DEVICE_UNABLE)
        q["accepted"] = err == 0
        assert srv.pending()[1] <= 2
    n_acc = sum(q["accepted"] for q in reqs)
    got = 0
    while got < n_acc:
        got += len(srv.progress())
        assert srv.pending()[1] <= 2
    st = srv.stats()
    assert st["peak_transient_bytes"] <= 24 << 20 and st["completed"] == n_acc
    srv.destroy()
