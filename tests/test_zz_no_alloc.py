"""Allocation-free norm2 (talshAmdTensorNorm2NoAlloc) against talshTensorNorm2.

The no-alloc entry point runs talshTensorNorm2's kernels with their partial
sums in a fixed per-GPU device array, so its result must equal the pooled
entry point's bit for bit.  Two calls follow each other on one GPU, the second
on a smaller input, so the second would read stale partials of the first if
the reduction read more of the array than its own launch wrote.  The pool's
free size must not change across no-alloc calls.
"""
import ctypes

import numpy as np
import pytest

from tal_common import C4, C8, R4, R8, tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def on_gpu(ts, dk, dims, data):
    t = ts.Tensor(dk, dims, ts.gpu_flat(0))
    h = ts.Tensor(dk, dims, 0, data=data)
    lab = ",".join("i%d" % k for k in range(len(dims)))
    ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), t, h), "upload")
    h.destruct()
    return t


def gpu_body(ts, t):
    body = ctypes.c_void_p()
    ts.check(ts.load().talshTensorGetBodyAccess(t.ptr, ctypes.byref(body), t.dk, 0, ts.DEV_NVIDIA_GPU),
             "talshTensorGetBodyAccess")
    return body.value


def scalar_bytes(dk, z):
    if dk in (R4, C4):
        return np.array([z.real], dtype=np.float32).tobytes()
    return np.array([z.real], dtype=np.float64).tobytes()


@pytest.mark.parametrize("dk", (R4, R8, C4, C8))
def test_norm2_noalloc_matches_pooled(talsh_gpu, oracle, dk):
    ts = talsh_gpu
    lib = ts.load()
    # first call large (every stage-1 block writes a partial), second smaller and ragged
    for step, dims in enumerate(([64, 64, 64, 64], [37, 21])):
        say("norm2 no-alloc", dk, dims, "call", step + 1)
        L = tensor(dk, dims, 31 + step)
        lt = on_gpu(ts, dk, dims, L)
        d_pool = ts.Tensor(dk, [], ts.gpu_flat(0))
        ts.check(ts.norm2(d_pool, lt), "talshTensorNorm2")
        want = d_pool.scalar()
        d_free = ts.Tensor(dk, [], ts.gpu_flat(0))
        free0 = lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU)
        err = lib.talshAmdTensorNorm2NoAlloc(dk, gpu_body(ts, lt), lt.volume(), gpu_body(ts, d_free), 0, None)
        assert err == 0, err
        ts.check(lib.talshAmdDeviceSynchronize(0), "sync")
        assert lib.talshDeviceBufferFreeSize(0, ts.DEV_NVIDIA_GPU) == free0
        got = d_free.scalar()
        assert scalar_bytes(dk, got) == scalar_bytes(dk, want), (dk, dims, got, want)
        ref = oracle.norm2(dk, L)
        tol = 1e-12 if dk in (R8, C8) else 1e-6
        assert abs(got.real - ref) <= tol * ref, (dk, dims, got, ref)
        for t in (lt, d_pool, d_free):
            t.destruct()


def test_norm2_noalloc_rejects_bad_args(talsh_gpu):
    ts = talsh_gpu
    lib = ts.load()
    t = on_gpu(ts, R8, [8], np.ones(8))
    d = ts.Tensor(R8, [], ts.gpu_flat(0))
    body, res = gpu_body(ts, t), gpu_body(ts, d)
    assert lib.talshAmdTensorNorm2NoAlloc(R8, body, 0, res, 0, None) == ts.TALSH_INVALID_ARGS
    assert lib.talshAmdTensorNorm2NoAlloc(7, body, 8, res, 0, None) == ts.TALSH_INVALID_ARGS
    assert lib.talshAmdTensorNorm2NoAlloc(R8, None, 8, res, 0, None) == ts.TALSH_INVALID_ARGS
    assert lib.talshAmdTensorNorm2NoAlloc(R8, body, 8, res, 5, None) == ts.TALSH_INVALID_ARGS
    assert lib.talshAmdTensorNorm2NoAlloc(R8, body, 8, res, 0, None) == 0
    ts.check(lib.talshAmdDeviceSynchronize(0), "sync")
    assert d.scalar().real == 8.0
    t.destruct()
    d.destruct()
