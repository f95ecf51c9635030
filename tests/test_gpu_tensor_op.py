"""Tensor slice views and tensor operations (talsh.h:312-431, talshc.cpp:1855-1935, :2599-3317) on the GPU.

An operation on boxes of existing tensors must leave D[box] = D[box] + alpha * L[box] * R[box] (the result
box is accumulated into the destination, talshc.cpp:2830-2854) and every element outside the box untouched;
compared with the oracle's contraction of the same boxes: R8/C8 normwise relative error <= 1e-12.
Talsh_tens_op_t's stage machine, the two-way decomposition and the FLOP / byte counts are checked too.
"""
import ctypes

import numpy as np
import pytest

from tal_common import C8, R8, tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def lib_of(ts):
    lib = ts.load()
    P = ctypes.POINTER
    lib.talshTensorSliceConstruct.argtypes = [P(ts.TalshTensSlice), P(ts.TalshTens), P(ctypes.c_size_t),
                                              P(ctypes.c_int), ctypes.c_void_p, ctypes.c_void_p]
    lib.talshTensorSliceVolume.restype = ctypes.c_size_t
    lib.talshTensorSliceVolume.argtypes = [P(ts.TalshTensSlice)]
    lib.talshTensorOpSetArgument.argtypes = [P(ts.TalshTensOp), P(ts.TalshTens), P(ctypes.c_size_t), P(ctypes.c_int)]
    lib.talshTensorOpSpecify.argtypes = [P(ts.TalshTensOp), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_double, ctypes.c_double]
    lib.talshTensorOpGetFlopCount.restype = ctypes.c_double
    lib.talshTensorOpGetByteCount.restype = ctypes.c_double
    lib.talshTensorOpGetByteCount.argtypes = [P(ts.TalshTensOp), ctypes.c_uint]
    lib.talshTensorOpGetArgSize.restype = ctypes.c_size_t
    return lib


def sz(v):
    return (ctypes.c_size_t * max(1, len(v)))(*v)


def iv(v):
    return (ctypes.c_int * max(1, len(v)))(*v)


def box(a, dims, off, sub):
    """Box [off, off+sub) of a flat dlf array with extents dims, as a flat dlf array."""
    x = a.reshape(dims, order="F")
    return np.ascontiguousarray(x[tuple(slice(o, o + s) for o, s in zip(off, sub))]).reshape(-1, order="F")


def put(a, dims, off, sub, v):
    x = a.reshape(dims, order="F").copy()
    x[tuple(slice(o, o + s) for o, s in zip(off, sub))] = v.reshape(sub, order="F")
    return x.reshape(-1, order="F")


def run(lib, op):
    """Progress until retired (talshTensorOpProgress), as talshTensorContractXL drives its operations."""
    done = ctypes.c_int(0)
    for _ in range(100000):
        e = lib.talshTensorOpProgress(ctypes.byref(op), ctypes.byref(done))
        assert e == 0, e
        if done.value == 1:
            return
    raise AssertionError("operation did not retire")


def test_slice_views(talsh_gpu):
    ts = talsh_gpu
    lib = lib_of(ts)
    t = ts.Tensor(R8, [5, 6, 7], 0)
    s = ts.TalshTensSlice()
    assert lib.talshTensorSliceClean(ctypes.byref(s)) == 0
    assert lib.talshTensorSliceConstruct(ctypes.byref(s), t.ptr, sz([1, 2, 3]), iv([4, 4, 4]), None, None) == 0
    assert lib.talshTensorSliceVolume(ctypes.byref(s)) == 64
    assert s.bases.num_dim == 3 and [s.bases.offsets[i] for i in range(3)] == [1, 2, 3]
    assert lib.talshTensorSliceDestruct(ctypes.byref(s)) == 0 and s.shape.num_dim == -1
    # a box that leaves the tensor
    assert lib.talshTensorSliceConstruct(ctypes.byref(s), t.ptr, sz([2, 0, 0]), iv([4, 6, 7]), None,
                                         None) == ts.TALSH_INVALID_ARGS
    t.destruct()


@pytest.mark.parametrize("where", ["host", "gpu"])
def test_tensor_op_contraction_on_boxes(talsh_gpu, oracle, where):
    ts = talsh_gpu
    lib = lib_of(ts)
    pat = "D(a,b,c)+=L(c,k,a)*R(b,k)"
    for dk in (R8, C8):
        dd, ld, rd = [20, 18, 16], [19, 30, 22], [21, 27]
        db, lbx, rb = [12, 10, 9], [9, 25, 12], [10, 25]       # box extents (k = 25, a = 12, b = 10, c = 9)
        do, lo, ro = [3, 5, 2], [7, 1, 4], [6, 2]
        say("tensor op", where, dk, pat)
        D, L, R = tensor(dk, dd, 11 + dk), tensor(dk, ld, 12 + dk), tensor(dk, rd, 13 + dk)
        dev = 0 if where == "host" else ts.gpu_flat(0)
        d = ts.Tensor(dk, dd, 0, data=D)
        l, r = ts.Tensor(dk, ld, 0, data=L), ts.Tensor(dk, rd, 0, data=R)
        if where == "gpu":
            for x in (d, l, r):
                x.place(dev)
        op = ts.TalshTensOp()
        assert lib.talshTensorOpClean(ctypes.byref(op)) == 0
        for t, o, b in ((d, do, db), (l, lo, lbx), (r, ro, rb)):
            assert lib.talshTensorOpSetArgument(ctypes.byref(op), t.ptr, sz(o), iv(b)) == 0
        alpha = complex(0.5, -0.25) if dk == C8 else 0.75
        pb = ctypes.create_string_buffer(pat.encode())  # the operation keeps a non-owning pointer to it
        assert lib.talshTensorOpSpecify(ctypes.byref(op), ts.TALSH_TENSOR_CONTRACT, dk, ctypes.addressof(pb),
                                        alpha.real, alpha.imag) == 0
        assert op.stage == ts.TALSH_OP_DEFINED
        cplx = 8.0 if dk == C8 else 2.0
        assert lib.talshTensorOpGetFlopCount(ctypes.byref(op)) == cplx * 12 * 10 * 9 * 25
        esz = 16 if dk == C8 else 8
        assert lib.talshTensorOpGetByteCount(ctypes.byref(op), 1) == esz * (12 * 10 * 9 + 9 * 25 * 12 + 10 * 25)
        assert lib.talshTensorOpGetArgSize(ctypes.byref(op), 2) == esz * 250
        run(lib, op)
        assert op.stage == ts.TALSH_OP_RETIRED
        a = (alpha.real, alpha.imag) if dk == C8 else (alpha, 0.0)
        want_box = oracle.contract(dk, pat, db, box(D, dd, do, db), lbx, box(L, ld, lo, lbx), rb,
                                   box(R, rd, ro, rb), a, 1)
        got = d.host_array()
        assert np.array_equal(put(got, dd, do, db, box(D, dd, do, db)), D)  # outside the box: untouched
        gb = box(got, dd, do, db)
        err = np.linalg.norm((gb - want_box).astype(np.complex128)) / np.linalg.norm(want_box.astype(np.complex128))
        assert err <= 1e-12, err
        assert np.array_equal(l.host_array(), L) and np.array_equal(r.host_array(), R)
        assert lib.talshTensorOpDestruct(ctypes.byref(op)) == 0 and op.stage == ts.TALSH_OP_EMPTY
        for x in (d, l, r):
            x.destruct()


def test_tensor_op_decompose2_children_add_up(talsh_gpu, oracle):
    """Two levels of talshTensorOpDecompose2 (the contracted group is the largest, so K is split first) give
    four operations whose results, accumulated into D, equal the whole contraction."""
    ts = talsh_gpu
    lib = lib_of(ts)
    pat = "D(a,b)+=L(k,a)*R(k,b)"
    dd, ld, rd = [24, 20], [300, 24], [300, 20]
    D, L, R = tensor(R8, dd, 1), tensor(R8, ld, 2), tensor(R8, rd, 3)
    d, l, r = ts.Tensor(R8, dd, 0, data=D), ts.Tensor(R8, ld, 0, data=L), ts.Tensor(R8, rd, 0, data=R)
    parent = ts.TalshTensOp()
    lib.talshTensorOpClean(ctypes.byref(parent))
    for t, dims in ((d, dd), (l, ld), (r, rd)):
        assert lib.talshTensorOpSetArgument(ctypes.byref(parent), t.ptr, sz([0] * len(dims)), iv(dims)) == 0
    pb = ctypes.create_string_buffer(pat.encode())  # the operations keep a non-owning pointer to it
    assert lib.talshTensorOpSpecify(ctypes.byref(parent), ts.TALSH_TENSOR_CONTRACT, R8, ctypes.addressof(pb), 1.0,
                                    0.0) == 0
    level, made = [parent], []
    for _ in range(2):
        made += level
        nxt = []
        for op in level:
            c1, c2 = ts.TalshTensOp(), ts.TalshTensOp()
            lib.talshTensorOpClean(ctypes.byref(c1))
            lib.talshTensorOpClean(ctypes.byref(c2))
            assert lib.talshTensorOpDecompose2(ctypes.byref(op), ctypes.byref(c1), ctypes.byref(c2)) == 0
            nxt += [c1, c2]
        level = nxt
    # first split halves K = 300 (150 + 150); second halves each K range again (75 + 75)
    ks = [op.tens_slice[1].shape.dims[0] for op in level]
    k0 = [op.tens_slice[1].bases.offsets[0] for op in level]
    assert ks == [75, 75, 75, 75] and k0 == [0, 75, 150, 225]
    for op in level:
        run(lib, op)
        assert lib.talshTensorOpDestruct(ctypes.byref(op)) == 0
    for op in made:  # parents: defined, never run
        assert lib.talshTensorOpDestruct(ctypes.byref(op)) == 0
    want = oracle.contract(R8, pat, dd, D, ld, L, rd, R, (1.0, 0.0), 1)
    got = d.host_array()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-12
    for x in (d, l, r):
        x.destruct()
