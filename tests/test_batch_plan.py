"""CPU tests of the batched-contraction planner (tal_sh_amd/csrc/batch_plan.cpp).

The batch kernel (batch.hip) reads operands straight from their dlf storage:
every row, column and contracted index becomes a byte offset through the
mixed-radix digit functions of batch_item.hpp, and indices past an edge get an
offset outside the operand so that the hardware range check of buffer loads
returns 0.  These tests build the real planner with g++ (tests/cxx/
batch_shim.cpp) and check, without a GPU:
  - the multiply-high division against integer division, exhaustively for
    divisors up to 4096 (the digit decomposition rests on it);
  - a thread-by-thread emulation of the kernel's workgroups (same address
    functions and staging maps, MFMA operand / accumulator layouts of each
    wave's quadrant, out-of-range loads read 0 and stores are dropped)
    reproduces the oracle's contraction for config-4-like contractions and
    every workgroup tile, with no offset
    between an operand's end and the out-of-range marker, every staged tile
    entry written exactly once per chunk and every D element stored once;
  - groups that cannot be fused into three dimensions, and hyper indices, are
    refused (the single-contraction path runs them);
  - the work queues hold every item exactly once, keep a small contraction's
    items side by side in one queue and balance the queues to within one unit
    (by the planner's own cost estimate).
"""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

from tal_common import R4, R8, batch_contraction, tensor

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TILES = ((32, 32), (32, 64), (64, 32), (64, 64))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("batch") / "libbatch_shim.so")
    csrc = os.path.join(REPO, "tal_sh_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + csrc, "-I" + os.path.join(REPO, "include"),
                    os.path.join(csrc, "batch_plan.cpp"), os.path.join(csrc, "plan.cpp"),
                    os.path.join(REPO, "tests", "cxx", "batch_shim.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    I, P, D, LL = ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_longlong
    IP = ctypes.POINTER(I)
    lib.batch_divider_check.restype = LL
    lib.batch_divider_check.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.batch_emulate.restype = I
    lib.batch_emulate.argtypes = [ctypes.c_char_p, IP, IP, IP, I, P, P, P, D, I, I, I, ctypes.POINTER(LL), IP,
                                  ctypes.POINTER(LL), IP]
    lib.batch_queue_probe.restype = LL
    lib.batch_queue_probe.argtypes = [I, IP, IP, IP, IP, IP, I, I, IP, IP, IP, ctypes.POINTER(D), IP]
    return lib


def _ints(v):
    return (ctypes.c_int * max(1, len(v)))(*v)


def emulate(lib, pat, dd, ld, rd, dk, L, R, D, alpha, acc, tile):
    esz = 8 if dk == R8 else 4
    D = np.array(D, copy=True)
    count = np.zeros(D.size, dtype=np.int32)
    bad, items = ctypes.c_longlong(), ctypes.c_longlong()
    nd = (ctypes.c_int * 3)()
    rc = lib.batch_emulate(pat.encode(), _ints(dd), _ints(ld), _ints(rd), esz, L.ctypes.data, R.ctypes.data,
                           D.ctypes.data, alpha, acc, tile[0], tile[1], ctypes.byref(bad),
                           count.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(items), nd)
    return rc, D, count, bad.value, items.value, list(nd)


def test_divider_exact(shim):
    assert shim.batch_divider_check(4096, 1 << 15) == 0


@pytest.mark.parametrize("dk", (R8, R4))
def test_emulated_kernel_matches_oracle(shim, oracle, dk):
    rng = random.Random(11 + dk)
    tol = 1e-12 if dk == R8 else 2e-5
    for t in range(36):
        pat, dd, ld, rd = batch_contraction(rng, 2, 9)
        L, R, D0 = tensor(dk, ld, 3 * t + 1), tensor(dk, rd, 3 * t + 2), tensor(dk, dd, 3 * t + 3)
        acc = t % 3 != 0
        alpha = 0.75 if t % 2 else 1.0
        want = oracle.contract(dk, pat, dd, D0, ld, L, rd, R, (alpha, 0.0), 1 if acc else 0)
        den = np.linalg.norm(want.astype(np.float64)) or 1.0
        for tile in TILES:
            rc, got, count, bad, items, nd = emulate(shim, pat, dd, ld, rd, dk, L, R, D0, alpha, 1 if acc else 0,
                                                    tile)
            assert rc == 0, (pat, dd, ld, rd, rc)
            assert bad == 0, (pat, dd, ld, rd, tile, bad)
            assert (count == 1).all(), (pat, tile, int(count.min()), int(count.max()))
            err = np.linalg.norm((got - want).astype(np.float64)) / den
            assert err <= tol, (pat, dd, ld, rd, tile, err)


def test_emulated_edge_shapes(shim, oracle):
    """Fully contracted L (M = 1), fully contracted R (N = 1), a dot product
    (D rank 0), an outer product (K = 1), extents of 1, and rows that fuse
    (dims adjacent in L and D) or are ordered by L's strides."""
    cases = [
        ("D(a,b)+=L(i,j)*R(j,a,i,b)", [5, 7], [3, 4], [4, 5, 3, 7]),
        ("D(a)+=L(a,i,j)*R(j,i)", [33], [33, 6, 2], [2, 6]),
        ("D()+=L(i,j)*R(j,i)", [], [9, 5], [5, 9]),
        ("D(a,b)+=L(a)*R(b)", [17, 19], [17], [19]),
        ("D(a,b,c)+=L(a,i,b)*R(c,i)", [1, 3, 1], [1, 5, 3], [1, 5]),
        ("D(a,b,c,d)+=L(a,b,i)*R(c,i,d)", [4, 5, 6, 3], [4, 5, 7], [6, 7, 3]),
        ("D(b,a,c)+=L(a,b,i)*R(i,c)", [6, 5, 20], [5, 6, 7], [7, 20]),
    ]
    for dk in (R8, R4):
        tol = 1e-12 if dk == R8 else 2e-5
        for t, (pat, dd, ld, rd) in enumerate(cases):
            L, R, D0 = tensor(dk, ld, 50 + t), tensor(dk, rd, 60 + t), tensor(dk, dd, 70 + t)
            want = oracle.contract(dk, pat, dd, D0, ld, L, rd, R, (1.0, 0.0), 1)
            den = np.linalg.norm(want.astype(np.float64)) or 1.0
            for tile in TILES:
                rc, got, count, bad, items, nd = emulate(shim, pat, dd, ld, rd, dk, L, R, D0, 1.0, 1, tile)
                assert rc == 0 and bad == 0, (pat, tile, rc, bad)
                assert (count == 1).all(), (pat, tile)
                assert np.linalg.norm((got - want).astype(np.float64)) / den <= tol, (pat, tile)
    # rows a, b adjacent in L and in D fuse into one dimension
    L, R, D0 = tensor(R8, [4, 5, 7], 1), tensor(R8, [6, 7, 3], 2), tensor(R8, [4, 5, 6, 3], 3)
    rc, _, _, _, _, nd = emulate(shim, "D(a,b,c,d)+=L(a,b,i)*R(c,i,d)", [4, 5, 6, 3], [4, 5, 7], [6, 7, 3], R8, L, R,
                                 D0, 1.0, 1, (32, 32))
    assert rc == 0 and nd == [1, 2, 1], nd


def test_geometry_refusals(shim):
    # four row dimensions that chain in L but not in D: no three-dimension form
    dd, ld, rd = [3, 4, 2, 2], [2, 3, 2, 4, 5], [5]
    L, R, D = np.zeros(240), np.zeros(5), np.zeros(48)
    rc, *_ = emulate(shim, "D(b,d,a,c)+=L(a,b,c,d,i)*R(i)", dd, ld, rd, R8, L, R, D, 1.0, 0, (32, 32))
    assert rc == 2
    # a hyper index (in D, L and R)
    rc, *_ = emulate(shim, "D(a,b,h)+=L(a,i,h)*R(i,b,h)", [3, 4, 2], [3, 5, 2], [5, 4, 2], R8, np.zeros(30),
                     np.zeros(40), np.zeros(24), 1.0, 0, (32, 32))
    assert rc == 2
    # extents that do not match the pattern
    rc, *_ = emulate(shim, "D(a,b)+=L(a,i)*R(i,b)", [3, 4], [3, 5], [6, 4], R8, np.zeros(15), np.zeros(24),
                     np.zeros(12), 1.0, 0, (32, 32))
    assert rc == 1


def test_queues_place_every_item_once(shim):
    rng = random.Random(5)
    ntask = 400
    M = [rng.choice([1, 3, 16, 40, 200, 1500]) for _ in range(ntask)]
    N = [rng.choice([1, 7, 16, 33, 300, 2000]) for _ in range(ntask)]
    K = [rng.randint(1, 300) for _ in range(ntask)]
    tiles = [TILES[rng.randrange(len(TILES))] for _ in range(ntask)]
    want = set()
    for t in range(ntask):
        for n0 in range(0, N[t], tiles[t][1]):
            for m0 in range(0, M[t], tiles[t][0]):
                want.add((t, m0, n0))
    nq = 8
    n = len(want)
    ot, om, on = (np.zeros(n, dtype=np.int32) for _ in range(3))
    qstart = np.zeros(nq + 1, dtype=np.int32)
    icost = np.zeros(n)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    got_n = shim.batch_queue_probe(ntask, _ints(M), _ints(N), _ints(K), _ints([x[0] for x in tiles]),
                                   _ints([x[1] for x in tiles]), 8, nq, ip(ot), ip(om), ip(on),
                                   icost.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ip(qstart))
    assert got_n == n
    got = list(zip(ot.tolist(), om.tolist(), on.tolist()))
    assert len(set(got)) == n and set(got) == want
    assert qstart[0] == 0 and qstart[-1] == n and (np.diff(qstart) >= 0).all()
    queue_of = np.zeros(n, dtype=np.int64)
    for q in range(nq):
        queue_of[qstart[q]:qstart[q + 1]] = q
    tcost = np.bincount(ot, weights=icost, minlength=ntask)
    chunk = icost.sum() / nq / 8
    for t in range(ntask):
        pos = np.nonzero(ot == t)[0]
        if tcost[t] < chunk:  # a small contraction is one unit: one queue, its items side by side
            assert len(set(queue_of[pos].tolist())) == 1, t
            assert pos.max() - pos.min() + 1 == len(pos), t
    # queues are loaded to within the largest unit (at most about 2 * chunk)
    qcost = np.array([icost[qstart[q]:qstart[q + 1]].sum() for q in range(nq)])
    assert qcost.max() - qcost.min() <= 2 * chunk + 1e-9, (qcost.max(), qcost.min(), chunk)
