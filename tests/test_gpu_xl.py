"""talshTensorContractXL (talshc.cpp:5394-5640) on the GPU against the CPU oracle.

The call cuts a contraction into pieces that fit the device memory granted to
it (talshAmdSetContractXLMemory shrinks that, so small cases split into many
pieces), brings operand boxes from the host image (strided DMA) or from the
GPU image (slice kernel), and leaves D with one image, on the host
(talshc.cpp:5460-5461).  Bars: R8/C8 normwise relative error <= 1e-12 against
the oracle's long-double contraction; R4 normwise error <= 1e-5 of the norm of
the same contraction over absolute values (a split reduction such as D() += L*R
cancels, so the result itself can be far smaller than its terms); with one
piece the result equals talshTensorContract's bit for bit (same plan, kernels).
"""
import ctypes
import random
import re

import numpy as np
import pytest

from tal_common import C4, C8, R4, R8, random_contraction, tensor

pytestmark = pytest.mark.gpu

_LIBC = ctypes.CDLL(None)


def say(*a):
    print("[case]", *a, flush=True)


def relerr(got, want):
    den = np.linalg.norm(want.astype(np.complex128))
    return np.linalg.norm((got - want).astype(np.complex128)) / (den if den > 0 else 1.0)


def place(ts, where, dk, dims, data):
    """Tensor holding `data`: 'arena' = Host image in the library's pinned host buffer,
    'ext' = Host image attached to caller (pageable) memory, 'gpu' = image on GPU 0 only."""
    if where == "arena":
        return ts.Tensor(dk, dims, 0, data=data)
    if where == "ext":
        return ts.Tensor(dk, dims, 0, ext=np.ascontiguousarray(data).copy())
    t = ts.Tensor(dk, dims, ts.gpu_flat(0))
    h = ts.Tensor(dk, dims, 0, data=data)
    lab = ",".join("i%d" % k for k in range(len(dims)))
    ts.check(ts.copy("D(%s)=L(%s)" % (lab, lab), t, h), "upload")
    h.destruct()
    return t


def xl_log(capfd, fn):
    """Runs fn with op logging on; returns (fn's result, pieces reported by the XL call)."""
    lib = __import__("tal_sh_amd").load()
    capfd.readouterr()
    lib.talshTensorOpLogStart()
    try:
        res = fn()
    finally:
        lib.talshTensorOpLogFinish()
        _LIBC.fflush(None)
    out = capfd.readouterr().out
    m = re.findall(r"talshTensorContractXL on GPU \d+: (\d+) pieces", out)
    return res, (int(m[-1]) if m else None)


def test_xl_streams_host_operands(talsh_gpu, oracle, capfd):
    """Operands on the host and a contraction that fits: the larger input is cut into 4 boxes along
    its slowest D index (one box crosses the host link while another is contracted) and the other
    input crosses once, whole."""
    ts = talsh_gpu
    cases = [("D(a,b,c,d)+=L(a,i,c,j)*R(j,b,i,d)", [12, 10, 16, 9], [12, 7, 16, 5], [5, 10, 7, 9]),
             ("D(a,b)+=L(k,b)*R(a,k)", [33, 64], [50, 64], [33, 50])]
    for t, (pat, dd, ld, rd) in enumerate(cases):
        for dk in (R8, C8):
            for acc in (1, 0):
                say("xl stream", dk, pat, "acc", acc)
                L, R, D = tensor(dk, ld, 7 * t + dk), tensor(dk, rd, 7 * t + dk + 1), tensor(dk, dd, 7 * t + dk + 2)
                lx, rx, dx = place(ts, "arena", dk, ld, L), place(ts, "ext", dk, rd, R), place(ts, "arena", dk, dd, D)
                err, pieces = xl_log(capfd, lambda: ts.contract_xl(pat, dx, lx, rx, alpha=0.5, accumulative=acc))
                assert err == 0 and pieces == 4, (err, pieces)
                assert dx.images() == [ts.DEV_HOST]
                want = oracle.contract_direct(dk, pat, dd, D, ld, L, rd, R, (0.5, 0.0), acc)
                assert relerr(dx.host_array(), want) <= 1e-12
                for x in (lx, rx, dx):
                    x.destruct()


def test_xl_one_piece_equals_contract(talsh_gpu, oracle, capfd):
    """GPU-resident operands that fit: one piece, bit-identical to talshTensorContract."""
    ts = talsh_gpu
    rng = random.Random(5)
    for t in range(8):
        dk = (R8, C8, R4, C4)[t % 4]
        pat, dd, ld, rd = random_contraction(rng, max_rank=4, ext_lo=2, ext_hi=12)
        acc = t % 2
        alpha = complex(0.5, -0.25) if dk in (C4, C8) else 0.75
        say("xl one piece", dk, pat, dd, ld, rd, "acc", acc)
        L, R, D = tensor(dk, ld, 3 * t), tensor(dk, rd, 3 * t + 1), tensor(dk, dd, 3 * t + 2)
        lx, rx, dx = place(ts, "gpu", dk, ld, L), place(ts, "gpu", dk, rd, R), place(ts, "gpu", dk, dd, D)
        err, pieces = xl_log(capfd, lambda: ts.contract_xl(pat, dx, lx, rx, alpha=alpha, accumulative=acc))
        assert err == 0 and pieces == 1, (err, pieces)
        assert dx.images() == [ts.DEV_HOST]
        lg, rg, dg = place(ts, "gpu", dk, ld, L), place(ts, "gpu", dk, rd, R), place(ts, "gpu", dk, dd, D)
        ts.check(ts.contract(pat, dg, lg, rg, alpha=alpha, accumulative=acc), "contract")
        got, ref = dx.host_array(), dg.host_array()
        assert np.array_equal(got, ref), (dk, pat)
        a = (alpha.real, alpha.imag) if dk in (C4, C8) else (alpha, 0.0)
        want = oracle.contract(dk, pat, dd, D, ld, L, rd, R, a, acc)
        assert relerr(got, want) <= (1e-12 if dk in (R8, C8) else 1e-5)
        for x in (lx, rx, dx, lg, rg, dg):
            x.destruct()


# (pattern, ddims, ldims, rdims): the largest index group differs from case to case,
# so the tiling splits contracted indices, D indices of L, D indices of R, a hyper
# index (in D, L and R), or several
SPLIT_CASES = [
    ("D(a,h,b)+=L(a,k,h)*R(k,h,b)", [30, 40, 20], [30, 25, 40], [25, 40, 20]),  # hyper index h
    ("D(a,b)+=L(a,k)*R(k,b)", [40, 36], [40, 3000], [3000, 36]),            # long K: contracted tiles
    ("D(a,b,c)+=L(c,k,a)*R(b,k)", [90, 70, 30], [30, 12, 90], [70, 12]),    # large D: D boxes
    ("D(a,b,c,d)+=L(c,i,b,j)*R(d,j,a,i)", [14, 13, 12, 11], [12, 9, 13, 10], [11, 10, 14, 9]),  # both
    ("D()+=L(i,j,k)*R(k,j,i)", [], [30, 40, 50], [50, 40, 30]),              # scalar result
    ("D(a,b)+=L(b)*R(a)", [700, 600], [600], [700]),                        # outer product (K = 1)
]


@pytest.mark.parametrize("where", ["arena", "ext", "gpu"])
def test_xl_split_matches_oracle(talsh_gpu, oracle, capfd, where):
    ts = talsh_gpu
    assert ts.set_xl_memory(1 << 17) == 0  # 128 KiB: every case below splits
    try:
        for t, (pat, dd, ld, rd) in enumerate(SPLIT_CASES):
            for dk in (R8, C8, R4):
                for acc in (1, 0):
                    seed = 100 * t + 10 * dk + acc
                    alpha = complex(-0.5, 0.75) if dk in (C4, C8) else 1.25
                    say("xl split", where, dk, pat, dd, ld, rd, "acc", acc)
                    L, R, D = tensor(dk, ld, seed), tensor(dk, rd, seed + 1), tensor(dk, dd, seed + 2)
                    lx, rx = place(ts, where, dk, ld, L), place(ts, where, dk, rd, R)
                    dx = place(ts, where, dk, dd, D)
                    err, pieces = xl_log(capfd, lambda: ts.contract_xl(pat, dx, lx, rx, alpha=alpha,
                                                                       accumulative=acc))
                    assert err == 0, err
                    assert pieces is not None and pieces > 1, pieces
                    assert dx.images() == [ts.DEV_HOST]
                    a = (alpha.real, alpha.imag) if dk in (C4, C8) else (alpha, 0.0)
                    want = oracle.contract_direct(dk, pat, dd, D, ld, L, rd, R, a, acc)
                    got = dx.host_array()
                    if dk in (R8, C8):
                        e = relerr(got, want)
                        assert e <= 1e-12, (dk, pat, acc, pieces, e)
                    else:  # FP32: error against the size of the terms (sums such as D() cancel)
                        mag = oracle.contract_direct(R8, pat, dd, np.abs(D).astype(np.float64), ld,
                                              np.abs(L).astype(np.float64), rd, np.abs(R).astype(np.float64),
                                              (abs(alpha), 0.0), acc)
                        e = np.linalg.norm((got - want).astype(np.complex128)) / np.linalg.norm(mag)
                        assert e <= 1e-5, (dk, pat, acc, pieces, e)
                    # inputs are untouched
                    assert np.array_equal(lx.host_array(), L) and np.array_equal(rx.host_array(), R)
                    for x in (lx, rx, dx):
                        x.destruct()
    finally:
        ts.set_xl_memory(0)


def test_xl_errors(talsh_gpu):
    ts = talsh_gpu
    d, l, r = ts.Tensor(R8, [4, 4], 0), ts.Tensor(R8, [4, 3], 0), ts.Tensor(R8, [3, 4], 0)
    lib = ts.load()
    pat = "D(a,b)+=L(a,k)*R(k,b)"
    assert ts.contract_xl(pat, d, l, r, dev_id=0, dev_kind=ts.DEV_HOST) == ts.TALSH_NOT_AVAILABLE
    assert ts.contract_xl(pat, l, l, r) == ts.TALSH_INVALID_ARGS          # D aliases L
    assert ts.contract_xl("D(a,b)+=L(a,k)*R(k)", d, l, r) == ts.TALSH_INVALID_ARGS
    h = ts.Tensor(R8, [4, 4], 0)
    empty = ts.TalshTens()
    assert lib.talshTensorClean(ctypes.byref(empty)) == 0
    assert lib.talshTensorContractXL(pat.encode(), ctypes.byref(empty), l.ptr, r.ptr, 1.0, 0.0, ts.DEV_DEFAULT,
                                     ts.DEV_DEFAULT, 1) == ts.TALSH_OBJECT_IS_EMPTY
    for x in (d, l, r, h):
        x.destruct()
