"""Shared helpers for the tests and the golden-fixture generator.

Seeded inputs, random contraction patterns and ctypes loaders for the two CPU
checkers under oracle/ (the C restatement and the reference build).  Nothing
here is product code; the product is reached only through tal_sh_amd.
"""
import ctypes
import os
import random

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libtalsh_ref.so")

R4, R8, C4, C8 = 4, 8, 14, 18
NP_DTYPE = {R4: np.float32, R8: np.float64, C4: np.complex64, C8: np.complex128}

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed, n):
    """Counter-based splitmix64 stream mapped to uniform[-1,1) (float64).

    Element i is mix(seed + (i+1)*gamma); identical to orc_fill_uniform."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def tensor(dk, dims, seed):
    """Seeded tensor body in dlf order as a flat array of the data kind.

    R4/C4 values are generated as float64 and rounded to float32; complex
    kinds draw the real and imaginary parts from seeds seed and seed+1."""
    n = int(np.prod(dims, dtype=np.int64)) if len(dims) else 1
    if dk in (R4, R8):
        return uniform(seed, n).astype(NP_DTYPE[dk])
    re = uniform(seed, n)
    im = uniform(seed + 1, n)
    return (re + 1j * im).astype(NP_DTYPE[dk])


def random_contraction(rng, max_rank=4, ext_lo=2, ext_hi=6, ncontr=None, allow_scalar=False):
    """Random binary contraction: returns (pattern, ddims, ldims, rdims).

    Labels are shuffled inside each operand, so every permutation pattern
    (including ones that need no transposes) can appear."""
    while True:
        nc = ncontr if ncontr is not None else rng.randint(1, 2)
        nl = rng.randint(0 if allow_scalar else 1, max(0, max_rank - nc))
        nr = rng.randint(0 if allow_scalar else 1, max(0, max_rank - nc))
        if nl + nr < (0 if allow_scalar else 1):
            continue
        labels = [chr(ord("a") + i) for i in range(nl + nr + nc)]
        ext = {x: rng.randint(ext_lo, ext_hi) for x in labels}
        lu, ru, cc = labels[:nl], labels[nl:nl + nr], labels[nl + nr:]
        d = lu + ru
        rng.shuffle(d)
        l = lu + cc
        rng.shuffle(l)
        r = ru + cc
        rng.shuffle(r)
        pat = "D(%s)+=L(%s)*R(%s)" % (",".join(d), ",".join(l), ",".join(r))
        return pat, [ext[x] for x in d], [ext[x] for x in l], [ext[x] for x in r]


def random_hyper_contraction(rng, max_rank=5, ext_lo=1, ext_hi=6, nhyper=None):
    """Random contraction with hyper indices (labels in D, L and R at once).

    Returns (pattern, ddims, ldims, rdims); at least one hyper index, zero or
    more contracted, left-only and right-only indices."""
    nh = nhyper if nhyper is not None else rng.randint(1, 2)
    while True:
        nc, nl, nr = rng.randint(0, 2), rng.randint(0, 2), rng.randint(0, 2)
        if nh + nc + max(nl, nr) > max_rank:
            continue
        labels = [chr(ord("a") + i) for i in range(nh + nc + nl + nr)]
        ext = {x: rng.randint(ext_lo, ext_hi) for x in labels}
        hy, cc = labels[:nh], labels[nh:nh + nc]
        lu, ru = labels[nh + nc:nh + nc + nl], labels[nh + nc + nl:]
        d = hy + lu + ru
        rng.shuffle(d)
        l = hy + lu + cc
        rng.shuffle(l)
        r = hy + ru + cc
        rng.shuffle(r)
        pat = "D(%s)+=L(%s)*R(%s)" % (",".join(d), ",".join(l), ",".join(r))
        return pat, [ext[x] for x in d], [ext[x] for x in l], [ext[x] for x in r]


def batch_contraction(rng, lo=2, hi=24):
    """Batched small contraction: D, L and R each of rank 2-4, 1-2 contracted indices, extents
    uniform in lo..hi, labels shuffled in every operand.  Returns (pattern, ddims, ldims, rdims).
    With every rank at most 4 an operand holds at most hi^4 elements (331,776 for hi = 24)."""
    while True:
        rl, rr = rng.randint(2, 4), rng.randint(2, 4)
        nc = rng.randint(1, min(2, rl, rr))
        nl, nr = rl - nc, rr - nc
        if not 2 <= nl + nr <= 4:
            continue
        labels = [chr(ord("a") + i) for i in range(nl + nr + nc)]
        ext = {x: rng.randint(lo, hi) for x in labels}
        lu, ru, cc = labels[:nl], labels[nl:nl + nr], labels[nl + nr:]
        d, l, r = lu + ru, lu + cc, ru + cc
        rng.shuffle(d)
        rng.shuffle(l)
        rng.shuffle(r)
        pat = "D(%s)+=L(%s)*R(%s)" % (",".join(d), ",".join(l), ",".join(r))
        return pat, [ext[x] for x in d], [ext[x] for x in l], [ext[x] for x in r]


def einsum_contract(pat, dd, D, ld, L, rd, R, alpha=1.0, accumulative=1, conj_l=False, conj_r=False):
    """numpy einsum of a TAL-SH pattern over Fortran-ordered (dlf) views; returns flat D."""
    lhs, rhs = pat.split("+=")
    dl = lhs[lhs.index("(") + 1:-1].replace(",", "")
    lp, rp = rhs.split("*")
    ll = lp[lp.index("(") + 1:-1].replace(",", "")
    rl = rp[rp.index("(") + 1:-1].replace(",", "")
    a = L.reshape(ld, order="F") if ld else L.reshape(())
    b = R.reshape(rd, order="F") if rd else R.reshape(())
    if conj_l:
        a = np.conj(a)
    if conj_r:
        b = np.conj(b)
    E = np.einsum("%s,%s->%s" % (ll, rl, dl), a.astype(np.complex128), b.astype(np.complex128))
    out = alpha * np.asarray(E).reshape(-1, order="F")
    return (D.astype(np.complex128) + out) if accumulative else out


def random_permutation_pattern(rng, rank):
    labels = [chr(ord("a") + i) for i in range(rank)]
    perm = labels[:]
    rng.shuffle(perm)
    return "D(%s)=L(%s)" % (",".join(labels), ",".join(perm)), perm, labels


def _intp(a):
    return (ctypes.c_int * max(1, len(a)))(*a)


class Oracle:
    """ctypes view of oracle/liboracle.so (the C restatement)."""

    def __init__(self, path=ORACLE_SO):
        self.lib = ctypes.CDLL(path)
        self.lib.orc_norm2.restype = ctypes.c_double
        self.lib.orc_norm2.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]

    def parse(self, pattern):
        dig = (ctypes.c_int * 128)()
        dr, lr, rr, cj = (ctypes.c_int() for _ in range(4))
        err = self.lib.orc_parse_pattern(pattern.encode(), dig, ctypes.byref(dr), ctypes.byref(lr),
                                         ctypes.byref(rr), ctypes.byref(cj))
        n = max(0, lr.value) + max(0, rr.value)
        return err, dr.value, lr.value, rr.value, cj.value, list(dig[:n]) if err == 0 else []

    def contr_permutations(self, gemm_tl, gemm_tr, lrank, rrank, dig, conj):
        dp, lp, rp = ((ctypes.c_int * 64)() for _ in range(3))
        ncd, nlu, nru = (ctypes.c_int() for _ in range(3))
        err = self.lib.orc_contr_permutations(gemm_tl, gemm_tr, lrank, rrank, _intp(dig), conj, dp, lp, rp,
                                              ctypes.byref(ncd), ctypes.byref(nlu), ctypes.byref(nru))
        return err, list(dp[:nlu.value + nru.value + 1]), list(lp[:lrank + 1]), list(rp[:rrank + 1]), \
            ncd.value, nlu.value, nru.value

    def copy(self, dk, ldims, o2n, l, conj=0):
        ddims = [0] * len(ldims)
        for k, t in enumerate(o2n):
            ddims[t - 1] = ldims[k]
        d = np.zeros(int(np.prod(ddims)) if ddims else 1, dtype=NP_DTYPE[dk])
        err = self.lib.orc_copy(dk, len(ldims), _intp(ldims), _intp(o2n), l.ctypes.data_as(ctypes.c_void_p),
                                d.ctypes.data_as(ctypes.c_void_p), conj)
        assert err == 0, err
        return d

    def add(self, dk, ldims, o2n, l, d, alpha=(1.0, 0.0), conj=0):
        d = d.copy()
        err = self.lib.orc_add(dk, len(ldims), _intp(ldims), _intp(o2n), l.ctypes.data_as(ctypes.c_void_p),
                               d.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(alpha[0]),
                               ctypes.c_double(alpha[1]), conj)
        assert err == 0, err
        return d

    def contract(self, dk, pattern, ddims, d, ldims, l, rdims, r, alpha=(1.0, 0.0), accumulative=1):
        err, dr, lr, rr, cj, dig = self.parse(pattern)
        assert err == 0, (pattern, err)
        d = d.copy()
        err = self.lib.orc_contract(dk, _intp(dig), cj, dr, _intp(ddims), d.ctypes.data_as(ctypes.c_void_p),
                                    lr, _intp(ldims), l.ctypes.data_as(ctypes.c_void_p), rr, _intp(rdims),
                                    r.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(alpha[0]),
                                    ctypes.c_double(alpha[1]), accumulative)
        assert err == 0, (pattern, err)
        return d

    def contract_element(self, dk, pattern, ddims, doff, ldims, l, rdims, r):
        err, dr, lr, rr, cj, dig = self.parse(pattern)
        assert err == 0
        re, im = ctypes.c_double(), ctypes.c_double()
        self.lib.orc_contract_element(dk, _intp(dig), cj, dr, _intp(ddims), ctypes.c_size_t(doff), lr,
                                      _intp(ldims), l.ctypes.data_as(ctypes.c_void_p), rr, _intp(rdims),
                                      r.ctypes.data_as(ctypes.c_void_p), ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def contract_direct(self, dk, pattern, ddims, d, ldims, l, rdims, r, alpha=(1.0, 0.0), accumulative=1):
        """Direct loop-nest contraction (hyper indices allowed); returns the new D."""
        err, dr, lr, rr, cj, dig = self.parse(pattern)
        assert err == 0
        out = np.array(d, dtype=NP_DTYPE[dk], copy=True).reshape(-1)
        e = self.lib.orc_contract_direct(dk, _intp(dig), cj, dr, _intp(ddims), out.ctypes.data_as(ctypes.c_void_p),
                                         lr, _intp(ldims), l.ctypes.data_as(ctypes.c_void_p), rr, _intp(rdims),
                                         r.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(alpha[0]),
                                         ctypes.c_double(alpha[1]), accumulative)
        assert e == 0, e
        return out

    def norm2(self, dk, a):
        return self.lib.orc_norm2(dk, a.size, a.ctypes.data_as(ctypes.c_void_p))

    def scale(self, dk, a, alpha):
        a = a.copy()
        self.lib.orc_scale(dk, ctypes.c_size_t(a.size), a.ctypes.data_as(ctypes.c_void_p),
                           ctypes.c_double(alpha[0]), ctypes.c_double(alpha[1]))
        return a

    def slice(self, dk, sdims, s, bdims, b, offsets, accumulative=0):
        """s (+)= b[offsets + i]; returns the new s."""
        s = s.copy()
        b = np.ascontiguousarray(b)
        err = self.lib.orc_slice(dk, len(sdims), _intp(sdims), s.ctypes.data_as(ctypes.c_void_p), _intp(bdims),
                                 b.ctypes.data_as(ctypes.c_void_p), _intp(offsets), accumulative)
        assert err == 0, err
        return s

    def insert(self, dk, sdims, s, bdims, b, offsets, accumulative=0):
        """b[offsets + i] (+)= s; returns the new b."""
        b = b.copy()
        s = np.ascontiguousarray(s)
        err = self.lib.orc_insert(dk, len(sdims), _intp(sdims), s.ctypes.data_as(ctypes.c_void_p), _intp(bdims),
                                  b.ctypes.data_as(ctypes.c_void_p), _intp(offsets), accumulative)
        assert err == 0, err
        return b


class Reference:
    """ctypes view of oracle/_ref/libtalsh_ref.so (TAL-SH's own CPU path)."""

    def __init__(self, path=REF_SO):
        self.lib = ctypes.CDLL(path)
        self.lib.ref_norm2.restype = ctypes.c_double
        assert self.lib.ref_init() == 0

    def parse(self, pattern):
        dig = (ctypes.c_int * 128)()
        dr, lr, rr, cj = (ctypes.c_int() for _ in range(4))
        err = self.lib.ref_str2dig(pattern.encode(), dig, ctypes.byref(dr), ctypes.byref(lr), ctypes.byref(rr),
                                   ctypes.byref(cj))
        n = max(0, lr.value) + max(0, rr.value)
        return err, dr.value, lr.value, rr.value, cj.value, list(dig[:n]) if err == 0 else []

    def contr_permutations(self, gemm_tl, gemm_tr, lrank, rrank, dig, conj):
        dp, lp, rp = ((ctypes.c_int * 64)() for _ in range(3))
        ncd, nlu, nru = (ctypes.c_int() for _ in range(3))
        err = self.lib.ref_contr_permutations(gemm_tl, gemm_tr, lrank, rrank, _intp(dig), conj, dp, lp, rp,
                                              ctypes.byref(ncd), ctypes.byref(nlu), ctypes.byref(nru))
        return err, list(dp[:nlu.value + nru.value + 1]), list(lp[:lrank + 1]), list(rp[:rrank + 1]), \
            ncd.value, nlu.value, nru.value

    def copy(self, dk, pattern, ddims, ldims, l):
        d = np.zeros(int(np.prod(ddims)) if ddims else 1, dtype=NP_DTYPE[dk])
        l = l.copy()
        err = self.lib.ref_copy(pattern.encode(), dk, len(ddims), _intp(ddims), d.ctypes.data_as(ctypes.c_void_p),
                                len(ldims), _intp(ldims), l.ctypes.data_as(ctypes.c_void_p))
        assert err == 0, err
        return d

    def add(self, dk, pattern, ddims, d, ldims, l, alpha=(1.0, 0.0)):
        d = d.copy()
        l = l.copy()
        err = self.lib.ref_add(pattern.encode(), dk, len(ddims), _intp(ddims), d.ctypes.data_as(ctypes.c_void_p),
                               len(ldims), _intp(ldims), l.ctypes.data_as(ctypes.c_void_p),
                               ctypes.c_double(alpha[0]), ctypes.c_double(alpha[1]))
        assert err == 0, err
        return d

    def contract(self, dk, pattern, ddims, d, ldims, l, rdims, r, alpha=(1.0, 0.0), accumulative=1):
        d = d.copy()
        l = l.copy()
        r = r.copy()
        err = self.lib.ref_contract(pattern.encode(), dk, len(ddims), _intp(ddims),
                                    d.ctypes.data_as(ctypes.c_void_p), len(ldims), _intp(ldims),
                                    l.ctypes.data_as(ctypes.c_void_p), len(rdims), _intp(rdims),
                                    r.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(alpha[0]),
                                    ctypes.c_double(alpha[1]), accumulative)
        assert err == 0, err
        return d

    def norm2(self, dk, dims, a):
        return self.lib.ref_norm2(dk, len(dims), _intp(dims), a.ctypes.data_as(ctypes.c_void_p))

    def scale(self, dk, dims, d, alpha):
        d = d.copy()
        err = self.lib.ref_scale(dk, len(dims), _intp(dims), d.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_double(alpha[0]), ctypes.c_double(alpha[1]))
        assert err == 0, err
        return d

    def slice(self, dk, sdims, s, bdims, b, offsets, accumulative=0):
        s, b = s.copy(), b.copy()
        err = self.lib.ref_slice(dk, len(sdims), _intp(sdims), s.ctypes.data_as(ctypes.c_void_p), _intp(bdims),
                                 b.ctypes.data_as(ctypes.c_void_p), _intp(offsets), accumulative)
        assert err == 0, err
        return s

    def insert(self, dk, sdims, s, bdims, b, offsets, accumulative=0):
        s, b = s.copy(), b.copy()
        err = self.lib.ref_insert(dk, len(sdims), _intp(sdims), s.ctypes.data_as(ctypes.c_void_p), _intp(bdims),
                                  b.ctypes.data_as(ctypes.c_void_p), _intp(offsets), accumulative)
        assert err == 0, err
        return b

    def pattern_sym(self, lrank, rrank, conj, dig):
        buf = ctypes.create_string_buffer(4096)
        cpl = ctypes.c_int()
        err = self.lib.ref_pattern_sym(lrank, rrank, conj, _intp(dig), buf, ctypes.byref(cpl))
        return err, buf.value.decode() if err == 0 else ""


def _split_volume(rng, vol, n, lo=2):
    """n extents (each >= lo) whose product is close to vol."""
    if n == 0:
        return []
    dims = []
    rem = float(vol)
    for i in range(n - 1):
        base = rem ** (1.0 / (n - i))
        e = max(lo, int(round(base * 2.0 ** rng.uniform(-0.5, 0.5))))
        dims.append(e)
        rem = max(1.0, rem / e)
    dims.append(max(lo, int(round(rem))))
    return dims


def serving_request(rng, vmin=1e3, vmax=1e8):
    """One serving task: operand volumes log-uniform in [vmin, vmax] elements,
    operand ranks 2-6, D rank 2-6, 1+ contracted indices, D volume <= vmax.
    Returns (pattern, ddims, ldims, rdims, flop)."""
    import math
    while True:
        rl, rr = rng.randint(2, 6), rng.randint(2, 6)
        cs = [c for c in range(1, min(rl, rr) + 1) if 2 <= rl + rr - 2 * c <= 6 and c < max(rl, rr)]
        if not cs:
            continue
        c = rng.choice(cs)
        nl, nr = rl - c, rr - c
        vl = math.exp(rng.uniform(math.log(vmin), math.log(vmax)))
        vr = math.exp(rng.uniform(math.log(vmin), math.log(vmax)))
        kmin = max(2.0 ** c, math.sqrt(vl * vr / vmax))
        kmax = min(vl / 2.0 ** nl, vr / 2.0 ** nr)
        if kmin > kmax:
            continue
        K = math.exp(rng.uniform(math.log(kmin), math.log(kmax)))
        kd = _split_volume(rng, K, c)
        Kv = float(np.prod(kd))
        md = _split_volume(rng, vl / Kv, nl)
        nd = _split_volume(rng, vr / Kv, nr)
        M, N = float(np.prod(md)) if md else 1.0, float(np.prod(nd)) if nd else 1.0
        if M * Kv > 1.3 * vmax or N * Kv > 1.3 * vmax or M * N > 1.3 * vmax:
            continue
        labels = [chr(ord("a") + i) for i in range(nl + nr + c)]
        lu, ru, cc = labels[:nl], labels[nl:nl + nr], labels[nl + nr:]
        ext = dict(zip(lu + ru + cc, md + nd + kd))
        d, l, r = lu + ru, lu + cc, ru + cc
        rng.shuffle(d)
        rng.shuffle(l)
        rng.shuffle(r)
        pat = "D(%s)+=L(%s)*R(%s)" % (",".join(d), ",".join(l), ",".join(r))
        return pat, [ext[x] for x in d], [ext[x] for x in l], [ext[x] for x in r], 2.0 * M * N * Kv
