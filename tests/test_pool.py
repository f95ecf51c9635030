"""CPU test of the device buffer pool's allocator (tal::Arena, tal_sh_amd/csrc/pool.cpp).

The arena replaces the reference's buffer tree (get_buf_entry / free_buf_entry,
mem_manager.cpp:544-692) with a best-fit, coalescing allocator over one region,
keeping TAL-SH's contract: TRY_LATER when a request fits the region but no free
range is large enough now, DEVICE_UNABLE when it can never fit.  Random
alloc/free churn is replayed against a brute-force model of the free list:
  - every live range is aligned, inside the region and disjoint from the others;
  - the chosen range starts a free range of the smallest length that fits (best fit);
  - TRY_LATER exactly when no free range fits, DEVICE_UNABLE exactly when the
    rounded request exceeds the region;
  - free bytes, largest free range and live entry count equal the model's;
  - releasing everything coalesces back into one free range of the whole region.
pool.cpp is compiled here with g++ together with tests/cxx/arena_shim.cpp.
"""
import ctypes
import os
import random
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRY_LATER, DEVICE_UNABLE = -918273645, -546372819


@pytest.fixture(scope="module")
def arena_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("arena") / "libarena_shim.so")
    csrc = os.path.join(REPO, "tal_sh_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I" + csrc, "-I" + os.path.join(REPO, "include"),
                    os.path.join(csrc, "pool.cpp"), os.path.join(REPO, "tests", "cxx", "arena_shim.cpp"), "-o", out],
                   check=True)
    lib = ctypes.CDLL(out)
    S, P, I = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
    for name, res, args in (("arena_new", P, [S, S]), ("arena_delete", None, [P]),
                            ("arena_alloc", I, [P, S, ctypes.POINTER(S), ctypes.POINTER(I)]),
                            ("arena_release", I, [P, I]), ("arena_free_bytes", S, [P]),
                            ("arena_largest_free", S, [P]), ("arena_length", S, [P, I]), ("arena_live", S, [P]),
                            ("arena_size", S, [P])):
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


class Model:
    """Brute-force free list: sorted disjoint [start, end) ranges, merged on release."""

    def __init__(self, size):
        self.size = size
        self.free = [(0, size)] if size else []

    def fits(self, n):
        return [r for r in self.free if r[1] - r[0] >= n]

    def take(self, off, n):
        for i, (a, b) in enumerate(self.free):
            if a <= off and off + n <= b:
                parts = [(a, off)] if off > a else []
                if off + n < b:
                    parts.append((off + n, b))
                self.free[i:i + 1] = parts
                return
        raise AssertionError("range not free")

    def give(self, off, n):
        self.free.append((off, off + n))
        self.free.sort()
        merged = [self.free[0]]
        for a, b in self.free[1:]:
            assert a >= merged[-1][1], "double free / overlap"
            if a == merged[-1][1]:
                merged[-1] = (merged[-1][0], b)
            else:
                merged.append((a, b))
        self.free = merged

    def free_bytes(self):
        return sum(b - a for a, b in self.free)

    def largest(self):
        return max((b - a for a, b in self.free), default=0)


@pytest.mark.parametrize("seed,size,align", [(1, 1 << 20, 256), (2, 10_000_000, 256), (3, 65536, 64), (4, 4096, 256)])
def test_arena_churn_matches_model(arena_lib, seed, size, align):
    lib = arena_lib
    rng = random.Random(seed)
    a = lib.arena_new(size, align)
    try:
        region = lib.arena_size(a)
        assert region == size // align * align
        m = Model(region)
        live = {}  # handle -> (offset, length)
        off, h = ctypes.c_size_t(), ctypes.c_int()
        counts = {0: 0, TRY_LATER: 0, DEVICE_UNABLE: 0}
        for step in range(4000):
            if live and (rng.random() < 0.45 or m.free_bytes() < region // 8):
                hd = rng.choice(sorted(live))
                o, n = live.pop(hd)
                assert lib.arena_release(a, hd) == 0
                assert lib.arena_release(a, hd) == -1  # a released handle is unknown
                m.give(o, n)
            else:
                kind = rng.random()
                if kind < 0.02:
                    req = region + rng.randint(1, 3 * align)  # can never fit
                elif kind < 0.1:
                    req = rng.randint(0, 2 * align)
                else:
                    req = int(region * rng.random() ** 3 / 2)
                n = max(1, (max(req, 1) + align - 1) // align * align) if req else align
                e = lib.arena_alloc(a, req, ctypes.byref(off), ctypes.byref(h))
                counts[e] += 1
                if n > region:
                    assert e == DEVICE_UNABLE, (step, req)
                    continue
                cands = m.fits(n)
                if not cands:
                    assert e == TRY_LATER, (step, req, m.largest())
                    continue
                assert e == 0, (step, req, e)
                o = off.value
                assert o % align == 0 and o + n <= region
                assert lib.arena_length(a, h.value) == n
                best = min(b - x for x, b in cands)
                home = [r for r in m.free if r[0] <= o < r[1]]
                assert len(home) == 1 and home[0][0] == o and home[0][1] - home[0][0] == best, "not best fit"
                for lo, ln in live.values():
                    assert o + n <= lo or lo + ln <= o, "overlap"
                m.take(o, n)
                assert h.value not in live
                live[h.value] = (o, n)
            assert lib.arena_free_bytes(a) == m.free_bytes()
            assert lib.arena_largest_free(a) == m.largest()
            assert lib.arena_live(a) == len(live)
        assert counts[0] > 500 and counts[TRY_LATER] > 0 and counts[DEVICE_UNABLE] > 0
        for hd in list(live):
            assert lib.arena_release(a, hd) == 0
        assert lib.arena_free_bytes(a) == region and lib.arena_largest_free(a) == region
        assert lib.arena_live(a) == 0
        assert lib.arena_alloc(a, region, ctypes.byref(off), ctypes.byref(h)) == 0 and off.value == 0
    finally:
        lib.arena_delete(a)
