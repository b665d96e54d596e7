"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; the HIP engine's results are compared with it.  Every
function here forwards to the C restatement in poy_oracle.c, which cites the
reference lines it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int)
u16p = C.POINTER(C.c_uint16)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the oracle's own Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "po_cm2_create": (vp, [C.c_int] * 5),
            "po_cm2_free": (None, [vp]),
            "po_cm2_cost": (C.c_int, [vp, C.c_int, C.c_int]),
            "po_cm2_worst": (C.c_int, [vp, C.c_int, C.c_int]),
            "po_cm2_median": (C.c_int, [vp, C.c_int, C.c_int]),
            "po_cm2_of_list": (vp, [i32p, C.c_int, C.c_int, C.c_int]),
            "po_cm2_of_sub_indel": (vp, [C.c_int, C.c_int]),
            "po_cm2_set_affine": (None, [vp, C.c_int, C.c_int]),
            "po_cm3_of_two_dim": (vp, [vp]),
            "po_cm3_free": (None, [vp]),
            "po_cm3_cost": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "po_cm3_median": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "po_count_gaps": (C.c_int, [u8p, C.c_int, C.c_int]),
            "po_deltaw": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp, C.c_int, C.c_int]),
            "po_algn_nw": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp, C.c_int, u16p, i32p, i32p]),
            "po_algn_nw_cells": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
            "po_backtrack_2d": (C.c_int, [u16p, i32p, i32p, u8p, C.c_int, u8p, C.c_int, vp,
                                          C.c_int, u8p, u8p, C.c_int]),
            "po_cost_2": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp, C.c_int, C.c_int]),
            "po_align_2": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp, u8p, u8p, C.c_int, i32p]),
            "po_median_2_no_gaps": (C.c_int, [u8p, u8p, C.c_int, vp, u8p, C.c_int]),
            "po_median_2_with_gaps": (C.c_int, [u8p, u8p, C.c_int, vp, u8p, C.c_int]),
            "po_ancestor_2": (C.c_int, [u8p, u8p, C.c_int, vp, u8p, C.c_int]),
            "po_union": (C.c_int, [u8p, u8p, C.c_int, u8p, C.c_int]),
            "po_union_make": (C.c_int, [u8p, u8p, u8p, C.c_int, C.c_int, u8p, i32p, C.c_int, u8p, i32p,
                                        C.c_int, u8p, i32p, i32p, i32p, C.c_int]),
            "po_verify_2": (C.c_int, [u8p, u8p, C.c_int, vp]),
            "po_worst_2": (C.c_int, [u8p, u8p, C.c_int, vp]),
            "po_align_affine_3": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp, u8p, u8p, u8p, u8p,
                                            C.c_int, i32p]),
            "po_cost_affine_3": (C.c_int, [u8p, C.c_int, u8p, C.c_int, vp]),
            "po_affine_cells": (C.c_int64, [C.c_int, C.c_int]),
            "po_align_3": (C.c_int, [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, vp, u8p, u8p, u8p,
                                     C.c_int, i32p]),
            "po_cost_3": (C.c_int, [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, vp]),
            "po_median_3": (C.c_int, [u8p, u8p, u8p, C.c_int, vp, u8p, C.c_int]),
            "po_nonadd_median": (C.c_int64, [u8p, u8p, u8p, C.c_int]),
            "po_nonadd_median_3": (None, [u8p, u8p, u8p, u8p, u8p, C.c_int]),
            "po_nonadd_median_w": (C.c_int64, [C.c_int, vp, vp, vp, C.c_int]),
            "po_nonadd_median_3_w": (None, [C.c_int, vp, vp, vp, vp, vp, C.c_int]),
            "po_nonadd_distance": (C.c_int64, [u8p, u8p, C.c_int]),
            "po_nonadd_dist_2": (C.c_int64, [u8p, u8p, u8p, C.c_int]),
            "po_nonadd_reroot_median": (None, [u8p, u8p, u8p, C.c_int]),
            "po_powell_3d": (C.c_int, [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int,
                                       C.c_int, u8p, u8p, u8p, C.c_int, i32p]),
            "po_cm2_gap_open": (C.c_int, [vp]),
            "po_powell_ukk": (C.c_int, [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, u8p, u8p, u8p, C.c_int, i32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _p8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


class CM2:
    """A two-dimensional cost matrix (src/cm.h:32-46) owned by the oracle."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("cost matrix construction failed")
        self.h = C.c_void_p(handle)

    def __del__(self):
        try:
            if self.h:
                lib().po_cm2_free(self.h)
        except Exception:
            pass

    @classmethod
    def sub_indel(cls, sub: int, indel: int, gap_open: int | None = None) -> "CM2":
        cm = cls(lib().po_cm2_of_sub_indel(sub, indel))
        if gap_open is not None:
            lib().po_cm2_set_affine(cm.h, 1, gap_open)
        return cm

    @classmethod
    def of_list(cls, mat) -> "CM2":
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.int32))
        w = m.shape[0]
        return cls(lib().po_cm2_of_list(m.ctypes.data_as(i32p), w, 1, (1 << w) - 1))

    def set_affine(self, model: int, gap_open: int) -> None:
        lib().po_cm2_set_affine(self.h, model, gap_open)

    def cost(self, a: int, b: int) -> int:
        return lib().po_cm2_cost(self.h, a, b)

    @property
    def gap_open(self) -> int:
        return lib().po_cm2_gap_open(self.h)

    def worst(self, a: int, b: int) -> int:
        return lib().po_cm2_worst(self.h, a, b)

    def median(self, a: int, b: int) -> int:
        return lib().po_cm2_median(self.h, a, b)

    def tables(self):
        """cost, median, prepend, tail, worst as 32x32 / 32 numpy arrays."""
        cost = np.zeros((32, 32), np.int32)
        med = np.zeros((32, 32), np.uint8)
        worst = np.zeros((32, 32), np.int32)
        for a in range(32):
            for b in range(32):
                cost[a, b] = self.cost(a, b)
                med[a, b] = self.median(a, b)
                worst[a, b] = self.worst(a, b)
        prepend = np.array([cost[16, x] if x else 0 for x in range(32)], np.int32)
        tail = np.array([cost[x, 16] if x else 0 for x in range(32)], np.int32)
        return cost, med, prepend, tail, worst


class CM3:
    def __init__(self, cm2: CM2):
        h = lib().po_cm3_of_two_dim(cm2.h)
        if not h:
            raise ValueError("3-D cost matrix construction failed")
        self.h = C.c_void_p(h)

    def __del__(self):
        try:
            if self.h:
                lib().po_cm3_free(self.h)
        except Exception:
            pass

    def cost(self, a, b, c):
        return lib().po_cm3_cost(self.h, a, b, c)

    def median(self, a, b, c):
        return lib().po_cm3_median(self.h, a, b, c)


def _tail(buf: np.ndarray, n: int) -> np.ndarray:
    return buf[len(buf) - n:].copy() if n > 0 else buf[:0].copy()


def deltaw(s1, s2, cm: CM2, dw: int | None = None) -> int:
    a, b = _u8(s1), _u8(s2)
    return lib().po_deltaw(_p8(a), len(a), _p8(b), len(b), cm.h, 0 if dw is None else 1,
                           0 if dw is None else dw)


def algn_nw(s1, s2, cm: CM2, deltawh: int, with_dm: bool = False):
    """algn_nw on (s1 longer-or-equal, s2); returns cost (and dm, band)."""
    a, b = _u8(s1), _u8(s2)
    if not with_dm:
        return lib().po_algn_nw(_p8(a), len(a), _p8(b), len(b), cm.h, deltawh, None, None, None)
    dm = np.zeros(len(a) * len(b), np.uint16)
    lo = np.zeros(len(a), np.int32)
    hi = np.zeros(len(a), np.int32)
    c = lib().po_algn_nw(_p8(a), len(a), _p8(b), len(b), cm.h, deltawh,
                         dm.ctypes.data_as(u16p), lo.ctypes.data_as(i32p), hi.ctypes.data_as(i32p))
    return c, dm.reshape(len(a), len(b)), lo, hi


def algn_nw_cells(l1: int, l2: int, deltawh: int) -> int:
    return lib().po_algn_nw_cells(l1, l2, deltawh)


def backtrack_2d(dm, lo, hi, s1, s2, cm: CM2, swaped: bool):
    a, b = _u8(s1), _u8(s2)
    cap = len(a) + len(b)
    r1 = np.zeros(cap, np.uint8)
    r2 = np.zeros(cap, np.uint8)
    dmf = np.ascontiguousarray(dm, dtype=np.uint16).reshape(-1)
    lo = np.ascontiguousarray(lo, dtype=np.int32)
    hi = np.ascontiguousarray(hi, dtype=np.int32)
    n = lib().po_backtrack_2d(dmf.ctypes.data_as(u16p), lo.ctypes.data_as(i32p),
                              hi.ctypes.data_as(i32p), _p8(a), len(a), _p8(b), len(b), cm.h,
                              int(bool(swaped)), _p8(r1), _p8(r2), cap)
    if n < 0:
        raise RuntimeError("backtrack left the band")
    return _tail(r1, n), _tail(r2, n)


def cost_2(s1, s2, cm: CM2, dw: int | None = None) -> int:
    a, b = _u8(s1), _u8(s2)
    return lib().po_cost_2(_p8(a), len(a), _p8(b), len(b), cm.h, 0 if dw is None else 1,
                           0 if dw is None else dw)


def align_2(s1, s2, cm: CM2):
    a, b = _u8(s1), _u8(s2)
    cap = len(a) + len(b) + 2
    o1 = np.zeros(cap, np.uint8)
    o2 = np.zeros(cap, np.uint8)
    n = C.c_int(0)
    cost = lib().po_align_2(_p8(a), len(a), _p8(b), len(b), cm.h, _p8(o1), _p8(o2), cap,
                            C.byref(n))
    if cost < 0:
        raise RuntimeError("align_2 failed")
    return cost, _tail(o1, n.value), _tail(o2, n.value)


def _two_in(fn, a, b, cm: CM2, extra: int = 1):
    x, y = _u8(a), _u8(b)
    cap = len(x) + extra
    out = np.zeros(cap, np.uint8)
    n = fn(_p8(x), _p8(y), len(x), cm.h, _p8(out), cap)
    return _tail(out, n)


def median_2(a, b, cm: CM2):
    return _two_in(lib().po_median_2_no_gaps, a, b, cm)


def median_2_with_gaps(a, b, cm: CM2):
    return _two_in(lib().po_median_2_with_gaps, a, b, cm)


def ancestor_2(a, b, cm: CM2):
    return _two_in(lib().po_ancestor_2, a, b, cm)


def union(a, b):
    x, y = _u8(a), _u8(b)
    out = np.zeros(len(x), np.uint8)
    n = lib().po_union(_p8(x), _p8(y), len(x), _p8(out), len(x))
    return _tail(out, n)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(i32p)


def union_leaf(seq):
    """Unions.leaf (src/sequence.ml:1870-1873): (items, offset, c1, c2)."""
    x = _u8(seq)
    pos = np.arange(len(x), dtype=np.int32)
    return x.copy(), np.zeros(len(x), np.int32), pos, pos.copy()


def union_make(a, b, median, gap, ua, ub):
    """union_merge (src/union.c:102-164) of the children's unions ua, ub
    (tuples whose first two entries are items and offsets) over the aligned
    pair (a, b) and its median with gaps.  Returns (items, offset, c1, c2)."""
    x, y, m = _u8(a), _u8(b), _u8(median)
    assert len(x) == len(y) == len(m)
    sa, oa = _u8(ua[0]), np.ascontiguousarray(ua[1], np.int32)
    sb, ob = _u8(ub[0]), np.ascontiguousarray(ub[1], np.int32)
    cap = len(sa) + len(sb) + 2
    out = np.zeros(cap, np.uint8)
    off, ca, cb = (np.zeros(cap, np.int32) for _ in range(3))
    n = lib().po_union_make(_p8(x), _p8(y), _p8(m), len(x), int(gap), _p8(sa), _p32(oa), len(sa), _p8(sb), _p32(ob),
                            len(sb), _p8(out), _p32(off), _p32(ca), _p32(cb), cap)
    if n < 0:
        raise ValueError("union_make: the children's unions do not match the alignment")
    return out[cap - n:].copy(), off[cap - n:].copy(), ca[cap - n:].copy(), cb[cap - n:].copy()


def verify_2(a, b, cm: CM2) -> int:
    x, y = _u8(a), _u8(b)
    return lib().po_verify_2(_p8(x), _p8(y), len(x), cm.h)


def worst_2(a, b, cm: CM2) -> int:
    x, y = _u8(a), _u8(b)
    return lib().po_worst_2(_p8(x), _p8(y), len(x), cm.h)


def align_affine_3(si, sj, cm: CM2):
    """Returns (median, resi, resj, cost, medianwg) like Sequence.Align.align_affine_3."""
    a, b = _u8(si), _u8(sj)
    cap = len(a) + len(b) + 2
    bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
    lens = np.zeros(4, np.int32)
    cost = lib().po_align_affine_3(_p8(a), len(a), _p8(b), len(b), cm.h, *[_p8(x) for x in bufs],
                                   cap, lens.ctypes.data_as(i32p))
    if cost < 0:
        raise RuntimeError("affine backtrace left the band")
    med, mwg, ri, rj = (_tail(bufs[k], int(lens[k])) for k in range(4))
    return med, ri, rj, cost, mwg


def cost_affine_3(si, sj, cm: CM2) -> int:
    a, b = _u8(si), _u8(sj)
    return lib().po_cost_affine_3(_p8(a), len(a), _p8(b), len(b), cm.h)


def affine_cells(li: int, lj: int) -> int:
    return lib().po_affine_cells(li, lj)


def align_3(s1, s2, s3, cm3: CM3):
    a, b, c = _u8(s1), _u8(s2), _u8(s3)
    cap = len(a) + len(b) + len(c)
    o = [np.zeros(cap, np.uint8) for _ in range(3)]
    n = C.c_int(0)
    cost = lib().po_align_3(_p8(a), len(a), _p8(b), len(b), _p8(c), len(c), cm3.h,
                            *[_p8(x) for x in o], cap, C.byref(n))
    return cost, _tail(o[0], n.value), _tail(o[1], n.value), _tail(o[2], n.value)


def cost_3(s1, s2, s3, cm3: CM3) -> int:
    a, b, c = _u8(s1), _u8(s2), _u8(s3)
    return lib().po_cost_3(_p8(a), len(a), _p8(b), len(b), _p8(c), len(c), cm3.h)


def median_3(a, b, c, cm3: CM3):
    x, y, z = _u8(a), _u8(b), _u8(c)
    cap = len(x) + 1
    out = np.zeros(cap, np.uint8)
    n = lib().po_median_3(_p8(x), _p8(y), _p8(z), len(x), cm3.h, _p8(out), cap)
    return _tail(out, n)


def nonadd_median(a, b):
    x, y = _u8(a), _u8(b)
    out = np.zeros(len(x), np.uint8)
    added = lib().po_nonadd_median(_p8(x), _p8(y), _p8(out), len(x))
    return out, int(added)


def nonadd_median_3(A, N, D1, D2):
    arrs = [_u8(v) for v in (A, N, D1, D2)]
    out = np.zeros(len(arrs[0]), np.uint8)
    lib().po_nonadd_median_3(*[_p8(v) for v in arrs], _p8(out), len(out))
    return out


_W_DTYPE = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def nonadd_median_w(a, b, width: int):
    """nonadd_median at 1, 2 or 4 bytes per character (NonaddCS8/16/32)."""
    t = _W_DTYPE[width]
    x, y = np.ascontiguousarray(a, t), np.ascontiguousarray(b, t)
    out = np.zeros(len(x), t)
    added = lib().po_nonadd_median_w(width, x.ctypes.data, y.ctypes.data, out.ctypes.data, len(x))
    return out, int(added)


def nonadd_median_3_w(A, N, D1, D2, width: int):
    t = _W_DTYPE[width]
    arrs = [np.ascontiguousarray(v, t) for v in (A, N, D1, D2)]
    out = np.zeros(len(arrs[0]), t)
    lib().po_nonadd_median_3_w(width, *[v.ctypes.data for v in arrs], out.ctypes.data, len(out))
    return out


def nonadd_distance(a, b) -> int:
    x, y = _u8(a), _u8(b)
    return int(lib().po_nonadd_distance(_p8(x), _p8(y), len(x)))


def nonadd_dist_2(a, b, c) -> int:
    x, y, z = _u8(a), _u8(b), _u8(c)
    return int(lib().po_nonadd_dist_2(_p8(x), _p8(y), _p8(z), len(x)))


def powell_ukk(a, b, c, mm: int, go: int, ge: int, planes_mode: int = 0):
    """The reference's own check-point Ukkonen search (src/ukk.checkp.c):
    its cost and the alignment it picks (three aligned sequences with the
    leading gap); None when a character has no base.  planes_mode 0 reuses
    cost planes as the reference does, 1 keeps every plane."""
    x, y, z = _u8(a), _u8(b), _u8(c)
    cap = len(x) + len(y) + len(z) + 2
    o = [np.zeros(cap, np.uint8) for _ in range(3)]
    n = C.c_int(0)
    cost = lib().po_powell_ukk(_p8(x), len(x), _p8(y), len(y), _p8(z), len(z), mm, go, ge, planes_mode,
                               _p8(o[0]), _p8(o[1]), _p8(o[2]), cap, C.byref(n))
    if cost < 0:
        return None
    return cost, o[0][:n.value].copy(), o[1][:n.value].copy(), o[2][:n.value].copy()


def powell_3d(a, b, c, mm: int, go: int, ge: int):
    """powell_3D_align's cost and one least-cost alignment (three aligned
    sequences with the leading gap); None when a character has no base."""
    x, y, z = _u8(a), _u8(b), _u8(c)
    cap = len(x) + len(y) + len(z) + 2
    o = [np.zeros(cap, np.uint8) for _ in range(3)]
    n = C.c_int(0)
    cost = lib().po_powell_3d(_p8(x), len(x), _p8(y), len(y), _p8(z), len(z), mm, go, ge,
                              _p8(o[0]), _p8(o[1]), _p8(o[2]), cap, C.byref(n))
    if cost < 0:
        return None
    return cost, o[0][:n.value].copy(), o[1][:n.value].copy(), o[2][:n.value].copy()
