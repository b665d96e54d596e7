"""Python host side of the engine: a ctypes binding of libpoyg.so.

Mirrors the reference's OCaml interface for the alignment hot path
(Sequence.Align in src/sequence.ml and the Cost_matrix.Two_D constructors);
every call goes through the C ABI declared in include/poyg.h and runs on the
GPU.  There is no CPU fallback: loading fails loudly when the library or a
GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpoyg.so")

POYG_MODE_LINEAR_COST = 0
POYG_MODE_LINEAR_ALIGN = 1
POYG_MODE_AFFINE_COST = 2
POYG_MODE_AFFINE_ALIGN = 3
POYG_MODE_TRIPLE = 4

_lib = None
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p


class PoygError(RuntimeError):
    pass


# exported symbols and signatures (kept in sync with include/poyg.h)
SIGNATURES = {
    "poyg_last_error": (C.c_char_p, []),
    "poyg_version": (C.c_char_p, []),
    "poyg_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "poyg_ctx_destroy": (None, [vp]),
    "poyg_ctx_sync": (C.c_int, [vp]),
    "poyg_ctx_stream": (vp, [vp]),
    "poyg_dev_alloc": (C.c_int, [vp, C.c_int64, C.POINTER(vp)]),
    "poyg_dev_free": (C.c_int, [vp, vp]),
    "poyg_dev_upload": (C.c_int, [vp, vp, vp, C.c_int64]),
    "poyg_dev_download": (C.c_int, [vp, vp, vp, C.c_int64]),
    "poyg_plan_status": (C.c_int, [vp, i32p]),
    "poyg_dev_zero": (C.c_int, [vp, vp, C.c_int64, vp]),
    "poyg_graph_begin": (C.c_int, [vp, C.POINTER(vp)]),
    "poyg_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "poyg_graph_launch": (C.c_int, [vp, vp]),
    "poyg_graph_destroy": (None, [vp]),
    "poyg_run_pairs_host": (C.c_int, [vp, vp, C.c_int, C.c_int32, u8p, C.c_int64, i64p, i32p, i64p, i32p, i32p, i32p,
                                      u8p, C.c_int64, i32p, i64p, i32p]),
    "poyg_cm_create": (C.c_int, [C.c_int] * 5 + [C.POINTER(vp)]),
    "poyg_cm_destroy": (None, [vp]),
    "poyg_cm_set_cost": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "poyg_cm_set_worst": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "poyg_cm_set_median": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "poyg_cm_set_prepend": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_set_tail": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_set_affine_flag": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_get_cost": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_get_median": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_get_worst": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_get_prepend": (C.c_int, [vp, C.c_int]),
    "poyg_cm_get_tail": (C.c_int, [vp, C.c_int]),
    "poyg_cm_get_gap": (C.c_int, [vp]),
    "poyg_cm_get_affine": (C.c_int, [vp]),
    "poyg_cm_get_gap_opening": (C.c_int, [vp]),
    "poyg_cm_of_list": (C.c_int, [C.POINTER(C.c_int), C.c_int, C.POINTER(vp)]),
    "poyg_cm_set_affine": (C.c_int, [vp, C.c_int, C.c_int]),
    "poyg_cm_of_sub_indel": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "poyg_cm3_of_two_dim": (C.c_int, [vp, C.POINTER(vp)]),
    "poyg_cm3_destroy": (None, [vp]),
    "poyg_cm3_get_cost": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "poyg_cm3_get_median": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "poyg_cm_serialized_size": (C.c_int64, [vp]),
    "poyg_cm_serialize": (C.c_int, [vp, u8p, C.c_int64]),
    "poyg_cm_deserialize": (C.c_int, [u8p, C.c_int64, C.POINTER(vp)]),
    "poyg_cm3_serialized_size": (C.c_int64, [vp]),
    "poyg_cm3_serialize": (C.c_int, [vp, u8p, C.c_int64]),
    "poyg_cm3_deserialize": (C.c_int, [u8p, C.c_int64, C.POINTER(vp)]),
    "poyg_cm_compare": (C.c_int, [vp, vp]),
    "poyg_algn_simple_2": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, C.c_int, i32p]),
    "poyg_algn_align_2d": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int, u8p, u8p,
                                     C.c_int, i32p, i32p]),
    "poyg_algn_align_affine_3": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, u8p, u8p, u8p, u8p,
                                           C.c_int, i32p, i32p]),
    "poyg_algn_cost_affine_3": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_align_3d": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, u8p, u8p, u8p,
                                     C.c_int, i32p, i32p]),
    "poyg_fitch_pairs_host": (C.c_int, [vp, u8p, u8p, C.c_int32, C.c_int32, u8p, i64p]),
    "poyg_powell_3d": (C.c_int, [vp, u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int, C.c_int,
                                 u8p, u8p, u8p, C.c_int, i32p, i32p]),
    "poyg_algn_median_2_no_gaps": (C.c_int, [vp, u8p, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_median_2_with_gaps": (C.c_int, [vp, u8p, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_ancestor_2": (C.c_int, [vp, u8p, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_union": (C.c_int, [u8p, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_median_3": (C.c_int, [vp, u8p, u8p, u8p, C.c_int, u8p, C.c_int, i32p]),
    "poyg_algn_worst_2": (C.c_int, [vp, u8p, u8p, C.c_int, i32p]),
    "poyg_algn_verify_2": (C.c_int, [vp, u8p, u8p, C.c_int, i32p]),
    "poyg_plan_pairs": (C.c_int, [vp, vp, C.c_int, C.c_int32, i64p, i32p, i64p, i32p, i32p,
                                  C.POINTER(vp)]),
    "poyg_plan_destroy": (None, [vp]),
    "poyg_plan_slots": (C.c_int, [vp]),
    "poyg_plan_out_bytes": (C.c_int64, [vp]),
    "poyg_plan_out_offsets": (C.c_int, [vp, i64p]),
    "poyg_plan_cells": (C.c_int64, [vp]),
    "poyg_plan_run": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "poyg_plan_last_times": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "poyg_plan_set_timing": (C.c_int, [vp, C.c_int]),
    "poyg_plan_triples": (C.c_int, [vp, vp, C.c_int32, i64p, i32p, i64p, i32p, i64p, i32p,
                                    C.POINTER(vp)]),
    "poyg_fitch_down": (C.c_int, [vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]),
    "poyg_fitch_up": (C.c_int, [vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp]),
    "poyg_nonadd_median": (C.c_int, [vp, u8p, u8p, u8p, C.c_int, i64p]),
    "poyg_nonadd_median_3": (C.c_int, [vp, u8p, u8p, u8p, u8p, u8p, C.c_int]),
    "poyg_ctx_device": (C.c_int, [vp]),
    "poyg_ctx_set_tuning": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "poyg_ctx_get_tuning": (C.c_int, [vp, C.c_char_p]),
    "poyg_ctx_device_name": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "poyg_fitch_down_w": (C.c_int, [vp, C.c_int, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]),
    "poyg_fitch_up_w": (C.c_int, [vp, C.c_int, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp]),
    "poyg_nonadd_median_w": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, i64p]),
    "poyg_nonadd_median_3_w": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "poyg_seq_stats": (C.c_int, [vp, vp, vp, vp, C.c_int32, vp, vp]),
    "poyg_union_make": (C.c_int, [vp, u8p, u8p, u8p, C.c_int, u8p, i32p, C.c_int, u8p, i32p, C.c_int, u8p, i32p,
                                  i32p, i32p, C.c_int, i32p]),
    "poyg_union_merge": (C.c_int, [vp, vp, C.c_int32, vp, vp, vp]),
    "poyg_algn_direction_matrix": (C.c_int, [vp, vp, u8p, C.c_int, u8p, C.c_int, C.c_int, vp, C.c_int64, i32p]),
    "poyg_mlp_param_count": (C.c_int64, []),
    "poyg_mlp_create": (C.c_int, [vp, C.c_int, C.c_uint64, C.POINTER(vp)]),
    "poyg_mlp_destroy": (None, [vp]),
    "poyg_mlp_set_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64]),
    "poyg_mlp_get_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64]),
    "poyg_mlp_forward": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
    "poyg_mlp_grad": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_float)]),
    "poyg_mlp_get_grads": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64]),
    "poyg_mlp_adam": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
    "poyg_kmer_counts": (C.c_int, [vp, vp, vp, vp, C.c_int32, vp, vp]),
    "poyg_pair_features": (C.c_int, [vp, vp, vp, vp, C.c_int64, C.c_float, vp, vp]),
    "poyg_noalloc_scratch_bytes": (C.c_int64, []),
    "poyg_pairs_cost_linear_noalloc": (C.c_int, [vp, vp, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]),
    "poyg_gather_rows": (C.c_int, [vp, vp, C.c_int64, vp, C.c_int64, vp, vp]),
    "poyg_gemm_bf16_f32": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int64, C.c_int64, vp, C.c_int64,
                                     C.c_int64, vp, C.c_int64, vp]),
}


def load(path: str = LIB_PATH):
    """Load libpoyg.so (no GPU needed); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise PoygError(f"{path} is missing: build it with `python -m poyd_amd.build`")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().poyg_last_error().decode(errors="replace")
        raise PoygError(f"{what} failed ({rc}): {msg}")


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _p8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device: int = 0):
        h = vp()
        _check(load().poyg_ctx_create(device, C.byref(h)), "poyg_ctx_create")
        self.h = h
        self.device = device

    def sync(self):
        _check(load().poyg_ctx_sync(self.h), "poyg_ctx_sync")

    def set_tuning(self, knob: str, value: int) -> None:
        """A launch-shape knob of later plans (poyg_ctx_set_tuning); -1
        restores the default."""
        _check(load().poyg_ctx_set_tuning(self.h, knob.encode(), int(value)), "poyg_ctx_set_tuning")

    def tuning(self, knob: str) -> int:
        return int(load().poyg_ctx_get_tuning(self.h, knob.encode()))

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        _check(load().poyg_ctx_device_name(self.h, buf, 256), "poyg_ctx_device_name")
        return buf.value.decode()

    def close(self):
        if self.h:
            load().poyg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation made through the engine (hipMalloc)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = vp()
        _check(load().poyg_dev_alloc(ctx.h, self.nbytes, C.byref(h)), "poyg_dev_alloc")
        self.ptr = h.value

    @classmethod
    def from_array(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(ctx, max(a.nbytes, 16))
        b.upload(a)
        return b

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _check(load().poyg_dev_upload(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "poyg_dev_upload")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(load().poyg_dev_download(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "poyg_dev_download")
        return out

    def free(self):
        if self.ptr:
            load().poyg_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CostMatrix:
    """Cost_matrix.Two_D.m (src/cost_matrix.ml), held by the engine."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def of_sub_indel(cls, sub: int, indel: int, gap_open: int | None = None) -> "CostMatrix":
        h = vp()
        _check(load().poyg_cm_of_sub_indel(sub, indel, -1 if gap_open is None else gap_open, C.byref(h)),
               "poyg_cm_of_sub_indel")
        return cls(h)

    @classmethod
    def of_list(cls, mat) -> "CostMatrix":
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.int32))
        h = vp()
        _check(load().poyg_cm_of_list(m.ctypes.data_as(C.POINTER(C.c_int)), m.shape[0], C.byref(h)),
               "poyg_cm_of_list")
        return cls(h)

    def serialize(self) -> bytes:
        """cm_CAML_serialize's byte stream (src/cm.c:909-932)."""
        n = load().poyg_cm_serialized_size(self.h)
        buf = np.zeros(n, np.uint8)
        _check(load().poyg_cm_serialize(self.h, _p8(buf), n), "poyg_cm_serialize")
        return buf.tobytes()

    @classmethod
    def deserialize(cls, data: bytes) -> "CostMatrix":
        buf = np.frombuffer(data, np.uint8).copy()
        h = vp()
        _check(load().poyg_cm_deserialize(_p8(buf), len(buf), C.byref(h)), "poyg_cm_deserialize")
        return cls(h)

    def compare(self, other: "CostMatrix") -> int:
        return load().poyg_cm_compare(self.h, other.h)

    def set_affine(self, model: int, gap_open: int) -> None:
        _check(load().poyg_cm_set_affine(self.h, model, gap_open), "poyg_cm_set_affine")

    def cost(self, a, b):
        return load().poyg_cm_get_cost(self.h, a, b)

    def median(self, a, b):
        return load().poyg_cm_get_median(self.h, a, b)

    def worst(self, a, b):
        return load().poyg_cm_get_worst(self.h, a, b)

    def prepend(self, a):
        return load().poyg_cm_get_prepend(self.h, a)

    def tail(self, a):
        return load().poyg_cm_get_tail(self.h, a)

    @property
    def affine(self) -> bool:
        return load().poyg_cm_get_affine(self.h) == 1

    @property
    def gap_opening(self) -> int:
        return load().poyg_cm_get_gap_opening(self.h)

    def __del__(self):
        try:
            if self.h:
                load().poyg_cm_destroy(self.h)
        except Exception:
            pass


class CostMatrix3:
    def __init__(self, cm2: CostMatrix):
        h = vp()
        _check(load().poyg_cm3_of_two_dim(cm2.h, C.byref(h)), "poyg_cm3_of_two_dim")
        self.h = h

    def cost(self, a, b, c):
        return load().poyg_cm3_get_cost(self.h, a, b, c)

    def median(self, a, b, c):
        return load().poyg_cm3_get_median(self.h, a, b, c)

    def serialize(self) -> bytes:
        """cm_CAML_serialize_3d's byte stream (src/cm.c:934-952)."""
        n = load().poyg_cm3_serialized_size(self.h)
        buf = np.zeros(n, np.uint8)
        _check(load().poyg_cm3_serialize(self.h, _p8(buf), n), "poyg_cm3_serialize")
        return buf.tobytes()

    @classmethod
    def deserialize(cls, data: bytes) -> "CostMatrix3":
        buf = np.frombuffer(data, np.uint8).copy()
        h = vp()
        _check(load().poyg_cm3_deserialize(_p8(buf), len(buf), C.byref(h)), "poyg_cm3_deserialize")
        obj = cls.__new__(cls)
        obj.h = h
        return obj

    def __del__(self):
        try:
            if self.h:
                load().poyg_cm3_destroy(self.h)
        except Exception:
            pass


def _tail(buf, n):
    return buf[len(buf) - n:].copy()


def count_gaps(s, gap: int = 16) -> int:
    """seq_CAML_count (src/seq.c:571-581): positions carrying the gap bit."""
    return int((np.asarray(s) & gap).astype(bool).sum())


def deltaw(s1, s2, dw: int | None = None, gap: int = 16) -> int:
    """Band parameter of Sequence.Align.cost_2 (src/sequence.ml:691-713)."""
    l1, l2 = len(s1), len(s2)
    gaps = max(count_gaps(s1, gap), count_gaps(s2, gap))
    longer, shorter = (l1, l2) if l1 >= l2 else (l2, l1)
    dif = longer - shorter
    lower = int(float(longer) * 0.10)
    if dw is None:
        extra = lower // 2 if dif < lower else 2
    else:
        extra = lower if dif < lower else dw
    return gaps + extra


class Align:
    """Per-call entry points with the semantics of Sequence.Align
    (src/sequence.ml:453-1141); each call is a one-pair GPU batch."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def cost_2(self, s1, s2, cm: CostMatrix, dw: int | None = None) -> int:
        a, b = _u8(s1), _u8(s2)
        out = C.c_int32(0)
        if cm.affine:
            _check(load().poyg_algn_cost_affine_3(self.ctx.h, cm.h, _p8(a), len(a), _p8(b), len(b),
                                                  C.byref(out)), "cost_affine_3")
            return out.value
        d = deltaw(a, b, dw)
        if len(a) < len(b):
            a, b = b, a
        _check(load().poyg_algn_simple_2(self.ctx.h, cm.h, _p8(a), len(a), _p8(b), len(b), d, C.byref(out)),
               "algn_simple_2")
        return out.value

    def align_2(self, s1, s2, cm: CostMatrix):
        """align_2 ~first_gap:true: (aligned s1, aligned s2, cost)."""
        if cm.affine:
            med, ri, rj, cost, mwg = self.align_affine_3(s1, s2, cm)
            return ri, rj, cost
        a, b = _u8(s1), _u8(s2)
        d = deltaw(a, b)
        swaped = len(a) >= len(b)
        x, y = (a, b) if swaped else (b, a)
        cap = len(a) + len(b) + 2
        r1 = np.zeros(cap, np.uint8)
        r2 = np.zeros(cap, np.uint8)
        n = C.c_int32(0)
        cost = C.c_int32(0)
        # the stub receives the longer sequence first and the swap flag
        _check(load().poyg_algn_align_2d(self.ctx.h, cm.h, _p8(x), len(x), _p8(y), len(y), d, int(swaped),
                                         _p8(r1), _p8(r2), cap, C.byref(n), C.byref(cost)), "algn_align_2d")
        o1, o2 = _tail(r1, n.value), _tail(r2, n.value)
        return (o1, o2, cost.value) if swaped else (o2, o1, cost.value)

    def align_affine_3(self, si, sj, cm: CostMatrix):
        """(median, resi, resj, cost, medianwg) as Sequence.Align.align_affine_3."""
        a, b = _u8(si), _u8(sj)
        cap = len(a) + len(b) + 2
        bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
        lens = np.zeros(4, np.int32)
        cost = C.c_int32(0)
        _check(load().poyg_algn_align_affine_3(self.ctx.h, cm.h, _p8(a), len(a), _p8(b), len(b),
                                               *[_p8(x) for x in bufs], cap, lens.ctypes.data_as(i32p),
                                               C.byref(cost)), "align_affine_3")
        med, mwg, ri, rj = (_tail(bufs[k], int(lens[k])) for k in range(4))
        return med, ri, rj, cost.value, mwg

    def direction_matrix(self, s1, s2, cm: CostMatrix, deltawh: int):
        """(cost, l1 x l2 uint16 direction matrix) of algn_fill_plane_2 on
        (s1 longer or equal, s2) with band parameter deltawh, as the readers
        of Matrix.default see it (0 outside the band)."""
        a, b = _u8(s1), _u8(s2)
        dm = np.zeros(len(a) * len(b), np.uint16)
        cost = C.c_int32(0)
        _check(load().poyg_algn_direction_matrix(self.ctx.h, cm.h, _p8(a), len(a), _p8(b), len(b), int(deltawh),
                                                 dm.ctypes.data, dm.size, C.byref(cost)), "algn_direction_matrix")
        return cost.value, dm.reshape(len(a), len(b))

    def median_2(self, a, b, cm: CostMatrix):
        x, y = _u8(a), _u8(b)
        out = np.zeros(len(x) + 1, np.uint8)
        n = C.c_int32(0)
        _check(load().poyg_algn_median_2_no_gaps(cm.h, _p8(x), _p8(y), len(x), _p8(out), len(out), C.byref(n)),
               "median_2")
        return _tail(out, n.value)


@dataclass
class PairResult:
    cost: np.ndarray
    out: np.ndarray | None
    out_len: np.ndarray | None
    out_off: np.ndarray | None
    cap: np.ndarray | None
    slots: int
    cells: int = 0          # DP cells the plan evaluated

    def seq(self, k: int, slot: int) -> np.ndarray:
        cap = int(self.cap[k])
        base = int(self.out_off[k]) + slot * cap
        n = int(self.out_len[k * self.slots + slot])
        return self.out[base + cap - n: base + cap]

    def seqs(self, slot: int) -> list:
        """seq(k, slot) of every pair."""
        end = (self.out_off + (slot + 1) * self.cap).tolist()
        n = self.out_len[slot::self.slots].tolist()
        out = self.out
        return [out[e - m: e] for e, m in zip(end, n)]


class PairPlan:
    """A batched pairwise plan over device-resident sequences (torch tensors
    on the GPU are used as plain device buffers)."""

    def __init__(self, ctx: Context, cm: CostMatrix, mode: int, off_a, len_a, off_b, len_b, dws=None):
        self.ctx = ctx
        self.cm = cm
        self.mode = mode
        self.n = len(len_a)
        self._keep = [np.ascontiguousarray(x, dtype=t) for x, t in
                      ((off_a, np.int64), (len_a, np.int32), (off_b, np.int64), (len_b, np.int32))]
        oa, la, ob, lb = self._keep
        d = None
        if dws is not None:
            d = np.ascontiguousarray(dws, dtype=np.int32)
            self._keep.append(d)
        h = vp()
        _check(load().poyg_plan_pairs(ctx.h, cm.h, mode, self.n, oa.ctypes.data_as(i64p), la.ctypes.data_as(i32p),
                                      ob.ctypes.data_as(i64p), lb.ctypes.data_as(i32p),
                                      d.ctypes.data_as(i32p) if d is not None else None, C.byref(h)),
               "poyg_plan_pairs")
        self.h = h
        self.slots = load().poyg_plan_slots(h)
        self.out_bytes = load().poyg_plan_out_bytes(h)
        self.cells = load().poyg_plan_cells(h)
        self.out_off = np.zeros(self.n, np.int64)
        if self.n:
            load().poyg_plan_out_offsets(h, self.out_off.ctypes.data_as(i64p))
        self.cap = ((la.astype(np.int64) + lb + 2 + 15) // 16) * 16

    def set_timing(self, on: bool = True):
        _check(load().poyg_plan_set_timing(self.h, int(on)), "set_timing")

    def last_times(self):
        f, t = C.c_float(0), C.c_float(0)
        load().poyg_plan_last_times(self.h, C.byref(f), C.byref(t))
        return f.value, t.value

    def status(self) -> int:
        """Pairs of the last run that failed (poyg_plan_status)."""
        bad = C.c_int32(0)
        _check(load().poyg_plan_status(self.h, C.byref(bad)), "poyg_plan_status")
        return int(bad.value)

    def run(self, d_seq, d_cost, d_out=None, d_out_len=None, stream=None) -> None:
        """d_* are device pointers (ints) of caller-owned buffers."""
        _check(load().poyg_plan_run(self.h, d_seq, d_cost, d_out or None, d_out_len or None, stream),
               "poyg_plan_run")

    def close(self):
        if getattr(self, "h", None):
            load().poyg_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TriplePlan(PairPlan):
    """A batched three-sequence plan (align_3: cost, three aligned sequences
    and the median) over device-resident sequences."""

    def __init__(self, ctx: Context, cm3: CostMatrix3, off_1, len_1, off_2, len_2, off_3, len_3):
        self.ctx = ctx
        self.cm = cm3
        self.mode = POYG_MODE_TRIPLE
        self.n = len(len_1)
        self._keep = [np.ascontiguousarray(x, dtype=t) for x, t in
                      ((off_1, np.int64), (len_1, np.int32), (off_2, np.int64), (len_2, np.int32),
                       (off_3, np.int64), (len_3, np.int32))]
        o1, l1, o2, l2, o3, l3 = self._keep
        h = vp()
        _check(load().poyg_plan_triples(ctx.h, cm3.h, self.n, o1.ctypes.data_as(i64p), l1.ctypes.data_as(i32p),
                                        o2.ctypes.data_as(i64p), l2.ctypes.data_as(i32p),
                                        o3.ctypes.data_as(i64p), l3.ctypes.data_as(i32p), C.byref(h)),
               "poyg_plan_triples")
        self.h = h
        self.slots = load().poyg_plan_slots(h)
        self.out_bytes = load().poyg_plan_out_bytes(h)
        self.cells = load().poyg_plan_cells(h)
        self.out_off = np.zeros(self.n, np.int64)
        if self.n:
            load().poyg_plan_out_offsets(h, self.out_off.ctypes.data_as(i64p))
        self.cap = ((l1.astype(np.int64) + l2 + l3 + 1 + 15) // 16) * 16


class UnionJob(C.Structure):
    """poyg_union_job: one merge of a batch (device pointers)."""
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("median", C.c_void_p),
                ("ua_seq", C.c_void_p), ("ub_seq", C.c_void_p), ("ua_end", C.c_void_p), ("ub_end", C.c_void_p),
                ("out_seq", C.c_void_p), ("out_end", C.c_void_p), ("out_ca", C.c_void_p), ("out_cb", C.c_void_p),
                ("len", C.c_int32), ("na", C.c_int32), ("nb", C.c_int32), ("cap", C.c_int32)]


def union_leaf(seq):
    """Unions.leaf (src/sequence.ml:1870-1873): (items, end, ca, cb)."""
    x = _u8(seq).copy()
    pos = np.arange(len(x), dtype=np.int32)
    return x, np.zeros(len(x), np.int32), pos, pos.copy()


def union_make(cm: "CostMatrix", a, b, median, ua, ub):
    """union_CAML_make (src/union.c:209-226) on the host: the union of the
    children's unions ua, ub ((items, end, ...) tuples) over the aligned pair
    (a, b) and its median with gaps (None: the matrix's column medians).
    Returns (items, end, ca, cb)."""
    x, y = _u8(a), _u8(b)
    m = None if median is None else _u8(median)
    sa, ea = _u8(ua[0]), np.ascontiguousarray(ua[1], np.int32)
    sb, eb = _u8(ub[0]), np.ascontiguousarray(ub[1], np.int32)
    cap = len(sa) + len(sb) + 2
    out = np.zeros(cap, np.uint8)
    end, ca, cb = (np.zeros(cap, np.int32) for _ in range(3))
    n = C.c_int32(0)
    _check(load().poyg_union_make(cm.h, _p8(x), _p8(y), None if m is None else _p8(m), len(x), _p8(sa),
                                  ea.ctypes.data_as(i32p), len(sa), _p8(sb), eb.ctypes.data_as(i32p), len(sb),
                                  _p8(out), end.ctypes.data_as(i32p), ca.ctypes.data_as(i32p),
                                  cb.ctypes.data_as(i32p), cap, C.byref(n)), "poyg_union_make")
    k = cap - n.value
    return out[k:].copy(), end[k:].copy(), ca[k:].copy(), cb[k:].copy()


def union_merge(ctx: "Context", cm: "CostMatrix", jobs, stream=None) -> np.ndarray:
    """poyg_union_merge over a list of UnionJob (device pointers); returns
    the union lengths (negative: the children did not match)."""
    n = len(jobs)
    if n == 0:
        return np.zeros(0, np.int32)
    arr = (UnionJob * n)(*jobs)
    d_jobs = DeviceBuffer(ctx, C.sizeof(arr))
    d_n = DeviceBuffer(ctx, 4 * n)
    try:
        _check(load().poyg_dev_upload(ctx.h, d_jobs.ptr, C.addressof(arr), C.sizeof(arr)), "poyg_dev_upload")
        _check(load().poyg_union_merge(ctx.h, cm.h, n, d_jobs.ptr, d_n.ptr, stream), "poyg_union_merge")
        ctx.sync()
        return d_n.download(np.int32, n)
    finally:
        d_jobs.free()
        d_n.free()


def powell_3d(ctx: Context, a, b, c, mm: int, go: int, ge: int):
    """powell_3D_align (src/ukkCommon.c:110-151): (cost, three aligned
    sequences with the leading gap)."""
    x, y, z = _u8(a), _u8(b), _u8(c)
    cap = len(x) + len(y) + len(z)
    out = [np.zeros(cap, np.uint8) for _ in range(3)]
    n, cost = C.c_int32(0), C.c_int32(0)
    _check(load().poyg_powell_3d(ctx.h, _p8(x), len(x), _p8(y), len(y), _p8(z), len(z), mm, go, ge,
                                 _p8(out[0]), _p8(out[1]), _p8(out[2]), cap, C.byref(n), C.byref(cost)),
           "poyg_powell_3d")
    return int(cost.value), [o[cap - n.value:].copy() for o in out]


def fitch_pairs(ctx: Context, a: np.ndarray, b: np.ndarray):
    """nonadd_median (src/nonaddCSc.c:422-462) of row pairs on the GPU: a and
    b are (n, len) u8 state sets; returns (medians (n, len), changes (n,))."""
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    n, m = a.shape
    if n == 0:
        return np.zeros((0, m), np.uint8), np.zeros(0, np.int64)
    # one call: one upload, the kernel, one download, one synchronization
    res = np.empty((n, m), np.uint8)
    added = np.empty(n, np.int64)
    _check(load().poyg_fitch_pairs_host(ctx.h, _p8(a), _p8(b), n, m, _p8(res), added.ctypes.data_as(i64p)),
           "poyg_fitch_pairs_host")
    return res, added


def batch_deltaw(a_batch, b_batch, dw=None) -> np.ndarray:
    """deltaw of Sequence.Align.cost_2 for every pair, vectorised; dw (one
    value per pair, or None) is the ?deltaw argument, as in deltaw()."""
    la = a_batch.ln.astype(np.int64)
    lb = b_batch.ln.astype(np.int64)
    ga = _gaps_per_item(a_batch)
    gb = _gaps_per_item(b_batch)
    longer = np.maximum(la, lb)
    dif = np.abs(la - lb)
    lower = np.floor(longer.astype(np.float64) * 0.10).astype(np.int64)
    if dw is None:
        extra = np.where(dif < lower, lower // 2, 2)
    else:
        extra = np.where(dif < lower, lower, np.asarray(dw, np.int64))
    return (np.maximum(ga, gb) + extra).astype(np.int32)


def _gaps_per_item(batch) -> np.ndarray:
    """count_gaps of every sequence of a SeqBatch (empty ones count 0)."""
    n = len(batch.ln)
    if n == 0:
        return np.zeros(0, np.int64)
    gap = (batch.seq & 16) != 0
    if (batch.ln > 0).all():   # reduceat sums [off[k], off[k+1]) when every item is non-empty
        return np.add.reduceat(gap, batch.off).astype(np.int64)
    pre = np.concatenate([[0], np.cumsum(gap, dtype=np.int64)])
    off = batch.off.astype(np.int64)
    return pre[off + batch.ln] - pre[off]


class Graph:
    """A recorded pass (poyg_graph_*): the launches made with the capture
    stream between record() and end(), replayed by launch()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None
        s = vp()
        _check(load().poyg_graph_begin(ctx.h, C.byref(s)), "poyg_graph_begin")
        self.stream = s.value   # pass this stream to the calls to record

    def end(self) -> "Graph":
        g = vp()
        _check(load().poyg_graph_end(self.ctx.h, C.byref(g)), "poyg_graph_end")
        self.h = g.value
        return self

    def launch(self, stream=None) -> None:
        _check(load().poyg_graph_launch(self.h, stream), "poyg_graph_launch")

    def close(self) -> None:
        if self.h:
            load().poyg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dev_zero(ctx: Context, ptr: int, nbytes: int, stream=None) -> None:
    """Zero nbytes (a multiple of 4) at a device address with a kernel."""
    _check(load().poyg_dev_zero(ctx.h, ptr, int(nbytes), stream), "poyg_dev_zero")


def run_pairs(ctx: Context, cm: CostMatrix, mode: int, a_batch, b_batch, dws=None,
              check_band: bool = True) -> PairResult:
    """Convenience: run two SeqBatch-like objects (seq/off/ln, host memory)
    as one batch through poyg_run_pairs_host (one upload, one
    synchronization) and return every result."""
    seq = np.ascontiguousarray(np.concatenate([a_batch.seq, b_batch.seq]), dtype=np.uint8)
    off_a = np.ascontiguousarray(a_batch.off, dtype=np.int64)
    off_b = np.ascontiguousarray(np.asarray(b_batch.off, np.int64) + len(a_batch.seq))
    la = np.ascontiguousarray(a_batch.ln, dtype=np.int32)
    lb = np.ascontiguousarray(b_batch.ln, dtype=np.int32)
    return _run_host(ctx, cm, mode, seq, off_a, la, off_b, lb, dws, check_band)


def run_pair_list(ctx: Context, cm: CostMatrix, mode: int, pairs, dws=None, check_band: bool = True) -> PairResult:
    """run_pairs over a list of (a, b) host sequences: both sides of every
    pair in one concatenation, in pair order."""
    flat = [x for p in pairs for x in p]
    ln = np.fromiter((len(x) for x in flat), np.int64, len(flat))
    off = np.zeros(len(flat), np.int64)
    if len(flat) > 1:
        np.cumsum(ln[:-1], out=off[1:])
    seq = np.concatenate(flat).astype(np.uint8, copy=False) if flat else np.zeros(0, np.uint8)
    return _run_host(ctx, cm, mode, seq, off[0::2].copy(), ln[0::2].astype(np.int32), off[1::2].copy(),
                     ln[1::2].astype(np.int32), dws, check_band)


def _run_host(ctx, cm, mode, seq, off_a, la, off_b, lb, dws, check_band) -> PairResult:
    n = len(la)
    # (linear modes without dws: the library computes the default band)
    dw = None if dws is None else np.ascontiguousarray(dws, dtype=np.int32)
    slots = {POYG_MODE_LINEAR_ALIGN: 3, POYG_MODE_AFFINE_ALIGN: 4}.get(mode, 0)
    # the plan's output layout: slots of round16(la + lb + 2) bytes per pair
    cap = ((la.astype(np.int64) + lb + 2 + 15) // 16) * 16
    out_off = np.zeros(n, np.int64)
    if n > 1:
        out_off[1:] = np.cumsum(slots * cap[:-1])
    out_bytes = int(slots * cap.sum()) if slots else 0
    cost = np.zeros(n, np.int32)
    out = np.empty(max(out_bytes, 16), np.uint8) if slots else None
    out_len = np.zeros(max(n * slots, 1), np.int32) if slots else None
    cells = C.c_int64(0)
    failed = C.c_int32(0)
    _check(load().poyg_run_pairs_host(ctx.h, cm.h, mode, n, _p8(seq), len(seq), off_a.ctypes.data_as(i64p),
                                      la.ctypes.data_as(i32p), off_b.ctypes.data_as(i64p), lb.ctypes.data_as(i32p),
                                      None if dw is None else dw.ctypes.data_as(i32p), cost.ctypes.data_as(i32p),
                                      None if out is None else _p8(out), out_bytes,
                                      None if out_len is None else out_len.ctypes.data_as(i32p), C.byref(cells),
                                      C.byref(failed)), "poyg_run_pairs_host")
    if check_band and failed.value:
        raise PoygError(f"{failed.value} pairs failed (a backtrace left the filled band, "
                        "or a fill's wait for the previous strip gave up)")
    return PairResult(cost, out, out_len, out_off if slots else None, cap if slots else None, slots,
                      int(cells.value))
