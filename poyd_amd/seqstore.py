"""Device-resident sequences for batched tree passes.

Every sequence lives in one device arena: leaves are uploaded once, and the
medians of a batch are written there by the alignment plan itself, so that a
later batch aligns them in place.  The host keeps per sequence only its
offset, length, gap count and emptiness (poyg_seq_stats), which is all that
planning the next batch needs:

* DOS.median (src/seqCS.ml:747-776): an empty side passes the other through
  at cost 0; otherwise linear matrices align with Sequence.Align.align_2 and
  keep the median (ancestor_2 / median_2 of the aligned pair), affine ones
  take align_affine_3's median; the band parameter is cost_2's default
  (src/sequence.ml:691-713).
* DOS.distance (src/seqCS.ml:818-864): 0 when a side is empty; linear
  matrices run cost_2 with ?deltaw = max(8, length difference), affine ones
  cost_2_affine.

Operations take and return integer sequence ids (numpy arrays), so a tree
pass composes them with vectorised index arithmetic.  OracleStore (in the
tests) implements the same interface on the CPU oracle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import engine

GAP = 16


def band_deltaw(la, lb, gaps_a, gaps_b, dw=None) -> np.ndarray:
    """deltaw of Sequence.Align.cost_2 (src/sequence.ml:691-713) from the
    lengths and gap counts: max gap count plus, when the length difference
    is below 10% of the longer length, that 10% (halved without ?deltaw),
    else 2 (or the given ?deltaw)."""
    la = np.asarray(la, np.int64)
    lb = np.asarray(lb, np.int64)
    longer = np.maximum(la, lb)
    dif = np.abs(la - lb)
    lower = np.floor(longer.astype(np.float64) * 0.10).astype(np.int64)
    if dw is None:
        extra = np.where(dif < lower, lower // 2, 2)
    else:
        extra = np.where(dif < lower, lower, np.asarray(dw, np.int64))
    return (np.maximum(np.asarray(gaps_a, np.int64), np.asarray(gaps_b, np.int64)) + extra).astype(np.int32)


class Aligned:
    """Arena offsets of the alignments behind a batch of medians: per pair
    the aligned children (a_off, b_off, `length` columns) and the median
    with gaps (mwg_off, -1 where the kernel takes the matrix's column
    medians); length -1 where no alignment ran (an empty side)."""

    def __init__(self, a_off, b_off, mwg_off, length):
        self.a_off, self.b_off, self.mwg_off, self.length = a_off, b_off, mwg_off, length

    @classmethod
    def none(cls, n: int) -> "Aligned":
        return cls(np.full(n, -1, np.int64), np.full(n, -1, np.int64), np.full(n, -1, np.int64),
                   np.full(n, -1, np.int32))


class DeviceStore:
    """Sequences in one device arena, aligned by libpoyg plans."""

    # pairs per plan: bounds a plan's direction bands (linear bands of the
    # tree workloads are a few hundred KB per pair)
    MEDIAN_CHUNK = 24576
    DISTANCE_CHUNK = 1 << 19

    def __init__(self, ctx: engine.Context, cm: engine.CostMatrix, capacity: int):
        self.ctx, self.cm = ctx, cm
        self.affine = cm.affine
        self.arena = engine.DeviceBuffer(ctx, capacity)
        self.capacity = int(capacity)
        self.top = 0
        self.off = np.zeros(0, np.int64)
        self.ln = np.zeros(0, np.int32)
        self.gaps = np.zeros(0, np.int32)
        self.empty = np.zeros(0, bool)
        self.cells = 0          # DP cells of every alignment run so far
        self.alignments = 0     # pairwise alignments run so far

    # ------------------------------------------------------------ arena
    def __len__(self) -> int:
        return len(self.ln)

    def _reserve(self, nbytes: int) -> int:
        at = (self.top + 255) & ~255
        if at + nbytes > self.capacity:
            raise engine.PoygError(f"sequence arena full ({self.capacity} bytes)")
        self.top = at + int(nbytes)
        return at

    def mark(self):
        return (len(self.ln), self.top)

    def release(self, mark) -> None:
        """Forget every sequence added after `mark` (their ids become free)."""
        n, top = mark
        self.off, self.ln = self.off[:n], self.ln[:n]
        self.gaps, self.empty = self.gaps[:n], self.empty[:n]
        self.top = top

    def add(self, seqs) -> np.ndarray:
        """Upload host sequences; returns their ids."""
        seqs = [np.ascontiguousarray(s, dtype=np.uint8) for s in seqs]
        ln = np.array([len(s) for s in seqs], np.int32)
        if len(seqs) == 0:
            return np.zeros(0, np.int64)
        blob = np.concatenate(seqs)
        at = self._reserve(len(blob))
        engine._check(engine.load().poyg_dev_upload(self.ctx.h, self.arena.ptr + at, blob.ctypes.data, len(blob)),
                      "poyg_dev_upload")
        off = at + np.concatenate([[0], np.cumsum(ln[:-1], dtype=np.int64)]).astype(np.int64)
        gaps = np.array([int(((s & GAP) != 0).sum()) for s in seqs], np.int32)
        empty = np.array([bool((s == GAP).all()) for s in seqs], bool)
        return self._append(off, ln, gaps, empty)

    def _append(self, off, ln, gaps, empty) -> np.ndarray:
        first = len(self.ln)
        self.off = np.concatenate([self.off, off])
        self.ln = np.concatenate([self.ln, ln])
        self.gaps = np.concatenate([self.gaps, gaps])
        self.empty = np.concatenate([self.empty, empty])
        return np.arange(first, len(self.ln), dtype=np.int64)

    def get(self, i: int) -> np.ndarray:
        n = int(self.ln[i])
        out = np.empty(n, np.uint8)
        engine._check(engine.load().poyg_dev_download(self.ctx.h, out.ctypes.data, self.arena.ptr + int(self.off[i]),
                                                      n), "poyg_dev_download")
        return out

    def _stats(self, off: np.ndarray, ln: np.ndarray):
        """Gap counts and emptiness of arena sequences (on the device)."""
        n = len(ln)
        if n == 0:
            return np.zeros(0, np.int32), np.zeros(0, bool)
        meta = np.zeros(3 * n, np.int64)   # offsets, then lengths (int32) and room for the stats
        meta[:n] = off
        lens = meta[n:].view(np.int32)
        lens[:n] = ln
        d = engine.DeviceBuffer.from_array(self.ctx, meta)
        st = d.ptr + 8 * n + 4 * n
        engine._check(engine.load().poyg_seq_stats(self.ctx.h, self.arena.ptr, d.ptr, d.ptr + 8 * n, n, st, None),
                      "poyg_seq_stats")
        back = np.empty(2 * n, np.int32)
        engine._check(engine.load().poyg_dev_download(self.ctx.h, back.ctypes.data, st, 8 * n), "poyg_dev_download")
        d.free()
        return back[0::2].copy(), back[1::2].astype(bool)

    # ------------------------------------------------------- alignments
    def _run(self, mode: int, a: np.ndarray, b: np.ndarray, dws):
        plan = engine.PairPlan(self.ctx, self.cm, mode, self.off[a], self.ln[a], self.off[b], self.ln[b], dws)
        d_small = None
        try:
            n = plan.n
            n_cost = (max(n, 1) + 63) // 64 * 64
            n_len = n * plan.slots
            d_small = engine.DeviceBuffer(self.ctx, 4 * (n_cost + max(n_len, 1)))
            out_at = self._reserve(plan.out_bytes) if plan.slots else 0
            plan.run(self.arena.ptr, d_small.ptr, self.arena.ptr + out_at if plan.slots else None,
                     d_small.ptr + 4 * n_cost if plan.slots else None)
            small = d_small.download(np.int32, n_cost + n_len)
            bad = plan.status()
            if bad:
                raise engine.PoygError(f"{bad} pairs failed (a backtrace left the filled band, "
                                       "or a fill's wait for the previous strip gave up)")
            self.cells += int(plan.cells)
            self.alignments += n
            lens = small[n_cost:n_cost + n_len].reshape(n, plan.slots) if plan.slots else None
            return small[:n].astype(np.int64), lens, out_at + plan.out_off, plan.cap
        finally:
            plan.close()
            if d_small is not None:
                d_small.free()

    def median(self, a, b, aligned: bool = False):
        """DOS.median of id pairs (a[k] first, the smaller taxon code side):
        returns (median ids, costs), and with aligned=True also an Aligned
        record of where each pair's alignment lies in the arena (the aligned
        children and the median with gaps that Unions.union reads)."""
        a = np.asarray(a, np.int64)
        b = np.asarray(b, np.int64)
        out = np.where(self.empty[a], b, a).astype(np.int64)
        cost = np.zeros(len(a), np.int64)
        al = Aligned.none(len(a)) if aligned else None
        work = np.nonzero(~self.empty[a] & ~self.empty[b])[0]
        for c0 in range(0, len(work), self.MEDIAN_CHUNK):
            idx = work[c0:c0 + self.MEDIAN_CHUNK]
            x, y = a[idx], b[idx]
            if self.affine:
                # slots: median, median with gaps, aligned a, aligned b
                costs, lens, base, cap = self._run(engine.POYG_MODE_AFFINE_ALIGN, x, y, None)
                slot, sa, sb, smwg = 0, 2, 3, 1
            else:
                # slots: aligned a, aligned b, median
                dws = band_deltaw(self.ln[x], self.ln[y], self.gaps[x], self.gaps[y])
                costs, lens, base, cap = self._run(engine.POYG_MODE_LINEAR_ALIGN, x, y, dws)
                slot, sa, sb, smwg = 2, 0, 1, None
            mlen = lens[:, slot].astype(np.int32)
            moff = base + slot * cap + cap - mlen
            gaps, empty = self._stats(moff, mlen)
            out[idx] = self._append(moff, mlen, gaps, empty)
            cost[idx] = costs
            if aligned:
                alen = lens[:, sa].astype(np.int32)
                al.length[idx] = alen
                al.a_off[idx] = base + sa * cap + cap - alen
                al.b_off[idx] = base + sb * cap + cap - alen
                if smwg is not None:
                    if not np.array_equal(lens[:, smwg], lens[:, sa]):
                        raise engine.PoygError("median with gaps and aligned pair differ in length")
                    al.mwg_off[idx] = base + smwg * cap + cap - alen
        return (out, cost, al) if aligned else (out, cost)

    def distance(self, a, b) -> np.ndarray:
        """DOS.distance of id pairs."""
        a = np.asarray(a, np.int64)
        b = np.asarray(b, np.int64)
        cost = np.zeros(len(a), np.int64)
        work = np.nonzero(~self.empty[a] & ~self.empty[b])[0]
        for c0 in range(0, len(work), self.DISTANCE_CHUNK):
            idx = work[c0:c0 + self.DISTANCE_CHUNK]
            x, y = a[idx], b[idx]
            if self.affine:
                cost[idx] = self._run(engine.POYG_MODE_AFFINE_COST, x, y, None)[0]
            else:
                la, lb = self.ln[x].astype(np.int64), self.ln[y].astype(np.int64)
                dws = band_deltaw(la, lb, self.gaps[x], self.gaps[y], np.maximum(8, np.abs(la - lb)))
                cost[idx] = self._run(engine.POYG_MODE_LINEAR_COST, x, y, dws)[0]
        return cost

    # ------------------------------------------------------- 4-mer profiles
    def kmer_counts(self, ids) -> engine.DeviceBuffer:
        """4-mer profiles (256 float counts each) of the given ids, on the
        device, in the order of `ids`."""
        ids = np.asarray(ids, np.int64)
        n = len(ids)
        meta = np.zeros(n + (n + 1) // 2, np.int64)
        meta[:n] = self.off[ids]
        meta[n:].view(np.int32)[:n] = self.ln[ids]
        d = engine.DeviceBuffer.from_array(self.ctx, meta)
        counts = engine.DeviceBuffer(self.ctx, 4 * 256 * max(n, 1))
        engine._check(engine.load().poyg_kmer_counts(self.ctx.h, self.arena.ptr, d.ptr, d.ptr + 8 * n, n, counts.ptr,
                                                     None), "poyg_kmer_counts")
        self.ctx.sync()
        d.free()
        return counts
