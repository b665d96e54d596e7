"""Learned join scorer: a 256 -> 256 -> 256 -> 1 network over 4-mer
profiles that predicts a join's per-character distance, so that a TBR round
scores exactly (poyd_amd.tbr.rescore_round) only the candidates it ranks
best.  The network's forward and backward passes, Adam and the feature
kernels are the library's own (poyg_mlp_*, poyg_kmer_counts,
poyg_pair_features); this module holds the host loop.

Features of a pair of sequences: the absolute difference of their 4-mer
count profiles, times X_SCALE, as bf16.  Target: the exact DOS.distance of
the pair (src/seqCS.ml:818-864) divided by Y_SCALE.  The model sees the
unbroken tree's edge medians in place of the broken components' ones (which
an exact round would have to compute first); it learns the exact distances
from them.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import engine, tbr

X_SCALE = 1.0 / 8.0
Y_SCALE = 64.0


class Scorer:
    def __init__(self, ctx: engine.Context, batch_max: int = 4096, seed: int = 1):
        self.ctx = ctx
        self.batch_max = batch_max
        h = C.c_void_p()
        engine._check(engine.load().poyg_mlp_create(ctx.h, batch_max, seed, C.byref(h)), "poyg_mlp_create")
        self.h = h
        self.n_params = int(engine.load().poyg_mlp_param_count())

    def close(self):
        if getattr(self, "h", None):
            engine.load().poyg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        engine._check(engine.load().poyg_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                        self.n_params), "poyg_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        engine._check(engine.load().poyg_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                        self.n_params), "poyg_mlp_set_params")

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        engine._check(engine.load().poyg_mlp_get_grads(self.h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                       self.n_params), "poyg_mlp_get_grads")
        return out

    def grad(self, d_x: int, d_t: int, n: int, want_loss: bool = False) -> float | None:
        loss = C.c_float(0)
        engine._check(engine.load().poyg_mlp_grad(self.h, d_x, d_t, n, C.byref(loss) if want_loss else None),
                      "poyg_mlp_grad")
        return float(loss.value) if want_loss else None

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        engine._check(engine.load().poyg_mlp_adam(self.h, lr, b1, b2, eps), "poyg_mlp_adam")

    def forward(self, d_x: int, n: int, d_y: int) -> None:
        engine._check(engine.load().poyg_mlp_forward(self.h, d_x, n, d_y, None), "poyg_mlp_forward")

    def train(self, x: engine.DeviceBuffer, t: engine.DeviceBuffer, n: int, epochs: int, lr: float,
              seed: int = 0, log=None) -> list:
        """Adam over shuffled batches of the n rows (x: bf16 rows of 256,
        t: f32 targets, both on the device).  Batches are row ranges of a
        permuted copy made on the device once per epoch."""
        rng = np.random.default_rng(seed)
        B = self.batch_max
        xs = engine.DeviceBuffer(self.ctx, 512 * n)
        ts = engine.DeviceBuffer(self.ctx, 4 * n)
        losses = []
        for ep in range(epochs):
            perm = rng.permutation(n).astype(np.int32)
            d_perm = engine.DeviceBuffer.from_array(self.ctx, perm)
            gather_rows(self.ctx, x.ptr, t.ptr, d_perm.ptr, n, xs.ptr, ts.ptr)
            tot, cnt = 0.0, 0
            lr_ep = lr * (0.5 ** (ep // max(1, epochs // 3)))
            for r in range(0, n - B + 1, B):
                want = (r // B) % 64 == 0
                l = self.grad(xs.ptr + 512 * r, ts.ptr + 4 * r, B, want)
                self.adam(lr_ep)
                if want:
                    tot += l
                    cnt += 1
            d_perm.free()
            losses.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.5f} (lr {lr_ep:g})")
        xs.free()
        ts.free()
        return losses

    def predict(self, d_x: int, n: int) -> np.ndarray:
        d_y = engine.DeviceBuffer(self.ctx, 4 * max(n, 1))
        self.forward(d_x, n, d_y.ptr)
        y = d_y.download(np.float32, n)
        d_y.free()
        return y


def gather_rows(ctx, d_x, d_t, d_perm, n, d_xo, d_to) -> None:
    """x_out[r] = x[perm[r]], t_out[r] = t[perm[r]] on the device."""
    L = engine.load()
    engine._check(L.poyg_gather_rows(ctx.h, d_x, 512, d_perm, n, d_xo, None), "poyg_gather_rows")
    engine._check(L.poyg_gather_rows(ctx.h, d_t, 4, d_perm, n, d_to, None), "poyg_gather_rows")


class Profiles:
    """4-mer profiles of a set of store sequences, on the device, with the
    row of each store id."""

    def __init__(self, store, ids):
        ids = np.unique(np.asarray(ids, np.int64).ravel())
        self.ids = ids
        self.counts = store.kmer_counts(ids)

    def rows(self, ids) -> np.ndarray:
        return np.searchsorted(self.ids, np.asarray(ids, np.int64)).astype(np.int32)

    def features(self, ctx, ia, ib) -> engine.DeviceBuffer:
        """bf16 features of the pairs (ia[k], ib[k]) of store ids."""
        n = len(ia)
        idx = np.concatenate([self.rows(ia), self.rows(ib)]).astype(np.int32)
        d_idx = engine.DeviceBuffer.from_array(ctx, idx)
        x = engine.DeviceBuffer(ctx, 512 * max(n, 1))
        engine._check(engine.load().poyg_pair_features(ctx.h, self.counts.ptr, d_idx.ptr, d_idx.ptr + 4 * n, n,
                                                       X_SCALE, x.ptr, None), "poyg_pair_features")
        d_idx.free()
        return x


def unbroken_edge_medians(store, sp: tbr.Space, dirs: tbr.Table, index: dict):
    """The unbroken tree's lazy edge medians (create_lazy_edge) and, for
    every global edge of the neighbourhood, the row of its edge's median."""
    edges = sp.t.edges()
    te = tbr.median_rows(store, dirs.take(np.array([index[(b, a)] for a, b in edges])),
                         dirs.take(np.array([index[(a, b)] for a, b in edges])))
    pos = {frozenset(e): i for i, e in enumerate(edges)}
    g_te = np.array([pos[frozenset(e)] for e in sp.gedge], np.int64)
    return te, g_te


def approx_ids(cand_refs, dirs: tbr.Table, te: tbr.Table, g_te: np.ndarray, K: int) -> np.ndarray:
    """Store ids the model reads for candidate rows: directional rows (the
    dissolved edges and leaf clades) exactly, other edges by the unbroken
    tree's median of the same edge."""
    refs = np.asarray(cand_refs, np.int64)
    out = np.empty((len(refs), dirs.ids.shape[1]), np.int64)
    d = refs < K
    out[d] = dirs.ids[refs[d]]
    out[~d] = te.ids[g_te[refs[~d] - K]]
    return out


def predict_joins(ctx, scorer: Scorer, prof: Profiles, ca: np.ndarray, ea: np.ndarray,
                  chunk: int = 1 << 20) -> np.ndarray:
    """Predicted per-character distances (n, L) of candidates whose model
    inputs are the store ids ca (clade) and ea (edge), (n, L) each."""
    n, L = ca.shape
    flat_a, flat_b = ca.ravel(), ea.ravel()
    out = np.empty(n * L, np.float32)
    for r in range(0, n * L, chunk):
        x = prof.features(ctx, flat_a[r:r + chunk], flat_b[r:r + chunk])
        m = min(chunk, n * L - r)
        out[r:r + m] = scorer.predict(x.ptr, m)
        x.free()
    return out.reshape(n, L) * Y_SCALE


def guided_round(store, scorer: Scorer, sp: tbr.Space, dirs: tbr.Table, index: dict, cand: tbr.Candidates,
                 frac: float = 0.08):
    """Scores every candidate with the model, then the best `frac` of them
    exactly.  Returns the exact round result of the picked candidates and
    the timings of the stages."""
    t0 = time.perf_counter()
    K = len(dirs)
    mark = store.mark()
    te, g_te = unbroken_edge_medians(store, sp, dirs, index)
    t1 = time.perf_counter()
    prof = Profiles(store, np.concatenate([dirs.ids.ravel(), te.ids.ravel()]))
    ca = approx_ids(cand.clade, dirs, te, g_te, K)
    ea = approx_ids(cand.edge, dirs, te, g_te, K)
    pred = predict_joins(store.ctx, scorer, prof, ca, ea)
    score = cand.base + pred.sum(1)
    n_pick = max(1, int(np.ceil(frac * len(score))))
    pick = np.argpartition(score, n_pick - 1)[:n_pick] if n_pick < len(score) else np.arange(len(score))
    t2 = time.perf_counter()
    store.release(mark)
    res = tbr.rescore_round(store, sp, dirs, cand, pick)
    t3 = time.perf_counter()
    prof.counts.free()
    return res, {"edge_medians_s": t1 - t0, "model_s": t2 - t1, "exact_s": t3 - t2, "total_s": t3 - t0,
                 "picked": int(n_pick), "predicted": pred}
