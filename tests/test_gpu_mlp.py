"""The learned join scorer's kernels against CPU references.

* The bf16 MFMA product tiles (poyg_gemm_bf16_f32) against a float64 numpy
  product of the same bf16 values, for row-major, transposed and ragged
  shapes.
* The network's gradient (poyg_mlp_grad: forward, mean squared error,
  backward) against two CPU references: a float64 restatement that rounds
  activations and back-propagated errors to bf16 where the kernels do
  (tight), and plain float32 autograd in torch over the bf16 weights and
  inputs (loose: the bf16 activations are the only difference).
* 4-mer profiles and pair features against numpy; a training run that
  learns a synthetic regression.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = 256
W1, B1, W2, B2, W3, B3, NP = 0, 65536, 65792, 131328, 131584, 131840, 131841


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def note(msg):
    print(f"[case] {msg}", flush=True)


def bf16(a):
    """float32 -> bf16 bits, round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def f32(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def rb(a):
    """round through bf16"""
    return f32(bf16(a)).astype(np.float64)


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, False, False), (256, 256, 256, False, False),
                                         (100, 70, 96, False, False), (256, 256, 512, True, False),
                                         (33, 129, 64, False, True), (256, 256, 4096, True, True)])
def test_gemm_tiles(ctx, M, N, K, ta, tb):
    from poyd_amd import engine
    note(f"gemm {M}x{N}x{K} ta={ta} tb={tb}")
    rng = np.random.default_rng(M * 7 + N + K)
    A = bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = bf16(rng.standard_normal((K, N)).astype(np.float32))
    # stored as given or transposed; strides describe A(i, k) and B(k, j)
    a_store = np.ascontiguousarray(A.T if ta else A)
    b_store = np.ascontiguousarray(B.T if tb else B)
    a_rs, a_cs = (1, M) if ta else (K, 1)
    b_rs, b_cs = (1, K) if tb else (N, 1)
    da = engine.DeviceBuffer.from_array(ctx, a_store)
    db = engine.DeviceBuffer.from_array(ctx, b_store)
    dc = engine.DeviceBuffer(ctx, 4 * M * N)
    rc = engine.load().poyg_gemm_bf16_f32(ctx.h, M, N, K, da.ptr, a_rs, a_cs, db.ptr, b_rs, b_cs, dc.ptr, N, None)
    assert rc == 0
    ctx.sync()
    got = dc.download(np.float32, M * N).reshape(M, N).astype(np.float64)
    want = f32(A).astype(np.float64) @ f32(B).astype(np.float64)
    scale = np.sqrt(K) + np.abs(want)
    assert np.max(np.abs(got - want) / scale) < 1e-5


def reference_grads(p, xb, t):
    """float64 restatement of the kernels' forward and backward passes with
    their bf16 roundings (activations, errors)."""
    x = f32(xb).astype(np.float64)
    w1 = rb(p[W1:B1].reshape(H, H))
    w2 = rb(p[W2:B2].reshape(H, H))
    w3 = rb(p[W3:B3])
    b1, b2, b3 = p[B1:W2].astype(np.float64), p[B2:W3].astype(np.float64), float(p[B3])
    h1 = rb(np.maximum(x @ w1 + b1, 0))
    h2 = rb(np.maximum(h1 @ w2 + b2, 0))
    y = h2 @ w3 + b3
    n = len(t)
    dy = 2 * (y - t) / n
    dz2 = rb(np.where(h2 > 0, dy[:, None] * w3[None, :], 0))
    dz1 = rb(np.where(h1 > 0, dz2 @ w2.T, 0))
    g = np.zeros(NP)
    g[W1:B1] = (x.T @ dz1).ravel()
    g[B1:W2] = dz1.sum(0)
    g[W2:B2] = (h1.T @ dz2).ravel()
    g[B2:W3] = dz2.sum(0)
    g[W3:B3] = h2.T @ dy
    g[B3] = dy.sum()
    return g, float(np.mean((y - t) ** 2))


def autograd_grads(p, xb, t):
    """Plain float32 autograd on the CPU over the bf16 weights and inputs."""
    import torch
    x = torch.tensor(f32(xb))
    w1 = torch.tensor(rb(p[W1:B1]).astype(np.float32).reshape(H, H), requires_grad=True)
    w2 = torch.tensor(rb(p[W2:B2]).astype(np.float32).reshape(H, H), requires_grad=True)
    w3 = torch.tensor(rb(p[W3:B3]).astype(np.float32), requires_grad=True)
    b1 = torch.tensor(p[B1:W2], requires_grad=True)
    b2 = torch.tensor(p[B2:W3], requires_grad=True)
    b3 = torch.tensor(p[B3:NP], requires_grad=True)
    y = torch.relu(torch.relu(x @ w1 + b1) @ w2 + b2) @ w3 + b3
    loss = torch.mean((y - torch.tensor(t, dtype=torch.float32)) ** 2)
    loss.backward()
    g = np.concatenate([w1.grad.numpy().ravel(), b1.grad.numpy(), w2.grad.numpy().ravel(), b2.grad.numpy(),
                        w3.grad.numpy(), b3.grad.numpy()])
    return g.astype(np.float64)


def segments():
    return {"W1": (W1, B1), "b1": (B1, W2), "W2": (W2, B2), "b2": (B2, W3), "w3": (W3, B3), "b3": (B3, NP)}


@pytest.mark.parametrize("n", [4096, 1000])
def test_gradient_matches_cpu_references(ctx, n):
    from poyd_amd import engine, surrogate
    note(f"mlp gradient, batch {n}")
    rng = np.random.default_rng(n)
    sc = surrogate.Scorer(ctx, batch_max=4096, seed=5)
    p = sc.params()
    p[B1:W2] = rng.standard_normal(H).astype(np.float32) * 0.1
    p[B2:W3] = rng.standard_normal(H).astype(np.float32) * 0.1
    p[B3] = 0.3
    sc.set_params(p)
    xb = bf16(np.abs(rng.standard_normal((n, H))).astype(np.float32))
    t = rng.standard_normal(n).astype(np.float32) * 2
    dx = engine.DeviceBuffer.from_array(ctx, xb)
    dt = engine.DeviceBuffer.from_array(ctx, t)
    loss = sc.grad(dx.ptr, dt.ptr, n, True)
    g = sc.grads().astype(np.float64)
    want, wloss = reference_grads(p, xb, t.astype(np.float64))
    assert abs(loss - wloss) <= 1e-4 * max(1.0, wloss)
    auto = autograd_grads(p, xb, t)
    for name, (a, b) in segments().items():
        rel = np.linalg.norm(g[a:b] - want[a:b]) / max(np.linalg.norm(want[a:b]), 1e-12)
        assert rel < 2e-3, (name, rel)
        rel32 = np.linalg.norm(g[a:b] - auto[a:b]) / max(np.linalg.norm(auto[a:b]), 1e-12)
        assert rel32 < 3e-2, (name, rel32)
    # the forward pass alone gives the same predictions as the loss pass
    dy = engine.DeviceBuffer(ctx, 4 * n)
    sc.forward(dx.ptr, n, dy.ptr)
    y = dy.download(np.float32, n)
    x = f32(xb).astype(np.float64)
    h1 = rb(np.maximum(x @ rb(p[W1:B1].reshape(H, H)) + p[B1:W2], 0))
    h2 = rb(np.maximum(h1 @ rb(p[W2:B2].reshape(H, H)) + p[B2:W3], 0))
    # equal up to the rare activation whose f32 sum lands on the other side
    # of a bf16 rounding boundary than the float64 one (one bf16 ulp)
    dev = np.abs(y - (h2 @ rb(p[W3:B3]) + p[B3]))
    assert np.median(dev) < 1e-5 and np.mean(dev < 1e-4) > 0.99 and dev.max() < 1e-2
    sc.close()


def test_kmer_profiles_and_features(ctx):
    from poyd_amd import engine, seqstore, surrogate
    note("4-mer profiles and pair features")
    rng = np.random.default_rng(9)
    seqs = []
    for k in range(40):
        n = int(rng.integers(0, 300))
        alphabet = [1, 2, 4, 8] if k % 3 else list(range(1, 32))
        seqs.append(np.concatenate([[16], rng.choice(alphabet, size=n)]).astype(np.uint8))
    store = seqstore.DeviceStore(ctx, engine.CostMatrix.of_sub_indel(1, 2), 1 << 20)
    ids = store.add(seqs)
    prof = surrogate.Profiles(store, ids)
    counts = prof.counts.download(np.float32, 256 * len(ids)).reshape(len(ids), 256)

    def base(x):
        b = int(x) & 15
        return (b & -b).bit_length() - 1 if b else -1

    for k, s in enumerate(seqs):
        want = np.zeros(256)
        bs = [base(x) for x in s]
        for i in range(1, len(s) - 3):
            w = bs[i:i + 4]
            if min(w) >= 0:
                want[w[0] * 64 + w[1] * 16 + w[2] * 4 + w[3]] += 1
        assert np.array_equal(counts[k], want), k
    ia = rng.integers(0, len(ids), 100)
    ib = rng.integers(0, len(ids), 100)
    x = prof.features(ctx, ids[ia], ids[ib])
    got = x.download(np.uint16, 100 * 256).reshape(100, 256)
    want = bf16((np.abs(counts[ia] - counts[ib]) * surrogate.X_SCALE).astype(np.float32))
    assert np.array_equal(got, want)


def test_training_learns_a_regression(ctx):
    from poyd_amd import engine, surrogate
    note("mlp training on a synthetic regression")
    rng = np.random.default_rng(4)
    n = 65536
    xf = np.abs(rng.standard_normal((n, H))).astype(np.float32)
    wtrue = rng.standard_normal(H) / 16
    t = (np.abs(xf @ wtrue) + 0.5 * xf[:, 0] * xf[:, 1]).astype(np.float32)
    dx = engine.DeviceBuffer.from_array(ctx, bf16(xf))
    dt = engine.DeviceBuffer.from_array(ctx, t)
    sc = surrogate.Scorer(ctx, batch_max=1024, seed=11)
    y0 = sc.predict(dx.ptr, n)
    losses = sc.train(dx, dt, n, epochs=6, lr=2e-3, seed=1)
    y1 = sc.predict(dx.ptr, n)
    base = float(np.mean((t - t.mean()) ** 2))
    assert float(np.mean((y1 - t) ** 2)) < 0.1 * base, (losses, base)
    assert float(np.mean((y1 - t) ** 2)) < float(np.mean((y0 - t) ** 2))
    sc.close()
