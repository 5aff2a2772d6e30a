"""Learned delta / warm start: the MLP kernels against the oracle and torch,
and the gain of the trained models over the plain rules.

* gb_mlp_loss_grad: loss and every gradient against the FP64 CPU oracle
  (orc_mlp_loss_grad, itself pinned by central finite differences in
  test_oracle.py), relative error <= 2e-4 (FP32 device sums).
* gb_mlp_forward: against a torch FP32 CPU forward of the same parameters.
* gb_mlp_adam: against torch.optim.Adam on the CPU.
* learned_models.json: delta model beats the plain rule max_w / avg_degree
  on unseen graphs (distances bit-identical), warm start cuts PageRank
  iterations at the same tolerance.
"""
import json
import os
import time

import numpy as np
import pytest

import oracle

gb = pytest.importorskip("graphbench_amd")
from graphbench_amd import learned  # noqa: E402


def _np_params(m):
    return [a.astype(np.float32) for a in m.split()]


def _torch():
    import torch
    return torch


def _data(n, din, dout, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, din)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0] * (k + 1)) + 0.3 * X[:, -1] ** 2 for k in range(dout)],
                 axis=1).astype(np.float32)
    return X, T


def test_models_file_format():
    if not os.path.exists(learned.MODELS_PATH):
        pytest.skip("learned_models.json not trained yet")
    d = json.load(open(learned.MODELS_PATH))
    for key, din, dout in (("delta", learned.DELTA_DIN, 2), ("warm_start", 3, 1)):
        m = d[key]
        assert m["din"] == din and m["dout"] == dout
        H = learned.HIDDEN
        total = H * din + H + H * H + H + dout * H + dout
        assert len(m["params"]) == total
        assert np.all(np.isfinite(m["params"]))
        assert len(m["x_mean"]) == din and len(m["x_std"]) == din
        assert min(m["x_std"]) > 0 and min(m["y_std"]) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("n,din,dout", [(1, 6, 2), (300, 3, 1), (10000, 6, 2), (9000, 16, 4)])
def test_mlp_loss_grad_vs_oracle(gpu, n, din, dout):
    torch = _torch()
    m = learned.MLP(din, dout, seed=n)
    X, T = _data(n, din, dout, seed=n + 1)
    loss = m.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    torch.cuda.synchronize()
    P = _np_params(m)
    ref_loss, ref_g = oracle.mlp_loss_grad(X, T, *P)
    assert abs(loss - ref_loss) <= 2e-5 * max(1.0, abs(ref_loss))
    got = m.split(m.grad.cpu().numpy())
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), got, ref_g):
        scale = max(np.abs(b).max(), 1e-6)
        err = np.abs(a.astype(np.float64) - b).max() / scale
        assert err <= 2e-4, (name, err)


@pytest.mark.gpu
def test_graph_features_multi_vs_host(gpu):
    """The delta model's inputs for several graphs in one launch equal the
    same statistics computed on the host from the CSR (log2 V, log2 average
    degree, log2 max and mean weight, log2 max out-degree / average degree,
    log2 sources), and the single-graph call gives the same row."""
    hs = [gb.gen_rmat(12, 8, seed=3, weights="uniform", wlo=1, whi=255),
          gb.gen_er(3000, 5, seed=4, weights="uniform", wlo=7, whi=7),
          gb.gen_rmat(14, 16, seed=5, weights="uniform", wlo=1, whi=65535)]
    gs = [gb.Graph.from_host(h, need_reverse=False) for h in hs]
    F = learned.graph_features_multi(gs, 8)
    for h, g, f in zip(hs, gs, F):
        deg = np.diff(h.row_ptr.astype(np.int64))
        avg = h.E / h.V
        ref = [np.log2(h.V), np.log2(avg), np.log2(max(1, int(h.w.max()))),
               np.log2(max(1.0, float(h.w.astype(np.float64).sum()) / h.E)),
               np.log2(max(1.0, deg.max() / avg)), np.log2(8)]
        np.testing.assert_allclose(f, np.array(ref, np.float32), rtol=1e-6, atol=1e-6)
        assert np.array_equal(learned.graph_features(g, 8), f)
    assert learned.graph_features_multi([], 8).shape == (0, 6)


@pytest.mark.gpu
def test_mlp_forward_vs_torch(gpu):
    torch = _torch()
    m = learned.MLP(6, 2, seed=3)
    X, _ = _data(5000, 6, 2, seed=4)
    Y = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    W1, b1, W2, b2, W3, b3 = [torch.from_numpy(a) for a in _np_params(m)]
    Xt = torch.from_numpy(X)
    ref = torch.relu(torch.relu(Xt @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    np.testing.assert_allclose(Y, ref.numpy(), rtol=1e-4, atol=1e-5)


def _bf16(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("din,dout,n", [(3, 1, 10007), (6, 2, 4096), (16, 4, 33)])
def test_mlp_forward_bf16_mfma(gpu, din, dout, n):
    """The MFMA inference against (a) an FP64 forward over the same bf16
    roundings of the layer-1 activations and of W2 -- any slip in the MFMA
    operand or result layouts mixes hidden units and fails it -- and (b) the
    FP32 forward, within bf16 rounding."""
    torch = _torch()
    m = learned.MLP(din, dout, seed=11)
    X, _ = _data(n, din, dout, seed=12)
    Y = m.forward(torch.from_numpy(X).cuda(), bf16=True).cpu().numpy()
    Y32 = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    W1, b1, W2, b2, W3, b3 = _np_params(m)
    h1 = np.maximum(X @ W1.T + b1, 0).astype(np.float32)
    z2 = _bf16(h1).astype(np.float64) @ _bf16(W2).astype(np.float64).T + b2
    ref = np.maximum(z2, 0) @ W3.T.astype(np.float64) + b3
    scale = max(1.0, float(np.abs(ref).max()))
    np.testing.assert_allclose(Y, ref, rtol=1e-4, atol=2e-5 * scale)
    np.testing.assert_allclose(Y, Y32, rtol=2e-2, atol=2e-2 * scale)


@pytest.mark.gpu
@pytest.mark.parametrize("din,dout,n", [(3, 1, 10007), (6, 2, 4096), (16, 4, 33)])
def test_mlp_forward_f32_mfma(gpu, din, dout, n):
    """The FP32 MFMA inference (the warm start's) against an FP64 forward and
    the per-row FP32 forward: any slip in the MFMA operand or result layouts
    mixes hidden units and fails it; otherwise only the FP32 summation order
    differs."""
    torch = _torch()
    m = learned.MLP(din, dout, seed=13)
    X, _ = _data(n, din, dout, seed=14)
    Y = m.forward(torch.from_numpy(X).cuda(), mfma=True).cpu().numpy()
    Y32 = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in _np_params(m)]
    h1 = np.maximum(X.astype(np.float64) @ W1.T + b1, 0)
    ref = np.maximum(h1 @ W2.T + b2, 0) @ W3.T + b3
    scale = max(1.0, float(np.abs(ref).max()))
    np.testing.assert_allclose(Y, ref, rtol=1e-5, atol=1e-5 * scale)
    np.testing.assert_allclose(Y, Y32, rtol=1e-5, atol=1e-5 * scale)


@pytest.mark.gpu
def test_mlp_adam_vs_torch_optim(gpu):
    torch = _torch()
    m = learned.MLP(3, 1, seed=5)
    X, T = _data(2000, 3, 1, seed=6)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    p_ref = torch.nn.Parameter(m.params.cpu().clone())
    opt = torch.optim.Adam([p_ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(5):
        m.loss_grad(Xd, Td)
        g = m.grad.cpu().clone()
        m.adam(lr=1e-2)
        p_ref.grad = g
        opt.step()
        torch.cuda.synchronize()
        np.testing.assert_allclose(m.params.cpu().numpy(), p_ref.detach().numpy(),
                                   rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_mlp_training_fits(gpu):
    torch = _torch()
    X, T = _data(8192, 3, 1, seed=7)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    m = learned.MLP(3, 1, seed=8)
    losses = m.fit(Xd, Td, steps=400, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0]
    # the loss the kernels report is the oracle's loss at the trained parameters
    ref_loss, _ = oracle.mlp_loss_grad(X, T, *_np_params(m))
    assert abs(m.loss_grad(Xd, Td) - ref_loss) <= 2e-5 * max(1.0, ref_loss)


def _models():
    if not os.path.exists(learned.MODELS_PATH):
        pytest.skip("learned_models.json not trained yet")
    return learned.LearnedModels.load()


def _best_of(fn, reps=3):
    torch = _torch()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return min(ts)


@pytest.mark.gpu
def test_learned_delta_beats_plain_rule(gpu):
    """Unseen R-MAT (scale 17, seed 4242, uniform weights [1, 255], 8 sources):
    the plain rule gives delta = 255 / 16 = 15; the learned delta must be at
    least 1.3x faster, with identical distances."""
    torch = _torch()
    models = _models()
    h = gb.gen_rmat(17, 16, seed=4242, weights="uniform", wlo=1, whi=255)
    g = gb.Graph.from_host(h, need_reverse=False)
    srcs = gb.pick_sources(h, 8, 4243)
    plain = learned.plain_delta(h)
    d, lm = models.predict_delta(g, len(srcs))
    out_plain = gb.sssp(g, srcs, delta=plain, light_max=plain)
    out_learn = gb.sssp(g, srcs, delta=d, light_max=lm)
    assert torch.equal(out_plain, out_learn)
    ref = oracle.dijkstra(h.row_ptr, h.col, h.w, int(srcs[0]))
    assert np.array_equal(gb.u32(out_learn[0]), ref)
    t_plain = _best_of(lambda: gb.sssp(g, srcs, delta=plain, light_max=plain))
    t_learn = _best_of(lambda: gb.sssp(g, srcs, delta=d, light_max=lm))
    print(f"plain delta {plain}: {t_plain*1e3:.2f} ms; learned ({d}, {lm}): {t_learn*1e3:.2f} ms")
    assert t_learn * 1.3 <= t_plain


@pytest.mark.gpu
def test_warm_start_table_matches_per_vertex(gpu):
    """gb_pr_warm_start runs the MLP once per (indeg, outdeg) pair with both
    below 512 and once per other vertex; its x0 matches one MLP row per vertex
    (features from the degrees, the same FP32 MFMA forward), normalised in
    FP64.  R-MAT 16 has vertices past 512 on either degree and dangling ones."""
    torch = _torch()
    models = _models()
    h = gb.gen_rmat(16, 16, seed=777, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    indeg, outdeg = np.diff(r.row_ptr), np.diff(h.row_ptr)
    assert (indeg >= 512).any() and (outdeg >= 512).any() and (outdeg == 0).any()
    p = models.data["warm_start"]
    xm, xs = np.array(p["x_mean"], np.float32), np.array(p["x_std"], np.float32)
    F = np.stack([np.log2(1 + indeg.astype(np.float32)), np.log2(1 + outdeg.astype(np.float32)),
                  (outdeg == 0).astype(np.float32)], 1)
    X = np.ascontiguousarray((F - xm) / xs, np.float32)
    Y = models.pr_mlp.forward(torch.from_numpy(X).cuda(), mfma=True).cpu().numpy()[:, 0]
    rr = np.exp2(Y.astype(np.float64) * p["y_std"][0] + p["y_mean"][0])
    x0 = models.warm_start(g).cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(x0, rr / rr.sum(), rtol=1e-5, atol=0)


@pytest.mark.gpu
def test_learned_warm_start_saves_iterations(gpu):
    """Unseen R-MAT (scale 18, seed 5151): warm start converges in fewer
    iterations to the same fixed point (L1 to the FP64 oracle <= 1e-6)."""
    torch = _torch()
    models = _models()
    h = gb.gen_rmat(18, 16, seed=5151, weights="none")
    r = h.transpose()
    g = gb.Graph.from_host(h, reverse=r)
    x0 = models.warm_start(g)
    s = float(x0.double().sum())
    assert abs(s - 1.0) < 1e-6 and float(x0.min()) > 0
    xu, ru = gb.pagerank(g, tol=1e-7, max_iter=200)
    xw, rw = gb.pagerank(g, tol=1e-7, max_iter=200, x0=x0)
    assert ru["converged"] and rw["converged"]
    ref, _, _ = oracle.pagerank_cfg(r.row_ptr, r.col, h.outdeg(), 0.85, 1e-12, 500)
    for x in (xu, xw):
        assert np.abs(x.cpu().numpy().astype(np.float64) - ref).sum() <= 1e-6
    print(f"iterations: uniform {ru['iterations']}, warm start {rw['iterations']}")
    assert rw["iterations"] < ru["iterations"]
