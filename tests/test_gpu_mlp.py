"""Learned rerank depth on the GPU: the MLP kernels against the FP64 oracle
(oracle/mlp_oracle.py), training convergence, the plain rerank of the
two-stage pipeline against the fused search rerank, and what the trained
network saves over the plain top-100 rerank at the same recall."""
import os
import sys

import numpy as np
import pytest

from conftest import gaussian_mixture

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _feats(rng, n):
    X = rng.normal(size=(n, 24)).astype(np.float32)
    bits = np.zeros((n, 32), np.uint16)
    bits[:, :24] = mo.bf16_bits(X)
    return bits, mo.bits_to_f32(bits[:, :24])


def _params(seed):
    rng = np.random.default_rng(seed)
    P = mo.init_params(seed)
    P[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.2, 64)
    P[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.2, 64)
    P[mo.OFF_B3] = 0.3
    return P


def test_forward_matches_oracle(gpu):
    import torch
    from quantization_graph_amd import learned
    rng = np.random.default_rng(5)
    bits, Xv = _feats(rng, 3001)
    P = _params(7)
    m = learned.MLP(P)
    y = m.forward(torch.from_numpy(bits.view(np.int16)).cuda()).cpu().numpy()
    yo, _ = mo.forward(P, Xv, exact=False)
    err = np.abs(y - yo) / (1.0 + np.abs(yo))
    # FP32 accumulation on the matrix cores vs FP64; a bf16 rounding of H1
    # can flip where Z1 lies within an FP32 ulp of a bf16 halfway point
    assert np.quantile(err, 0.999) < 1e-5
    assert err.max() < 2e-2


def test_gradient_matches_oracle(gpu):
    import torch
    from quantization_graph_amd import learned
    rng = np.random.default_rng(6)
    bits, Xv = _feats(rng, 2500)
    t = np.abs(rng.normal(size=2500)).astype(np.float32)
    P = _params(8)
    m = learned.MLP(P)
    loss, g = m.grad(torch.from_numpy(bits.view(np.int16)).cuda(), torch.from_numpy(t).cuda(), 0.9)
    lo, go = mo.backward(P, Xv, t, 0.9, exact=False)
    g = g.cpu().numpy()
    assert abs(float(loss.item()) - lo) <= 1e-5 * (1 + abs(lo))
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert rel < 1e-3, rel
    for a, b in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_W2, mo.OFF_B2), (mo.OFF_W3, mo.OFF_B3)):
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-3 * np.linalg.norm(go[a:b]) + 1e-9


def test_adam_step_matches_oracle(gpu):
    import torch
    from quantization_graph_amd import _lib, learned
    import ctypes as C
    from quantization_graph_amd.space import _ptr, _stream
    rng = np.random.default_rng(9)
    bits, Xv = _feats(rng, 1024)
    t = np.abs(rng.normal(size=1024)).astype(np.float32)
    P = _params(10)
    m = learned.MLP(P)
    _, go = mo.backward(P, Xv, t, 0.95, exact=False)
    mm = torch.zeros(learned.N_PARAMS, device="cuda")
    vv = torch.zeros(learned.N_PARAMS, device="cuda")
    xb = torch.from_numpy(bits.view(np.int16)).cuda()
    tb = torch.from_numpy(t).cuda()
    ws = m._ws(1024)
    lr = 1e-2
    _lib.check(_lib.lib().qg_mlp_adam_step(_ptr(m.params), _ptr(mm), _ptr(vv), _ptr(xb), _ptr(tb), 1024,
                                           C.c_float(0.95), C.c_float(lr), C.c_float(0.9), C.c_float(0.999),
                                           C.c_float(1e-8), 1, None, _ptr(ws), ws.numel(), _stream()))
    new = m.params.cpu().numpy().astype(np.float64)
    # first Adam step: m/(1-b1) = g, v/(1-b2) = g^2 -> P - lr * g / (|g| + eps)
    expect = P - lr * go / (np.abs(go) + 1e-8)
    big = np.abs(go) > 1e-6 * np.abs(go).max()
    assert np.allclose(new[big], expect[big], rtol=0, atol=1e-6)


def test_training_reduces_pinball_loss(gpu):
    import torch
    from quantization_graph_amd import learned
    rng = np.random.default_rng(11)
    bits, Xv = _feats(rng, 40000)
    t = (np.abs(Xv[:, 0]) * 0.5 + 0.1 * np.abs(Xv[:, 3] * Xv[:, 5])).astype(np.float32)
    m = learned.MLP(seed=1)
    xb = torch.from_numpy(bits.view(np.int16)).cuda()
    tb = torch.from_numpy(t).cuda()
    hist = m.fit(xb, tb, tau=0.9, epochs=6, batch=4096, lr=3e-3)
    assert hist[-1] < 0.4 * hist[0], hist
    # the trained quantile covers about tau of the targets
    y = m.forward(xb).cpu().numpy()
    assert 0.8 < float((t <= y).mean()) < 0.97


def _powerlaw(n, dim, nq):
    import torch
    import bench
    import quantization_graph_amd as qg
    from quantization_graph_amd import learned
    print("case: power-law base %d x %d, %d queries" % (n, dim, nq), flush=True)
    x, q = bench.make_powerlaw(torch, n, dim, nq, 4)
    gt = bench.exact_topk(torch, x, q, 10)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(x.double().mean(0).float().cpu().numpy())
    ix = sp.encode(x)
    r = 32 if dim <= 1024 else 64
    deg, edges, _ = sp.build_graph(ix, r=r, l_build=100, alpha=1.2, batch=2048, refine_passes=1, prune_cap=128)
    rstats = learned.residual_stats(sp, ix)
    return dict(dim=dim, x=x, q=q, gt=gt, sp=sp, ix=ix, deg=deg, edges=edges, rstats=rstats)


@pytest.fixture(scope="module")
def powerlaw_case(gpu):
    return _powerlaw(20000, 256, 3500)


@pytest.fixture(scope="module", params=[256, 1536], ids=lambda d: "D%d" % d)
def parity_case(gpu, request, powerlaw_case):
    """The D=256 case, and D=1536 (code dimension 2048, the learned
    workload's) on a degree-64 graph."""
    if request.param == 256:
        return powerlaw_case
    return _powerlaw(8000, 1536, 400)


@pytest.mark.parametrize("codec", ["full", "int8"])
def test_plain_rerank_equals_fused_search(gpu, parity_case, codec):
    from quantization_graph_amd import learned
    c = parity_case
    sp, ix = c["sp"], c["ix"]
    qv = sp.prepare(c["q"][:300], codec)
    ids, dists, st = sp.search(ix, qv, c["deg"], c["edges"], k=10, l_search=64, rerank=100)
    pid, pd, st2 = learned.search_pool(sp, ix, qv, c["deg"], c["edges"], 64)
    ids2, dists2, depth = learned.rerank(sp, ix, qv, pid, pd, None)
    assert (ids.cpu() == ids2.cpu()).all()
    assert (dists.cpu().view(dtype=__import__("torch").int32) == dists2.cpu().view(dtype=__import__("torch").int32)).all()
    assert (st.cpu() == st2.cpu()).all()
    assert int(depth.min()) == 64 and int(depth.max()) == 64  # min(|pool|, 100) with L = 64


def test_learned_depth_saves_rerank_work(gpu, powerlaw_case):
    """Trained on (estimate, exact) pairs from 3000 training queries, the
    calibrated depth reranks far fewer than 100 candidates per query on 300
    unseen queries while Recall@10 stays within 0.002 of the plain rerank."""
    from quantization_graph_amd import learned
    c = powerlaw_case
    sp, ix, q, gt = c["sp"], c["ix"], c["q"], c["gt"]
    train, val, test = slice(0, 3000), slice(3000, 3200), slice(3200, 3500)
    feats, targets = learned.training_pairs(sp, ix, c["deg"], c["edges"], c["x"], q[train], 128, "full", c["rstats"])
    m = learned.MLP(seed=0)
    hist = m.fit(feats, targets, tau=0.95, epochs=4, batch=8192, lr=2e-3)
    assert hist[-1] < hist[0]
    qv = sp.prepare(q[val], "full")
    pid, pd, _ = learned.search_pool(sp, ix, qv, c["deg"], c["edges"], 128)
    preds = m.forward(learned.features(sp, ix, qv, pid, pd, c["rstats"]))
    kmult, table = learned.calibrate(sp, ix, qv, pid, pd, preds, gt[val])
    qv = sp.prepare(q[test], "full")
    pid, pd, _ = learned.search_pool(sp, ix, qv, c["deg"], c["edges"], 128)
    preds = m.forward(learned.features(sp, ix, qv, pid, pd, c["rstats"]))
    plain, _, _ = learned.rerank(sp, ix, qv, pid, pd, None)
    ids, _, depth = learned.rerank(sp, ix, qv, pid, pd, preds, kmult)
    r_plain, r_learned = learned.recall_at_k(plain, gt[test]), learned.recall_at_k(ids, gt[test])
    print("kmult %.2f plain %.4f learned %.4f mean depth %.1f" % (kmult, r_plain, r_learned, depth.float().mean()))
    assert r_learned >= r_plain - 0.002
    assert float(depth.float().mean()) < 60.0


def _query_norms(qv):
    qs = qv.qs.cpu().numpy().view(np.float32)
    return qs[:, 0], qs[:, 1]


def test_features_match_oracle(gpu, parity_case):
    """Feature rows of full pools, pools cut short (empty slots inside the
    depth) and an empty pool, against the FP64 restatement: each bf16 value
    within one bf16 ulp."""
    import torch
    from quantization_graph_amd import learned
    c = parity_case
    sp, ix = c["sp"], c["ix"]
    qv = sp.prepare(c["q"][:160], "full")
    pid, pd, _ = learned.search_pool(sp, ix, qv, c["deg"], c["edges"], 128)
    pid, pd = pid.clone(), pd.clone()
    for r in range(40):  # pools of 3 .. 120 entries
        pid[r, 3 * r + 3:] = -1
        pd[r, 3 * r + 3:] = float("inf")
    pid[40, :] = -1
    pd[40, :] = float("inf")
    f = learned.features(sp, ix, qv, pid, pd, c["rstats"])
    got = mo.bits_to_f32(f.cpu().numpy().view(np.uint16).reshape(160, 100, 32))
    qn, qns = _query_norms(qv)
    ref = mo.features(pid.cpu().numpy().view(np.uint32), pd.cpu().numpy(), qn, qns, ix.hdr.cpu().numpy(),
                      ix.sfac.cpu().numpy(), ix.rfac.cpu().numpy(), c["rstats"].cpu().numpy(), sp.code_dim)
    assert (got[:, :, 24:] == 0).all()
    diff = np.abs(got[:, :, :24] - ref)
    assert (diff <= 2.0 ** -7 * np.abs(ref) + 1e-30).all(), np.unravel_index(np.argmax(diff), diff.shape)
    assert (got[40] == 0).all() and (got[0, 3:] == 0).all()


@pytest.mark.parametrize("codec", ["full", "int8"])
def test_learned_depth_rule_and_rerank(gpu, parity_case, codec):
    """Per-query depths from given predictions equal the oracle's rule, and
    the learned rerank's ids and estimates equal, bit for bit, the plain
    rerank of pools cut at those depths."""
    import torch
    from quantization_graph_amd import learned
    c = parity_case
    sp, ix = c["sp"], c["ix"]
    qv = sp.prepare(c["q"][:300], codec)
    pid, pd, _ = learned.search_pool(sp, ix, qv, c["deg"], c["edges"], 128)
    rng = np.random.default_rng(11)
    preds_np = (rng.normal(size=(300, 100)) * 0.01 + 0.004).astype(np.float32)
    preds = torch.from_numpy(preds_np).cuda()
    ids, dists, depth = learned.rerank(sp, ix, qv, pid, pd, preds, 1.5)
    _, qns = _query_norms(qv)
    want = mo.rerank_depth(pid.cpu().numpy().view(np.uint32), pd.cpu().numpy(), preds_np, qns, 1.5)
    got = depth.cpu().numpy()
    assert (got == want).all(), (got[got != want][:5], want[got != want][:5])
    assert 10 <= got.min() and got.max() <= 100 and got.min() < got.max()
    cut = torch.arange(pid.shape[1], device="cuda")[None, :] >= depth.long()[:, None]
    pid_t = torch.where(cut, torch.full_like(pid, -1), pid)
    pd_t = torch.where(cut, torch.full_like(pd, float("inf")), pd)
    ids2, dists2, depth2 = learned.rerank(sp, ix, qv, pid_t, pd_t, None)
    assert (depth2.cpu() == depth.cpu()).all()
    assert (ids.cpu() == ids2.cpu()).all()
    assert (dists.cpu().view(torch.int32) == dists2.cpu().view(torch.int32)).all()


@pytest.mark.parametrize("codec", ["full", "int8"])
@pytest.mark.parametrize("L", [64, 128])
def test_fused_learned_search_equals_pipeline(gpu, parity_case, codec, L):
    """qg_search_learned (features, MLP and depth rule inside the search
    kernel, after each query's traversal) equals the four-launch pipeline
    (traversal with pool write-out, qg_rerank_features, qg_mlp_forward,
    qg_learned_rerank) bit for bit: ids, estimates, depths and counters."""
    import torch
    from quantization_graph_amd import learned
    c = parity_case
    sp, ix = c["sp"], c["ix"]
    print("case: fused learned search, D %d, %s, L %d" % (c["dim"], codec, L), flush=True)
    qv = sp.prepare(c["q"][:300], codec)
    for seed, km in ((0, 0.5), (3, 2.0)):
        lr = learned.LearnedRerank(learned.MLP(_params(seed)), km)
        ids, d, depth, st = lr.search(sp, ix, qv, c["deg"], c["edges"], L, c["rstats"])
        ids2, d2, depth2, st2 = lr.search_fused(sp, ix, qv, c["deg"], c["edges"], L, c["rstats"])
        assert torch.equal(depth.cpu(), depth2.cpu().to(depth.dtype))
        assert torch.equal(ids.cpu(), ids2.cpu())
        assert torch.equal(d.cpu().view(torch.int32), d2.cpu().view(torch.int32))
        assert torch.equal(st.cpu(), st2.cpu())
        dd = depth2.cpu().numpy()
        assert dd.min() >= 10 and dd.max() <= 100
