"""Fixed-candidate pipeline on the GPU: the one-launch rerank (workgroup per
query) returns the same final ids and estimates, bit for bit, as the
three-launch rerank built from the oracle-checked scan kernels (FP32
restore for INT8 queries, residual intervals, top-k by (estimate, id))."""
import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fixed_case(gpu):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(6000, 960, 24, 0.3, 9, nq=96)
    sp = qg.RabitqSpace(960, 100)
    sp.set_center(x.mean(0).astype(np.float32))
    ix = sp.encode(gpu.from_numpy(x).cuda())
    rng = np.random.default_rng(3)
    cand = np.stack([rng.choice(x.shape[0], 400, replace=False) for _ in range(q.shape[0])]).astype(np.int32)
    return dict(sp=sp, ix=ix, q=gpu.from_numpy(q).cuda(), cand=gpu.from_numpy(cand).cuda())


@pytest.mark.parametrize("codec", ["int8", "full"])
@pytest.mark.parametrize("depth,k", [(100, 10), (128, 10), (37, 20), (10, 10)])
def test_rerank_topk_equals_rerank(gpu, fixed_case, codec, depth, k):
    from quantization_graph_amd import fixed
    c = fixed_case
    sp, ix = c["sp"], c["ix"]
    qv = sp.prepare(c["q"], codec)
    first = fixed.score(sp, ix, qv, c["cand"], codec=codec)
    ids1, d1, _, _ = fixed.rerank(sp, ix, qv, c["cand"], first, k=k, depth=depth, restore_full=codec == "int8")
    ids2, d2 = fixed.rerank_topk(sp, ix, qv, c["cand"], first, k=k, depth=depth)
    assert (ids1.cpu() == ids2.cpu()).all()
    assert (d1.cpu().view(gpu.int32) == d2.cpu().view(gpu.int32)).all()


def test_search_falls_back_past_the_workgroup_limit(gpu, fixed_case):
    """Depths beyond 128 take the three-launch path and agree with it."""
    from quantization_graph_amd import fixed
    c = fixed_case
    sp, ix = c["sp"], c["ix"]
    qv = sp.prepare(c["q"], "int8")
    ids, d = fixed.search(sp, ix, qv, c["cand"], k=10, depth=300, codec="int8")
    first = fixed.score(sp, ix, qv, c["cand"], codec="int8")
    ids1, d1, _, _ = fixed.rerank(sp, ix, qv, c["cand"], first, k=10, depth=300)
    assert (ids.cpu() == ids1.cpu()).all()
    assert (d.cpu().view(gpu.int32) == d1.cpu().view(gpu.int32)).all()
