"""GPU Vamana build against the oracle's restatement of the reference build
(rabitq_build_vamana_graph, rabitq_bridge.cpp:224-303): identical degrees,
identical adjacency lists in order, identical distance_evaluations. The
oracle build is itself pinned to the reference build (golden fixture and,
where it is built, the reference bridge) in test_oracle.py."""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _gpu_build(dim, c, comp, **kw):
    import quantization_graph_amd as qg
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    ix = qg.DeviceIndex.from_records(sp, comp)
    deg, edges, evals = sp.build_graph(ix, **kw)
    return deg.cpu().numpy().astype(np.uint32), edges.cpu().numpy().astype(np.uint32), evals


def _same_graph(d1, e1, d2, e2):
    assert np.array_equal(d1, d2), "degrees differ at %s" % np.nonzero(d1 != d2)[0][:10]
    for i in range(d1.size):
        assert np.array_equal(e1[i, :d1[i]], e2[i, :d2[i]]), "adjacency of node %d differs" % i


def test_build_matches_golden_fixture(gpu, golden):
    g = golden
    print("case: golden n=300 dim=128 R=16 L=64 batch=64 refine=1 cap=128", flush=True)
    d, e, ev = _gpu_build(128, g["g_center"], g["g_compact"], r=16, l_build=64, alpha=1.2, beam_width=1,
                          batch=64, refine_passes=1, prune_cap=128)
    _same_graph(d, e, g["g_degrees"], g["g_edges"])
    assert ev == int(g["g_evals"][0])


CASES = [
    # n, dim, R, L, alpha, beam, batch, refine, cap, early_stop
    (1500, 128, 32, 100, 1.2, 1, 256, 0, 0, 0),
    (1500, 128, 16, 50, 1.5, 2, 100, 2, 40, 0),
    (1200, 200, 24, 80, 2.0, 3, 300, 1, 0, 3),
    (800, 960, 32, 64, 1.2, 1, 128, 1, 0, 0),
    (600, 1536, 64, 96, 1.2, 1, 200, 1, 0, 0),
    (2600, 64, 16, 32, 1.2, 1, 2500, 0, 0, 0),  # a back-edge group above the LDS pool capacity
    (1500, 64, 32, 320, 1.2, 1, 256, 1, 300, 0),  # refine pools above the LDS Gram triangle (overflow path)
    (1200, 96, 24, 400, 1.2, 1, 300, 0, 0, 0),    # uncapped pools too large for the Gram triangle
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_R%d_L%d_a%g_b%d_bs%d_rf%d_cap%d_es%d" % c)
def test_build_matches_oracle(gpu, case):
    n, dim, R, L, alpha, beam, batch, refine, cap, es = case
    print("case:", case, flush=True)
    x, _ = gaussian_mixture(n, dim, 8, 0.3, 5, nq=1)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, _ = orc.split(orc.encode_full(x))
    d0, e0, v0 = orc.build_graph(comp, R, L, alpha, beam, batch, refine, cap, es)
    d1, e1, v1 = _gpu_build(dim, c, comp, r=R, l_build=L, alpha=alpha, beam_width=beam, batch=batch,
                            refine_passes=refine, prune_cap=cap, early_stop_hops=es)
    _same_graph(d1, e1, d0, e0)
    assert v1 == v0


def test_build_with_duplicate_points(gpu):
    """Repeated vectors give symmetric distances that round to zero or just
    below it; the memoised-negative rule of the prune then matters."""
    n, dim = 900, 64
    x, _ = gaussian_mixture(n, dim, 4, 0.3, 3, nq=1)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, _ = orc.split(orc.encode_full(x))
    from oracle_bind import _p
    negs = [a for a in range(n) if orc.lib.ro_symmetric_distance(orc.sp, _p(comp[a]), _p(comp[a])) < 0]
    assert len(negs) >= 2
    x[100:120] = x[negs[0]]
    x[300:310] = x[negs[1]]
    comp, _ = orc.split(orc.encode_full(x))
    d0, e0, v0 = orc.build_graph(comp, 16, 48, 1.2, 1, 128, 1, 0, 0)
    d1, e1, v1 = _gpu_build(dim, c, comp, r=16, l_build=48, alpha=1.2, batch=128, refine_passes=1)
    _same_graph(d1, e1, d0, e0)
    assert v1 == v0


def test_build_rejects_bad_parameters(gpu):
    import quantization_graph_amd as qg
    x, _ = gaussian_mixture(64, 64, 2, 0.3, 1, nq=1)
    sp = qg.RabitqSpace(64, 100)
    ix = qg.DeviceIndex.from_records(sp, Oracle(64, 100).split(Oracle(64, 100).encode_full(x))[0])
    for kw in (dict(r=0), dict(r=65), dict(l_build=0), dict(l_build=2000), dict(alpha=0.9),
               dict(alpha=float("nan")), dict(beam_width=0)):
        with pytest.raises(qg._lib.QGError):
            sp.build_graph(ix, **kw)
