"""A device index and its GPU-built graph saved in the reference's own
formats (EXRBTQ31 space state, VMNARBQ1 index) pass through the reference's
loaders unchanged (oracle/_ref/libindexio_ref.so, where it travelled with
the tree), and load back onto the device with identical records, graph and
search results."""
import os
import tempfile

import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import RefIndexIO

pytestmark = pytest.mark.gpu


def _bytes(p):
    with open(p, "rb") as f:
        return f.read()


def test_index_and_state_round_trip(gpu):
    import quantization_graph_amd as qg
    from quantization_graph_amd import formats
    print("formats round trip: 1200 x 200, R 24", flush=True)
    x, q = gaussian_mixture(1200, 200, 8, 0.3, 4, nq=16)
    sp = qg.RabitqSpace(200, 100)
    sp.set_center(x.mean(0))
    ix = sp.encode(gpu.from_numpy(x).cuda())
    deg, edges, _ = sp.build_graph(ix, r=24, l_build=64, alpha=1.2, batch=256, refine_passes=1)
    with tempfile.TemporaryDirectory() as td:
        st, vm = os.path.join(td, "space.bin"), os.path.join(td, "index.vmn")
        sp.save_state(st)
        formats.save_index(vm, ix, deg, edges, 24, 64, 1.2)
        if RefIndexIO.available():
            io = RefIndexIO()
            back_st, back_vm = os.path.join(td, "s2.bin"), os.path.join(td, "i2.vmn")
            assert io.load_save_state(200, st, back_st) == 0
            assert _bytes(back_st) == _bytes(st)
            h = io.space(200, 100, x.mean(0).astype(np.float32))
            try:
                assert io.load_save_index(h, vm, back_vm) == 0
            finally:
                io.free(h)
            assert _bytes(back_vm) == _bytes(vm)
        sp2 = qg.RabitqSpace.load_state(st)
        assert np.array_equal(sp2.host_constants()[1], sp.host_constants()[1])
        _, res = ix.to_records()  # residual records live in their own store
        ix2, deg2, edges2, head = formats.load_index(vm, sp2, residual=res)
    assert head["count"] == 1200 and head["r"] == 24
    assert np.array_equal(deg2.cpu().numpy(), deg.cpu().numpy())
    assert np.array_equal(edges2.cpu().numpy(), edges.cpu().numpy())
    c1, _ = ix.to_records()
    c2, _ = ix2.to_records()
    assert np.array_equal(c1.cpu().numpy(), c2.cpu().numpy())
    assert np.array_equal(ix.msb.cpu().numpy(), ix2.msb.cpu().numpy())
    for name in ("rscale", "rcodes", "rfac"):
        assert np.array_equal(getattr(ix, name).cpu().numpy(), getattr(ix2, name).cpu().numpy()), name
    qv = sp.prepare(gpu.from_numpy(q).cuda(), "full")
    qv2 = sp2.prepare(gpu.from_numpy(q).cuda(), "full")
    a = sp.search(ix, qv, deg, edges, k=10, l_search=64)
    b = sp2.search(ix2, qv2, deg2, edges2, k=10, l_search=64)
    for u, v in zip(a[:2], b[:2]):
        assert np.array_equal(u.cpu().numpy(), v.cpu().numpy())


def test_state_round_trip_with_centroids(gpu):
    """A space of 4 centroids: its EXRBTQ31 file holds the 4 centroids, the
    reference's loadState / saveState (a space made with K = 4) writes it
    back byte for byte, and the space loaded here encodes the same records."""
    import quantization_graph_amd as qg
    x, _ = gaussian_mixture(600, 200, 4, 0.3, 7)
    cents = np.stack([x[i::4].mean(0) for i in range(4)]).astype(np.float32)
    sp = qg.RabitqSpace(200, 100, 4)
    sp.set_centroids(cents)
    with tempfile.TemporaryDirectory() as td:
        st = os.path.join(td, "space4.bin")
        sp.save_state(st)
        if RefIndexIO.available():
            back = os.path.join(td, "back4.bin")
            assert RefIndexIO().load_save_state(200, st, back, centroid_count=4) == 0
            assert _bytes(back) == _bytes(st)
        sp2 = qg.RabitqSpace.load_state(st)
    assert sp2.centroid_count == 4
    xg = gpu.from_numpy(x).cuda()
    a, _ = sp.encode(xg).to_records()
    b, _ = sp2.encode(xg).to_records()
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_load_state_rejects_foreign_signs(gpu):
    import quantization_graph_amd as qg
    from quantization_graph_amd import formats
    sp = qg.RabitqSpace(100, 100)
    signs, center = sp.host_constants()
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.bin")
        formats.write_space_state(p, 100, 100, center, -signs)
        with pytest.raises(ValueError, match="rotation signs"):
            qg.RabitqSpace.load_state(p)
