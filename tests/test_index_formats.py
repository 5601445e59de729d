"""The reference's own serialisation formats, written and read here
byte-for-byte as the reference writes and reads them:

* the ExRaBitQ space state "EXRBTQ31" (RaBitQSpace::saveState / loadState,
  Ours/core/hnswlib/space_rabitq.h:3139-3345);
* the Vamana index file "VMNARBQ1" (VamanaIndex::saveIndex / loadIndex,
  Ours/core/hnswlib/vamana_index.h:727-829).

Files the compiled reference writes (oracle/_ref/libindexio_ref.so) are
parsed and compared with the space, records and graph they hold; files
written here must equal the reference's byte for byte and pass through the
reference's loaders unchanged.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle_bind import Oracle, Ref, RefIndexIO
from quantization_graph_amd import formats

pytestmark = pytest.mark.skipif(not RefIndexIO.available(), reason="oracle/_ref/libindexio_ref.so not built")


def _bytes(p):
    with open(p, "rb") as f:
        return f.read()


@pytest.mark.parametrize("dim,mse,fp16", [(960, True, True), (100, False, True), (1536, True, False)])
def test_space_state_matches_reference(dim, mse, fp16):
    io = RefIndexIO()
    rng = np.random.default_rng(dim)
    center = rng.standard_normal(dim).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        ref_path, mine, back = (os.path.join(td, n) for n in ("ref.bin", "mine.bin", "back.bin"))
        assert io.save_state(dim, 100, ref_path, center, mse, fp16) == 0
        head, seed, cents, signs = formats.read_space_state(ref_path)
        assert head == formats.space_state_fields(dim, 1, 16, mse, fp16)
        assert seed == 100
        assert np.array_equal(cents[0], center)
        assert np.array_equal(signs, Oracle(dim, 100).signs())
        formats.write_space_state(mine, dim, 100, center, signs, residual_mse=mse, residual_fp16=fp16)
        assert _bytes(mine) == _bytes(ref_path)
        assert io.load_save_state(dim, mine, back, mse, fp16) == 0
        assert _bytes(back) == _bytes(mine)


def test_space_state_rejected_when_corrupt():
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.bin")
        formats.write_space_state(p, 64, 7, np.zeros(64, np.float32), np.ones(64, np.float32))
        raw = bytearray(_bytes(p))
        with open(p, "wb") as f:
            f.write(bytes(raw[:-4]))
        with pytest.raises(ValueError):
            formats.read_space_state(p)
        raw[7] = ord("0")  # EXRBTQ30 is an older layout
        with open(p, "wb") as f:
            f.write(bytes(raw))
        with pytest.raises(ValueError):
            formats.read_space_state(p)
        with pytest.raises(ValueError):
            formats.write_space_state(p, 64, 7, np.zeros(64, np.float32), np.ones(32, np.float32))


@pytest.mark.parametrize("refine", [0, 1])
def test_vamana_index_matches_reference(refine):
    io = RefIndexIO()
    dim, n = 128, 300
    rng = np.random.default_rng(3 + refine)
    x = rng.standard_normal((n, dim)).astype(np.float32)
    center = x.mean(0).astype(np.float32)
    h = io.space(dim, 100, center)
    try:
        comp = io.compact_records(h, x)
        with tempfile.TemporaryDirectory() as td:
            ref_path, mine, back = (os.path.join(td, k) for k in ("ref.vmn", "mine.vmn", "back.vmn"))
            assert io.build_save_index(h, comp, 16, 32, 1.2, 1, 64, refine, ref_path) == 0
            got = formats.read_vamana_index(ref_path)
            assert got["header"] == dict(max_elements=n, count=n, data_size=comp.shape[1], r=16, l_build=32,
                                         alpha=np.float32(1.2), beam_width=1, start_node=0)
            assert np.array_equal(got["compact"], comp)
            assert np.array_equal(got["labels"], np.arange(n, dtype=np.uint64))
            # the graph is the bridge's build of the same records
            if Ref.available():
                ref = Ref(dim, 100)
                ref.set_center(center)
                deg, edges, _ = ref.build_graph(comp, 16, 32, 1.2, beam=1, batch=64, refine=refine)[:3]
                assert np.array_equal(got["degrees"], deg)
                for i in range(n):
                    assert np.array_equal(got["edges"][i, :deg[i]], edges[i, :deg[i]])
            formats.write_vamana_index(mine, got["compact"], got["degrees"], got["edges"], 16, 32, 1.2, 1, 0)
            assert _bytes(mine) == _bytes(ref_path)
            assert io.load_save_index(h, mine, back) == 0
            assert _bytes(back) == _bytes(mine)
    finally:
        io.free(h)


def test_vamana_index_rejected_when_corrupt():
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "i.vmn")
        comp = np.zeros((3, 8), np.uint8)
        formats.write_vamana_index(p, comp, [1, 1, 0], [[1], [0], [0]], 4, 8, 1.2)
        got = formats.read_vamana_index(p)
        assert got["degrees"].tolist() == [1, 1, 0] and got["edges"][:2, 0].tolist() == [1, 0]
        raw = _bytes(p)
        with open(p, "wb") as f:
            f.write(raw[:-2])
        with pytest.raises(ValueError):
            formats.read_vamana_index(p)
        with open(p, "wb") as f:
            f.write(b"X" + raw[1:])
        with pytest.raises(ValueError):
            formats.read_vamana_index(p)
