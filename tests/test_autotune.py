"""The search autotuner's helpers on the CPU: residency from LDS and
registers, the lognormal list generator, and the workload key."""
import pytest

torch = pytest.importorskip("torch")


def test_resident_slots_follow_lds_and_registers():
    from quantization_graph_amd import autotune
    from quantization_graph_amd import _lib
    lds = lambda cd, L: autotune.STATIC_LDS_BYTES + _lib.lib().qg_search_lds_bytes(cd, L)
    # cd 2048, L 100: about 37 KB -> 4 workgroups per CU; 4 waves each = 16 waves, the register limit
    assert autotune.resident_slots(2048, 100, 4, cus=256) == 256 * min(160 * 1024 // lds(2048, 100), 4) == 1024
    # L 4096 needs about 64 KB -> 2 per CU
    assert autotune.resident_slots(1024, 4096, 4, cus=256) == 256 * (160 * 1024 // lds(1024, 4096)) == 512
    # two-wave workgroups: up to 8 per CU, LDS permitting
    assert autotune.resident_slots(1024, 64, 2, cus=10) == 10 * min(160 * 1024 // lds(1024, 64), 8)


def test_lognormal_lists_are_seeded_and_clipped(monkeypatch):
    from quantization_graph_amd import autotune
    monkeypatch.setattr(torch.Tensor, "cuda", lambda self: self)
    a = autotune.lognormal_lists(20000, 400, 0.6, 64, 4096, 3)
    b = autotune.lognormal_lists(20000, 400, 0.6, 64, 4096, 3)
    assert torch.equal(a, b) and a.dtype == torch.int32
    assert int(a.min()) >= 64 and int(a.max()) <= 4096
    assert abs(float(a.float().median()) - 400) < 15
    c = autotune.lognormal_lists(20000, 400, 0.6, 64, 4096, 4)
    assert not torch.equal(a, c)


def test_order_candidates():
    from quantization_graph_amd import autotune
    lq = torch.tensor([5, 90, 12, 90, 40], dtype=torch.int32)
    perm = autotune.SearchTuner.order_of(lq, "longest_first")
    assert perm.tolist() == [1, 3, 4, 2, 0]  # stable for equal lengths
    assert autotune.SearchTuner.order_of(lq, "given") is None
