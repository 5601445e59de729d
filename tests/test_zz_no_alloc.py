"""qg_search_noalloc, the search entry point that allocates nothing (its
workspace is a file-scope device array), against qg_search_ex with a
caller-allocated workspace: the same ids, estimates, counters and pools, bit
for bit, on two calls in a row, the second over a smaller index with fewer
queries (so the scratch holds the first call's leftovers). Both are called
through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu


def _case(torch, n, nq, seed, dim=960):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(n, dim, 12, 0.3, seed, nq=nq)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(x.mean(0).astype(np.float32))
    ix = sp.encode(torch.from_numpy(x).cuda())
    deg, edges, _ = sp.build_graph(ix, r=32, l_build=64, alpha=1.2, batch=512, refine_passes=1)
    return sp, ix, q, deg, edges


def _outputs(torch, nq, k, L):
    return (torch.full((nq, k), -7, dtype=torch.int32, device="cuda"),
            torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda"),
            torch.full((nq, 6), -7, dtype=torch.int64, device="cuda"),
            torch.full((nq, L), -7, dtype=torch.int32, device="cuda"),
            torch.full((nq, L), -7.0, dtype=torch.float32, device="cuda"))


def _search(torch, fn, sp, ix, qv, deg, edges, k, L, lq, slots, ws=None):
    from quantization_graph_amd import _lib
    from quantization_graph_amd.space import _ptr, _stream
    out = _outputs(torch, qv.nq, k, L)
    di, dq = ix.desc(), qv.desc()
    args = [sp.handle, C.byref(di), C.byref(dq), qv.codec, _ptr(deg), _ptr(edges), edges.shape[1], k, L, _ptr(lq),
            C.c_float(1.9), 100, slots, 4] + [_ptr(t) for t in out]
    if ws is not None:
        args += [_ptr(ws), ws.numel()]
    _lib.check(fn(*args, _stream()))
    torch.cuda.synchronize()
    return [t.cpu() for t in out]


def test_noalloc_search_equals_workspace_search(gpu):
    import torch
    from quantization_graph_amd import _lib
    L_ = _lib.lib()
    assert L_.qg_search_noalloc_scratch_bytes() == 32 << 20
    cases = [(3000, 12, 41, 8, 200), (1200, 5, 42, 3, 64)]  # the second call is smaller
    for n, nq, seed, slots, L in cases:
        print("case: noalloc vs workspace search, n %d, %d queries, %d slots, L %d" % (n, nq, slots, L), flush=True)
        sp, ix, q, deg, edges = _case(torch, n, nq, seed)
        lq = torch.tensor(([L, 16, 37, 100] * nq)[:nq], dtype=torch.int32, device="cuda")
        for codec in ("int8", "full"):
            qv = sp.prepare(torch.from_numpy(q).cuda(), codec)
            ws = torch.empty(L_.qg_search_workspace_bytes(n, slots), dtype=torch.uint8, device="cuda")
            want = _search(torch, L_.qg_search_ex, sp, ix, qv, deg, edges, 10, L, lq, slots, ws)
            got = _search(torch, L_.qg_search_noalloc, sp, ix, qv, deg, edges, 10, L, lq, slots)
            for w, g in zip(want, got):
                assert torch.equal(w.view(torch.int32) if w.dtype == torch.float32 else w,
                                   g.view(torch.int32) if g.dtype == torch.float32 else g)
            assert (want[0][:, 0] >= 0).all()


def test_noalloc_search_rejects_oversized_scratch(gpu):
    """Slots whose visited bitsets exceed the 32 MiB array fail with an
    error instead of running."""
    import torch
    from quantization_graph_amd import _lib
    L_ = _lib.lib()
    sp, ix, q, deg, edges = _case(torch, 600, 2, 43)
    qv = sp.prepare(torch.from_numpy(q).cuda(), "full")
    lq = torch.tensor([32, 32], dtype=torch.int32, device="cuda")
    slots = (32 << 20) // ((600 + 32) // 32 * 4) + 1
    with pytest.raises(_lib.QGError, match="above its"):
        _search(torch, L_.qg_search_noalloc, sp, ix, qv, deg, edges, 10, 32, lq, slots)
    # the same call with few slots runs
    _search(torch, L_.qg_search_noalloc, sp, ix, qv, deg, edges, 10, 32, lq, 4)
