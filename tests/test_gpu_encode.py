"""The encode kernel against the oracle on inputs chosen to reach every
branch of its critical-point scan: exact ties between events (rows whose
|o| are all equal after the rotation, two-valued and quantised rows, the
identity rotation of seed 0), sparse and one-hot rows, a row equal to the
center (zero residual), tiny components at and around the 1e-5 cut, rows
of widely different scales, and dimensions from 48 to 1536. Every record
byte must be equal.
"""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _adversarial_rows(dim, seed):
    rng = np.random.default_rng(seed)
    rows = []
    rows.append(np.zeros(dim))                                   # the center itself (below)
    one = np.zeros(dim); one[3 % dim] = 1.0; rows.append(one)    # one-hot
    rows.append(np.ones(dim))                                    # constant
    rows.append(np.where(np.arange(dim) % 2 == 0, 1.0, 2.0))     # two values
    rows.append(np.floor(rng.normal(size=dim) * 3) / 3)          # quantised: many ties
    sp = rng.normal(size=dim); sp[np.arange(dim) % 7 != 0] = 0.0; rows.append(sp)
    tiny = rng.normal(size=dim); tiny[::3] = 2.0 ** -rng.integers(10, 40, size=tiny[::3].shape); rows.append(tiny)
    cut = rng.normal(size=dim) * 1e-5; cut[0] = 1.0; rows.append(cut)   # |o| around eps
    rows.append(rng.normal(size=dim) ** 3)                        # heavy tails
    rows.append(rng.normal(size=dim) * 1e6)
    rows.append(rng.normal(size=dim) * 1e-6)
    rows.append(np.full(dim, -0.5))
    return np.stack(rows).astype(np.float32)


@pytest.mark.parametrize("dim", [48, 64, 100, 200, 960, 1536])
@pytest.mark.parametrize("seed", [100, 0])
def test_encode_adversarial_rows(gpu, dim, seed):
    import quantization_graph_amd as qg
    x = _adversarial_rows(dim, dim + seed)
    base, _ = gaussian_mixture(64, dim, 4, 0.3, dim)
    x = np.concatenate([x, base]).astype(np.float32)
    c = np.zeros(dim, np.float32)
    sp = qg.RabitqSpace(dim, seed)
    sp.set_center(c)
    orc = Oracle(dim, seed)
    orc.set_center(c)
    comp, res = sp.encode(gpu.from_numpy(x).cuda()).to_records()
    comp_o, res_o = orc.split(orc.encode_full(x))
    comp, res = comp.cpu().numpy(), res.cpu().numpy()
    for i in range(x.shape[0]):
        assert np.array_equal(comp[i], comp_o[i]), "compact record of row %d" % i
        assert np.array_equal(res[i], res_o[i]), "residual record of row %d" % i


@pytest.mark.parametrize("dim", [960, 1536])
def test_encode_many_rows(gpu, dim):
    """2048 rows of a Gaussian mixture (a batch of many workgroups per CU)."""
    import quantization_graph_amd as qg
    x, _ = gaussian_mixture(2048, dim, 32, 0.3, 7)
    c = x.mean(0).astype(np.float32)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, res = sp.encode(gpu.from_numpy(x).cuda()).to_records()
    comp_o, res_o = orc.split(orc.encode_full(x))
    assert np.array_equal(comp.cpu().numpy(), comp_o)
    assert np.array_equal(res.cpu().numpy(), res_o)
