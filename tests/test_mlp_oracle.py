"""Pins the MLP oracle (oracle/mlp_oracle.py): its backward against central
finite differences of its own FP64 forward, and its bf16 rounding against
the IEEE definition on representable and halfway values."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    P = mo.init_params(3).astype(np.float64)
    P[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    P[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 64)
    X = rng.normal(size=(64, mo.F))
    t = np.abs(rng.normal(size=64))
    tau = 0.9
    _, g = mo.backward(P, X, t, tau, exact=True)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 40), rng.choice(np.arange(mo.OFF_B1, mo.OFF_W2), 10),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 40), rng.choice(np.arange(mo.OFF_B2, mo.OFF_W3), 10),
                          rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 10), [mo.OFF_B3]])
    h = 1e-7
    for p in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        lp = mo.pinball(mo.forward(Pp, X, exact=True)[0], t, tau).mean()
        lm = mo.pinball(mo.forward(Pm, X, exact=True)[0], t, tau).mean()
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p])) + 1e-8, (p, fd, g[p])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 65504.0], np.float32)
    out = mo.bf16_round(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert list(out) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, 65536.0]
    assert np.array_equal(mo.bits_to_f32(mo.bf16_bits(vals)), out)
