"""CPU oracle of the learned rerank-depth MLP (quantization_graph_amd/csrc/qg_mlp.hip).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the GPU kernels; the
product never imports it. The MLP has no counterpart in the reference, so
this file restates the kernels' own definition:

    params (FP32) = W1[64][24] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
    Z1 = X W1b^T + b1          X, W1b: values rounded to bf16 (nearest even)
    H1 = bf16(relu(Z1))
    Z2 = H1 W2b^T + b2         W2b: W2 rounded to bf16
    H2 = relu(Z2)
    y  = H2 . W3 + b3
    loss = mean_i pinball_tau(t_i - y_i)

`forward` / `backward` take `exact=True` for the same formulas in FP64
without any rounding (smooth, so finite differences pin the backward), and
`exact=False` for the kernels' rounding points with FP64 arithmetic in
between (the comparison target of the GPU, within FP32 accumulation error).

`features` and `rerank_depth` restate the feature rows and the per-query
depth rule of the learned rerank (k_rerank_features, k_learned_rerank).
Parity pins: forward/backward against finite differences of the exact
form; features against the kernel's bf16 output within one bf16 ulp;
rerank_depth exactly (same FP32 operations).
"""
import numpy as np

F, H = 24, 64
OFF_W1, OFF_B1 = 0, H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def bf16_round(x):
    """float32 values rounded to bf16 (round to nearest even), as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def bits_to_f32(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[OFF_W1:OFF_B1].reshape(H, F), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3], P[OFF_B3])


def init_params(seed):
    """Uniform fan-in initialisation (the scheme of quantization_graph_amd.learned.MLP)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(N_PARAMS, np.float32)
    P[OFF_W1:OFF_B1] = rng.uniform(-1, 1, H * F) / np.sqrt(F)
    P[OFF_W2:OFF_B2] = rng.uniform(-1, 1, H * H) / np.sqrt(H)
    P[OFF_W3:OFF_B3] = rng.uniform(-1, 1, H) / np.sqrt(H)
    return P


def forward(P, X, exact=False):
    """X: [n, 24] (bf16-valued float32 unless exact). Returns y and the
    activations needed by backward."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    if not exact:
        W1 = bf16_round(W1.astype(np.float32)).astype(np.float64)
        W2 = bf16_round(W2.astype(np.float32)).astype(np.float64)
    Z1 = X @ W1.T + b1
    H1 = np.maximum(Z1, 0.0)
    if not exact:
        H1 = bf16_round(H1.astype(np.float32)).astype(np.float64)
    Z2 = H1 @ W2.T + b2
    H2 = np.maximum(Z2, 0.0)
    y = H2 @ W3 + b3
    return y, (X, W1, W2, Z1, H1, Z2, H2)


def pinball(y, t, tau):
    u = np.asarray(t, np.float64) - y
    return np.where(u >= 0, tau * u, (tau - 1.0) * u)


def backward(P, X, t, tau, exact=False):
    """Mean pinball loss and its gradient [N_PARAMS] (FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    y, (Xd, W1u, W2u, Z1, H1, Z2, H2) = forward(P, X, exact)
    n = y.shape[0]
    u = np.asarray(t, np.float64) - y
    dy = np.where(u >= 0, -tau, 1.0 - tau) / n
    g = np.zeros(N_PARAMS, np.float64)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = H2.T @ dy
    dZ2 = (dy[:, None] * W3[None, :]) * (Z2 > 0)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    g[OFF_W2:OFF_B2] = (dZ2.T @ H1).reshape(-1)
    dZ1 = (dZ2 @ W2u) * (Z1 > 0)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W1:OFF_B1] = (dZ1.T @ Xd).reshape(-1)
    return pinball(y, t, tau).mean(), g


def features(pid, pd, qn, qns, hdr, sfac, rfac, rstats, cd, depth=100):
    """The 24 rerank features of each query's first `depth` pool entries
    (k_rerank_features), in FP64 before the bf16 rounding: [nq, depth, 24],
    rows past the pool (first empty id within the depth) all zero.
    pid [nq, L] (uint32, 0xFFFFFFFF = empty), pd [nq, L]; qn / qns the
    query norm and its square; hdr / sfac / rfac / rstats per base id."""
    pid = np.asarray(pid, np.uint32)
    nq, L = pid.shape
    out = np.zeros((nq, depth, F), np.float64)
    lim = min(L, depth)
    sqcd, nbk = np.sqrt(cd), cd / 16.0
    for qi in range(nq):
        empty = np.nonzero(pid[qi, :lim] == 0xFFFFFFFF)[0]
        n = int(empty[0]) if empty.size else lim
        if n == 0:
            continue
        S = float(qns[qi]) if qns[qi] > 0 else 1.0
        q = float(qn[qi])
        e_all = pd[qi].astype(np.float64)
        e0, e9 = e_all[0], e_all[min(9, n - 1)]
        for i in range(n):
            idx = int(pid[qi, i])
            e = e_all[i]
            hx, hy = float(hdr[idx, 0]), float(hdr[idx, 1])
            sx, sy = float(sfac[idx, 0]), float(sfac[idx, 1])
            bound = sy * q
            f = out[qi, i]
            f[0] = e / S
            f[1] = (e - e0) / S
            f[2] = (e - e9) / S
            f[3] = i / 100.0
            f[4] = np.log1p(i) / np.log(101.0)
            f[5] = ((e_all[i + 1] if i + 1 < n else e) - e) / S
            f[6] = (e - (e_all[i - 1] if i > 0 else e)) / S
            f[7] = bound / S
            f[8] = hx / S
            f[9] = np.sqrt(max(hx, 0.0)) * q / S
            f[10] = hy * q / S
            f[11] = sx * q / S
            f[12] = e0 / S
            f[13] = e9 / S
            f[14] = float(rfac[idx, 0]) / S
            f[15] = float(rfac[idx, 1]) / S
            f[16] = float(rstats[idx, 0]) * q * sqcd / S
            f[17] = float(rstats[idx, 1]) * q * sqcd / S
            f[18] = float(rstats[idx, 2])
            f[19] = 4.0 * q * float(rstats[idx, 0]) * np.sqrt(nbk) / S
            f[20] = i / n
            f[21] = (e_all[min(i + 5, n - 1)] - e_all[max(i - 5, 0)]) / S
            f[22] = e / (e9 if e9 > 0 else 1.0)
            f[23] = (e - e9) / (bound + 1e-12)
    bad = ~np.isfinite(out)
    out[bad] = 0.0
    return np.clip(out, -1e4, 1e4)


def rerank_depth(pid, pd, preds, qns, kmult, depth=100, k=10, min_depth=10):
    """Per-query learned rerank depth (k_learned_rerank): rc = min(|pool|,
    max(depth, k)), |pool| = the first empty slot; ub = max over the first min(rc, min_depth) entries of
    estimate + kmult * max(pred, 0) * S; the depth is the last position whose
    estimate - kmult * max(pred, 0) * S <= ub, at least min(rc, min_depth).
    Evaluated in FP32 as the kernel does."""
    pid = np.asarray(pid, np.uint32)
    nq, L = pid.shape
    out = np.zeros(nq, np.int64)
    km = np.float32(kmult)
    for qi in range(nq):
        empty = np.nonzero(pid[qi] == 0xFFFFFFFF)[0]
        n = int(empty[0]) if empty.size else L
        rc = min(n, max(depth, k))
        if rc == 0:
            continue
        S = np.float32(qns[qi]) if qns[qi] > 0 else np.float32(1.0)
        top = min(rc, min_depth)
        p = np.maximum(preds[qi, :rc].astype(np.float32), np.float32(0.0))
        e = pd[qi, :rc].astype(np.float32)
        ub = np.max(e[:top] + (km * p[:top]) * S)
        lb = e - (km * p) * S
        ok = np.nonzero(lb <= ub)[0]
        last = int(ok[-1]) + 1 if ok.size else 0
        out[qi] = max(last, top)
    return out
