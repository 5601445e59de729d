"""Writes tests/golden/ref_vectors.npz from the reference's own C-ABI bridge
compiled from source (oracle/_ref/librabitq_ref.so, built by oracle/Makefile
in the container that holds the reference checkout).

The fixture holds inputs and the reference's outputs only:
  * encode: 48 base vectors (D = 960 and D = 1536, seed 100, MSE + fp16
    scales, mean center) -> full / compact / residual records, MSB sidecars
    and paper factors;
  * 4 queries per dimension, Full and Int8 codecs -> mode-0 distances,
    1-bit gate estimates (eps0 = 1.9), mode-1/mode-2 remaining-bit
    distances, residual-4 intervals;
  * symmetric distances for 64 record pairs;
  * a Vamana graph built by rabitq_build_vamana_graph on 300 vectors
    (R = 16, L_build = 64, alpha = 1.2);
  * rabitq_paper_estimate (the gate from the compact record) for the Full
    and Int8 codecs, and rabitq_query_distance_b1 under the B1 codec;
  * rabitq_train_centroids on 2000 x 37 random rows (k = 1, 16, 256) and on
    a three-row near tie whose assignment depends on the contraction of
    d += diff * diff.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_bind import Ref  # noqa: E402


def data(n, dim, seed, nq):
    rng = np.random.default_rng(seed)
    centers = rng.normal(size=(16, dim)).astype(np.float32)
    x = (centers[rng.integers(0, 16, n)] + 0.3 * rng.normal(size=(n, dim))).astype(np.float32)
    q = (x[rng.integers(0, n, nq)] + 0.05 * rng.normal(size=(nq, dim))).astype(np.float32)
    return x, q


def main():
    if not Ref.available():
        raise SystemExit("oracle/_ref/librabitq_ref.so is not built")
    out = {}
    for dim in (960, 1536):
        x, q = data(48, dim, dim, 4)
        c = x.mean(0).astype(np.float32)
        ref = Ref(dim, 100)
        ref.set_center(c)
        full = ref.encode_full(x)
        comp, res = ref.split(full)
        msb, fac = ref.sidecar(comp)
        fac["pad"] = 0  # the reference leaves the struct padding uninitialised
        p = "d%d_" % dim
        out.update({p + "x": x, p + "q": q, p + "center": c, p + "full": full, p + "compact": comp,
                    p + "residual": res, p + "msb": msb, p + "factors": fac.view(np.uint8).reshape(-1, 20)})
        ids = np.arange(x.shape[0], dtype=np.uint32)
        for codec, name in ((0, "full"), (3, "int8")):
            ref.set_codec(codec)
            d0, est, rem, ivs = [], [], [], []
            for i in range(q.shape[0]):
                pq = ref.prepare(q[i])
                a = ref.distance_batch(pq, comp, ids, np.zeros(ids.size, np.uint8))
                e = ref.gate_batch(pq, msb, fac, ids, 1.9)
                mode = 2 if codec == 3 else 1
                r = ref.distance_batch(pq, comp, ids, np.full(ids.size, mode, np.uint8), e["short_ip"])
                iv = ref.residual_batch(pq, comp, res, ids, a)
                d0.append(a)
                est.append(e.view(np.uint8).reshape(-1, 24))
                rem.append(r)
                ivs.append(iv.view(np.uint8).reshape(-1, 12))
            out[p + name + "_dist"] = np.stack(d0)
            out[p + name + "_gate"] = np.stack(est)
            out[p + name + "_rem"] = np.stack(rem)
            out[p + name + "_residual"] = np.stack(ivs)
        for codec, name in ((0, "full"), (3, "int8"), (1, "b1")):
            ref.set_codec(codec)
            rec = []
            for i in range(q.shape[0]):
                pq = ref.prepare(q[i])
                if codec == 1:
                    rec.append(ref.b1_batch(pq, comp, ids))
                else:
                    rec.append(ref.paper_estimate_batch(pq, comp, ids, 1.9).view(np.uint8).reshape(-1, 24))
            out[p + name + ("_dist" if codec == 1 else "_record_gate")] = np.stack(rec)
        rng = np.random.default_rng(7)
        pairs = rng.integers(0, x.shape[0], size=(64, 2)).astype(np.uint32)
        out[p + "sym_pairs"] = pairs
        out[p + "sym"] = np.array([ref.symmetric(comp[a], comp[b]) for a, b in pairs], np.float32)
    # small Vamana build
    x, _ = data(300, 128, 11, 1)
    c = x.mean(0).astype(np.float32)
    ref = Ref(128, 100)
    ref.set_center(c)
    comp, res = ref.split(ref.encode_full(x))
    deg, edges, msb, fac, evals = ref.build_graph(comp, 16, 64, 1.2, beam=1, batch=64, refine=1, cap=128)
    out.update({"g_x": x, "g_center": c, "g_compact": comp, "g_degrees": deg, "g_edges": edges,
                "g_evals": np.array([evals], np.uint64)})
    # k-means (rabitq_train_centroids)
    ref = Ref(64, 100)
    km_x = np.random.default_rng(5).standard_normal((2000, 37)).astype(np.float32)
    out["km_x"] = km_x
    for k, sc, it in ((1, 2000, 3), (16, 1000, 5), (256, 1500, 2)):
        out["km_k%d" % k] = ref.train_centroids(km_x, k, sc, 11, it)
    tie = np.array([[0.5413387, 1.6302696], [0.76907164, 1.5360193], [0.0, 0.0]], np.float32)
    out["km_tie_x"] = tie
    out["km_tie"] = ref.train_centroids(tie, 2, 3, 33, 1)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
