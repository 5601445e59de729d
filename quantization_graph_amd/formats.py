"""On-disk formats of the reference's callers.

* fvecs / ivecs: per row a little-endian int32 width followed by that many
  float32 / int32 values (the TEXMEX layout read by diskann_runner.rs:1847-1920
  and written by Ours/experiments/run_ours.py:52-87).
* the canonical DiskANN graph file (storage::bin::save_graph / load_graph,
  baselines/diskann/diskann-providers/src/storage/bin.rs:234-380): a 24-byte
  header {u64 file size, u32 max degree, u32 start point, u64 number of
  frozen start points}, then for every node, frozen points last, a u32
  neighbour count followed by the u32 neighbour ids.

* the 1-bit sign-plane database cache of the query-codec scan
  (experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp:226-339):
  a 64-byte header {char magic[16] = "GIST1M1BITBLK1", u32 version = 1,
  u32 input dim, u32 code dim, u32 batch = 32, u64 vector count, u64 size
  of the source fvecs file, u64 seed, u64 payload bytes}, then the u64 words
  in blocks of 32 vectors, word-major inside a block.

Readers memory-map the files, so a 1M x 1536 base set is not copied twice.
"""
from __future__ import annotations

import os
import struct
from pathlib import Path

import numpy as np

GRAPH_HEADER_BYTES = 24
SIGN_CACHE_MAGIC = b"GIST1M1BITBLK1"
SIGN_CACHE_HEADER = struct.Struct("<16sIIIIQQQQ")


def _vecs_shape(path, itemsize=4):
    path = Path(path)
    size = path.stat().st_size
    if size < 4:
        raise ValueError("%s is empty" % path)
    with path.open("rb") as f:
        width = struct.unpack("<i", f.read(4))[0]
    if width <= 0:
        raise ValueError("%s has invalid first dimension %d" % (path, width))
    row = 4 + itemsize * width
    if size % row:
        raise ValueError("%s has size %d, not divisible by the row size %d" % (path, size, row))
    return size // row, width


def vecs_count(path):
    """Number of rows of an fvecs / ivecs file (from its size and first width)."""
    return _vecs_shape(path)[0]


def read_fvecs(path, mmap=True):
    """float32 [n, dim]; every row must carry the same width."""
    n, dim = _vecs_shape(path)
    raw = np.memmap(path, dtype=np.int32, mode="r") if mmap else np.fromfile(path, dtype=np.int32)
    rows = raw.reshape(n, dim + 1)
    if not np.all(rows[:, 0] == dim):
        raise ValueError("%s: fvecs dimension changes between rows" % path)
    return rows[:, 1:].view(np.float32)


def read_ivecs(path):
    """int32 [n, k] (all rows of one width)."""
    n, k = _vecs_shape(path)
    rows = np.fromfile(path, dtype=np.int32).reshape(n, k + 1)
    if not np.all(rows[:, 0] == k):
        raise ValueError("%s: ivecs width changes between rows" % path)
    return rows[:, 1:]


def read_ivecs_topk(path, k):
    """The first k ids of every ground-truth row, as uint32 [n, k]; rows
    narrower than k are rejected (read_ivecs_topk, diskann_runner.rs:1891-1920)."""
    gt = read_ivecs(path)
    if gt.shape[1] < k:
        raise ValueError("%s groundtruth k=%d is smaller than requested recall@%d" % (path, gt.shape[1], k))
    return np.ascontiguousarray(gt[:, :k]).astype(np.uint32)


def write_fvecs(path, arr):
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    n, dim = arr.shape
    out = np.empty((n, dim + 1), np.int32)
    out[:, 0] = dim
    out[:, 1:] = arr.view(np.int32)
    out.tofile(path)


def write_ivecs(path, arr):
    arr = np.ascontiguousarray(arr, dtype=np.int32)
    n, k = arr.shape
    out = np.empty((n, k + 1), np.int32)
    out[:, 0] = k
    out[:, 1:] = arr
    out.tofile(path)


def write_graph(path, degrees, edges, start_point=0, frozen_lists=((0,),), max_degree=None):
    """Write adjacency lists (degrees [n], edges [n, R]) as a DiskANN graph
    file, with `frozen_lists` appended as the lists of the frozen start
    points. The header's max degree is `max_degree` when given (the
    provider's configured degree), else the largest list. Returns the file
    size, as save_graph does."""
    degrees = np.asarray(degrees, dtype=np.int64).reshape(-1)
    edges = np.asarray(edges)
    n = degrees.size
    if edges.shape[0] != n or (n and int(degrees.max()) > edges.shape[1]):
        raise ValueError("degrees and edges do not match")
    lists = [np.asarray(f, dtype=np.uint32) for f in frozen_lists]
    observed = max([int(degrees.max()) if n else 0] + [l.size for l in lists])
    total = GRAPH_HEADER_BYTES + 4 * (n + int(degrees.sum()) + len(lists) + sum(l.size for l in lists))
    # one u32 stream: for node i, [deg_i, edges_i[:deg_i]]
    width = edges.shape[1]
    body = np.zeros((n, width + 1), np.uint32)
    body[:, 0] = degrees
    body[:, 1:] = edges.astype(np.uint32, copy=False)
    keep = np.arange(width + 1)[None, :] <= degrees[:, None]
    with open(path, "wb") as f:
        f.write(struct.pack("<QIIQ", total, int(max_degree if max_degree is not None else observed),
                            int(start_point), len(lists)))
        f.write(body[keep].tobytes())
        for l in lists:
            f.write(struct.pack("<I", l.size))
            f.write(l.tobytes())
    return total


def read_graph_header(path):
    with open(path, "rb") as f:
        size, max_degree, start, frozen = struct.unpack("<QIIQ", f.read(GRAPH_HEADER_BYTES))
    return dict(file_size=size, max_degree=max_degree, start_point=start, frozen_points=frozen)


def read_graph(path):
    """Adjacency of a DiskANN graph file: (degrees uint32 [m], edges uint32
    [m, D], header), m counting every list in the file (frozen points
    included, last) and D the largest list."""
    head = read_graph_header(path)
    if head["file_size"] != os.path.getsize(path):
        raise ValueError("%s: header size %d, file size %d" % (path, head["file_size"], os.path.getsize(path)))
    words = np.fromfile(path, dtype=np.uint32, offset=GRAPH_HEADER_BYTES)
    starts, pos = [], 0
    while pos < words.size:
        starts.append(pos)
        pos += 1 + int(words[pos])
    if pos != words.size:
        raise ValueError("%s: adjacency lists overrun the file" % path)
    starts = np.asarray(starts, np.int64)
    degrees = words[starts].astype(np.uint32)
    width = int(degrees.max()) if degrees.size else 0
    edges = np.zeros((degrees.size, width), np.uint32)
    col = np.arange(width)[None, :]
    mask = col < degrees[:, None]
    edges[mask] = words[(starts[:, None] + 1 + col)[mask]]
    return degrees, edges, head


def graph_edge_count(path):
    """Out-edges over every list of the file (count_diskann_graph_edges,
    diskann_runner.rs:1814-1832)."""
    return int(read_graph(path)[0].astype(np.int64).sum())


def write_sign_cache(path, db, input_dim, code_dim, data_count, source_size, seed):
    """Writes a sign-plane database (u64 words in the blocked layout) as the
    reference's cache file (save_cache, :314-339)."""
    words = np.ascontiguousarray(db).view(np.uint64).reshape(-1)
    if words.size != data_count * (code_dim // 64):
        raise ValueError("payload holds %d words, expected %d" % (words.size, data_count * (code_dim // 64)))
    head = SIGN_CACHE_HEADER.pack(SIGN_CACHE_MAGIC, 1, input_dim, code_dim, 32, data_count, source_size, seed,
                                  words.nbytes)
    with open(path, "wb") as f:
        f.write(head)
        words.tofile(f)


def read_sign_cache(path, mmap=True):
    """(header dict, u64 words) of a sign-plane database cache file; raises
    on the checks of load_cache (:291-312) that do not need the source
    file: magic, version, batch and payload size."""
    with open(path, "rb") as f:
        raw = f.read(SIGN_CACHE_HEADER.size)
    if len(raw) != SIGN_CACHE_HEADER.size:
        raise ValueError("%s: truncated header" % path)
    magic, version, input_dim, code_dim, batch, count, source_size, seed, payload = SIGN_CACHE_HEADER.unpack(raw)
    if magic[:15] != SIGN_CACHE_MAGIC + b"\0" or version != 1 or batch != 32:
        raise ValueError("%s: not a version-1 sign-plane cache" % path)
    if code_dim % 64 or payload != count * (code_dim // 64) * 8:
        raise ValueError("%s: payload size does not match the header" % path)
    if os.path.getsize(path) < SIGN_CACHE_HEADER.size + payload:
        raise ValueError("%s: truncated payload" % path)
    head = dict(input_dim=input_dim, code_dim=code_dim, batch=batch, data_count=count, source_size=source_size,
                seed=seed, payload_bytes=payload)
    if mmap:
        words = np.memmap(path, dtype=np.uint64, mode="r", offset=SIGN_CACHE_HEADER.size, shape=(payload // 8,))
    else:
        words = np.fromfile(path, dtype=np.uint64, offset=SIGN_CACHE_HEADER.size, count=payload // 8)
    return head, words
