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

* the ExRaBitQ space state "EXRBTQ31" (RaBitQSpace::saveState / loadState,
  Ours/core/hnswlib/space_rabitq.h:3139-3345): an 8-byte magic, a packed
  header of record sizes and flags, the u32 rotation seed, the K centroids
  and the code_dim rotation signs (little-endian, no padding);
* the Vamana index file "VMNARBQ1" (VamanaIndex::saveIndex / loadIndex,
  Ours/core/hnswlib/vamana_index.h:727-829): an 8-byte magic, u32 version 1,
  u64 max elements, count, record size, R, L_build, f32 alpha, u64 beam
  width and start node, then the compact records, the u64 labels and for
  every node a u32 degree followed by its u32 neighbour ids.

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


# --------------------------------------------------- EXRBTQ31 space state

_STATE_HEAD = struct.Struct("<8sQQIIIQQQQIQQQQQBQIQ")


def space_state_fields(dim, centroid_count=1, residual_block_size=16, residual_mse=True, residual_fp16=True):
    """Header values saveState writes for a space of rabitq_space_new(dim,
    seed, K, 4, block, mse, fp16) (K = centroid_count; residual records
    stored externally): the record sizes of the compact (hot) and residual
    (cold) records and the residual flags."""
    from .space import code_dim_for
    cd = code_dim_for(dim)
    blocks = (cd + residual_block_size - 1) // residual_block_size
    scale_bytes = blocks * (2 if residual_fp16 else 4)
    compact_code = (cd + 1) // 2
    residual_code = (cd * 4 + 7) // 8
    data_size = 8 + 8 + compact_code + 1
    full_size = 8 + 8 + 8 + scale_bytes + compact_code + 1 + residual_code
    flags = 1 | 2 | (4 if residual_mse else 0) | (8 if residual_fp16 else 0)
    return dict(dim=dim, code_dim=cd, total_bits=4, short_bits=1, remaining_bits=3, encoded_header_size=8,
                short_factor_size=8, residual_factor_size=8, residual_scale_bytes=scale_bytes, residual_bits=4,
                full_code_bytes=compact_code, residual_code_bytes=residual_code, data_size=data_size,
                full_data_size=full_size, residual_disk_record_bytes=8 + scale_bytes + residual_code,
                external_residual_storage=1, residual_block_size=residual_block_size, flags=flags,
                centroid_count=centroid_count)


def write_space_state(path, dim, seed, centroids, signs, residual_block_size=16, residual_mse=True,
                      residual_fp16=True):
    """Writes the space state saveState writes: centroids [K, dim] float32
    (K = 1: the mean center), signs [code_dim] float32 (the rotation signs,
    +-1/sqrt(code_dim), as qg_space_signs holds them)."""
    centroids = np.ascontiguousarray(centroids, np.float32).reshape(-1, dim)
    signs = np.ascontiguousarray(signs, np.float32).reshape(-1)
    f = space_state_fields(dim, centroids.shape[0], residual_block_size, residual_mse, residual_fp16)
    if signs.size != f["code_dim"]:
        raise ValueError("expected %d rotation signs, got %d" % (f["code_dim"], signs.size))
    head = _STATE_HEAD.pack(b"EXRBTQ31", *[f[k] for k in _STATE_KEYS])
    with open(path, "wb") as out:
        out.write(head)
        out.write(struct.pack("<I", seed))
        centroids.tofile(out)
        signs.tofile(out)


_STATE_KEYS = ("dim", "code_dim", "total_bits", "short_bits", "remaining_bits", "encoded_header_size",
               "short_factor_size", "residual_factor_size", "residual_scale_bytes", "residual_bits",
               "full_code_bytes", "residual_code_bytes", "data_size", "full_data_size",
               "residual_disk_record_bytes", "external_residual_storage", "residual_block_size", "flags",
               "centroid_count")


def read_space_state(path):
    """(header dict, seed, centroids [K, dim], signs [code_dim]) of an
    EXRBTQ31 space state; raises on another magic or a short file."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < _STATE_HEAD.size + 4 or raw[:8] != b"EXRBTQ31":
        raise ValueError("%s: not an EXRBTQ31 space state" % path)
    vals = _STATE_HEAD.unpack_from(raw, 0)
    head = dict(zip(_STATE_KEYS, vals[1:]))
    (seed,) = struct.unpack_from("<I", raw, _STATE_HEAD.size)
    k, dim, cd = head["centroid_count"], head["dim"], head["code_dim"]
    off = _STATE_HEAD.size + 4
    if len(raw) != off + 4 * (k * dim + cd):
        raise ValueError("%s: payload size does not match the header" % path)
    centroids = np.frombuffer(raw, np.float32, k * dim, off).reshape(k, dim).copy()
    signs = np.frombuffer(raw, np.float32, cd, off + 4 * k * dim).copy()
    return head, seed, centroids, signs


# ---------------------------------------------------- VMNARBQ1 index file

_INDEX_HEAD = struct.Struct("<8sIQQQQQfQQ")


def write_vamana_index(path, compact, degrees, edges, r, l_build, alpha, beam_width=1, start_node=0,
                       labels=None, max_elements=None):
    """Writes the index saveIndex writes: compact records [n, data_size]
    uint8 (the reference record layout, as DeviceIndex.to_records returns),
    the graph as degrees [n] and edges [n, >= max degree], labels [n] (default
    0..n-1, the bulk build's labels)."""
    compact = np.ascontiguousarray(compact, np.uint8)
    n, data_size = compact.shape
    degrees = np.asarray(degrees, np.uint32).reshape(-1)
    edges = np.asarray(edges, np.uint32).reshape(n, -1)
    if degrees.size != n:
        raise ValueError("degrees has %d entries for %d records" % (degrees.size, n))
    if n and int(degrees.max()) > edges.shape[1]:
        raise ValueError("a degree exceeds the edge row width")
    labels = np.arange(n, dtype=np.uint64) if labels is None else np.asarray(labels, np.uint64).reshape(-1)
    max_elements = n if max_elements is None else max_elements
    with open(path, "wb") as out:
        out.write(_INDEX_HEAD.pack(b"VMNARBQ1", 1, max_elements, n, data_size, r, l_build, alpha, beam_width,
                                   start_node))
        compact.tofile(out)
        labels.tofile(out)
        col = np.arange(edges.shape[1])[None, :]
        lists = np.concatenate([degrees[:, None], edges], axis=1)
        keep = np.concatenate([np.ones((n, 1), bool), col < degrees[:, None]], axis=1)
        lists[keep].astype(np.uint32).tofile(out)


def read_vamana_index(path):
    """dict(header, compact [n, data_size] uint8, labels [n] uint64, degrees
    [n] uint32, edges [n, max degree] uint32) of a VMNARBQ1 index file."""
    raw = np.fromfile(path, np.uint8)
    if raw.size < _INDEX_HEAD.size:
        raise ValueError("%s: truncated header" % path)
    magic, version, max_el, n, data_size, r, l_build, alpha, beam, start = _INDEX_HEAD.unpack_from(raw.tobytes()[
        :_INDEX_HEAD.size])
    if magic != b"VMNARBQ1" or version != 1:
        raise ValueError("%s: invalid Vamana index header" % path)
    off = _INDEX_HEAD.size
    end = off + n * data_size + 8 * n
    if raw.size < end:
        raise ValueError("%s: truncated index" % path)
    compact = raw[off:off + n * data_size].reshape(n, data_size)
    labels = raw[off + n * data_size:end].view(np.uint64)
    words = raw[end:]
    if words.size % 4:
        raise ValueError("%s: adjacency lists are not whole words" % path)
    words = words.view(np.uint32)
    starts = np.empty(n, np.int64)
    pos = 0
    for i in range(n):
        if pos >= words.size:
            raise ValueError("%s: truncated index" % path)
        starts[i] = pos
        pos += 1 + int(words[pos])
    if pos != words.size:
        raise ValueError("%s: trailing bytes after the adjacency lists" % path)
    degrees = words[starts].astype(np.uint32) if n else np.zeros(0, np.uint32)
    width = int(degrees.max()) if n else 0
    edges = np.full((n, width), 0xFFFFFFFF, np.uint32)
    col = np.arange(width)[None, :]
    mask = col < degrees[:, None]
    edges[mask] = words[(starts[:, None] + 1 + col)[mask]]
    head = dict(max_elements=max_el, count=n, data_size=data_size, r=r, l_build=l_build, alpha=alpha,
                beam_width=beam, start_node=start)
    return dict(header=head, compact=compact, labels=labels, degrees=degrees, edges=edges)


def save_index(path, index, degrees, edges, r, l_build, alpha, beam_width=1, start_node=0):
    """VamanaIndex::saveIndex of a device index and its graph: the compact
    records (DeviceIndex.to_records) and adjacency (degrees [n], edges
    [n, >= R], device or host) as the VMNARBQ1 file the reference loads."""
    import torch
    comp, _ = index.to_records()
    deg = degrees.cpu().numpy() if isinstance(degrees, torch.Tensor) else np.asarray(degrees)
    edg = edges.cpu().numpy() if isinstance(edges, torch.Tensor) else np.asarray(edges)
    write_vamana_index(path, comp.cpu().numpy(), deg.view(np.uint32) if deg.dtype == np.int32 else deg,
                       edg.view(np.uint32) if edg.dtype == np.int32 else edg, r, l_build, alpha, beam_width,
                       start_node)


def load_index(path, space, residual=None):
    """VamanaIndex::loadIndex onto the device: (DeviceIndex, degrees [n]
    int32, edges [n, R] int32 with unused slots 0xFFFFFFFF, header) from a
    VMNARBQ1 file; residual records (the reference keeps them in their own
    store) may be given as [n, residual record bytes]."""
    import torch
    from .space import DeviceIndex
    got = read_vamana_index(path)
    head = got["header"]
    if head["data_size"] != space.compact_record_bytes:
        raise ValueError("%s: record size %d does not match the space's %d" % (path, head["data_size"],
                                                                               space.compact_record_bytes))
    ix = DeviceIndex.from_records(space, np.ascontiguousarray(got["compact"]), residual)
    width = max(int(head["r"]), got["edges"].shape[1])
    edges = np.full((head["count"], width), 0xFFFFFFFF, np.uint32)
    edges[:, :got["edges"].shape[1]] = got["edges"]
    deg = torch.from_numpy(got["degrees"].view(np.int32).copy()).cuda()
    return ix, deg, torch.from_numpy(edges.view(np.int32)).cuda(), head
