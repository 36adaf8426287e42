"""The search that hands the graph over as it goes (pq_find_overlaps_export):
node records and edge_t records equal what pq_graph_export gives after the
search, and both equal the oracle's graph, for any batch size (several batches
each add their part of the CSR), in both search modes, with sides beyond the
16 staged edges (the shared edge pool), and with an edges buffer too small."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu

EDGE_DT = {4: np.dtype([("dest", "<u4"), ("len", "<u2"), ("type", "u1"), ("pad", "u1")]),
           8: np.dtype([("dest", "<u8"), ("len", "<u2"), ("type", "u1"), ("pad", "V5")])}


def _edges(buf, E, ib):
    got = buf.array[:E * EDGE_DT[ib].itemsize].view(EDGE_DT[ib])
    return np.stack([got["dest"].astype(np.int64), got["len"].astype(np.int64),
                     got["type"].astype(np.int64)], 1)


def _after_search(ctx, mode, ib):
    """Node records, edges and length sum by a search followed by pq_graph_export."""
    ctx.set_search_batch(0)
    ctx.find_overlaps(mode)
    n, E = ctx.graph_size()
    nb = ctx.host_buffer(n * 8)
    eb = ctx.host_buffer(max(1, E) * EDGE_DT[ib].itemsize)
    s = ctx.graph_export(nb, eb, ib)
    out = (nb.array.copy(), _edges(eb, E, ib), s)
    nb.free()
    eb.free()
    return out


def _check_batches(ctx, mode, ib, want, batches, where):
    n = ctx.n
    wn, we, ws = want
    for batch in batches:
        ctx.set_search_batch(batch)
        nb = ctx.host_buffer(n * 8)
        eb = ctx.host_buffer(max(1, len(we)) * EDGE_DT[ib].itemsize)
        E, s = ctx.find_overlaps_export(nb, eb, ib, mode)
        assert E == len(we), f"{where} batch {batch}: {E} edges, want {len(we)}"
        assert s == ws, f"{where} batch {batch}: overlap sum"
        assert np.array_equal(nb.array, wn), f"{where} batch {batch}: node records"
        got = _edges(eb, E, ib)
        if not np.array_equal(got, we):
            bad = np.nonzero((got != we).any(1))[0][:5]
            raise AssertionError(f"{where} batch {batch}: edges differ at {bad.tolist()}")
        # the result also stays on the device as after pq_find_overlaps
        assert ctx.graph_size() == (n, len(we))
        nb.free()
        eb.free()
    ctx.set_search_batch(0)


def _oracle_edges(g):
    return np.stack([np.asarray(g["dest"]).astype(np.int64), np.asarray(g["len"]).astype(np.int64),
                     np.asarray(g["type"]).astype(np.int64)], 1)


@pytest.mark.parametrize("ib", [4, 8])
def test_export_during_search_plain_and_learned(ctx, ib):
    rng = np.random.default_rng(41)
    codes, off, _ = _random_reads(rng, 20000, 6000, 100, 100, err=0.004)
    ctx.upload(codes, off, 63)
    ctx.build_index()
    want = _after_search(ctx, pq.SEARCH_PLAIN, ib)
    ex = orc.search_fm(orc.build_index(codes, off, 63))
    assert np.array_equal(want[1], _oracle_edges(ex))
    _check_batches(ctx, pq.SEARCH_PLAIN, ib, want, [0, 1000, 1777, 5999], "plain")
    ctx.learned_train()
    wl = _after_search(ctx, pq.SEARCH_LEARNED, ib)
    assert np.array_equal(wl[1], want[1]) and np.array_equal(wl[0], want[0])
    _check_batches(ctx, pq.SEARCH_LEARNED, ib, want, [0, 999, 4096], "learned")


def test_export_during_search_pool_and_uneven(ctx):
    # uneven lengths, and 40 copies of one read: the sides that hold the copies
    # as one candidate group have more than 16 edges and go through the shared
    # edge pool, whose offsets span the batches
    rng = np.random.default_rng(43)
    _, _, reads = _random_reads(rng, 3000, 2500, 40, 70, err=0.002)
    reads = reads + [reads[5].copy() for _ in range(40)]
    codes = np.concatenate(reads).astype(np.uint8)
    off = np.zeros(len(reads) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    ctx.upload(codes, off, 21)
    ctx.build_index()
    want = _after_search(ctx, pq.SEARCH_PLAIN, 8)
    ex = orc.search_explicit(codes, off, 21)
    assert np.array_equal(want[1], _oracle_edges(ex))
    assert (np.diff(np.asarray(ex["edge_off"]).astype(np.int64)) > 32).any()
    _check_batches(ctx, pq.SEARCH_PLAIN, 8, want, [0, 257, 1000], "pool")
    # buffers reserved ahead of the search (as construct_index does): same graph
    ctx.reserve_search(8)
    _check_batches(ctx, pq.SEARCH_PLAIN, 8, want, [0], "reserved")


def test_export_edges_buffer_too_small(ctx):
    rng = np.random.default_rng(44)
    codes, off, _ = _random_reads(rng, 8000, 3000, 80, 80, err=0.003)
    ctx.upload(codes, off, 41)
    ctx.build_index()
    want = _after_search(ctx, pq.SEARCH_PLAIN, 8)
    n = ctx.n
    E = len(want[1])
    assert E > 100
    ctx.set_search_batch(700)
    nb = ctx.host_buffer(n * 8)
    eb = ctx.host_buffer((E // 2) * 16)
    with pytest.raises(pq.PqError):
        ctx.find_overlaps_export(nb, eb, 8)
    # the node records arrived; the graph stays on the device for the export
    assert np.array_equal(nb.array, want[0])
    assert ctx.graph_size() == (n, E)
    big = ctx.host_buffer(E * 16)
    assert ctx.graph_export(None, big, 8) == want[2]
    assert np.array_equal(_edges(big, E, 8), want[1])
    for b in (nb, eb, big):
        b.free()
    ctx.set_search_batch(0)
