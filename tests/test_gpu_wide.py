"""64-bit rows and sliced suffix sorting -- the layout of texts of 2^32 codes or
more -- forced on small read sets: suffix arrays, BWT, occurrence table,
reads_index, learned table and overlap graphs equal the oracle's, whatever the
row width, the slice size and the table-build chunk."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_gpu_index import _check_against_oracle
from test_gpu_search import _identical_copies, _same_graph
from test_gpu_table import _host_table
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wctx():
    c = pq.Context()
    yield c
    c.close()


def _reads(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "ties":          # tiny genome: long runs of identical suffixes
        return _random_reads(rng, 70, 3000, 30, 30)[:2]
    if kind == "varlen":
        return _random_reads(rng, 20000, 3000, 36, 250, err=0.003)[:2]
    if kind == "long250":
        return _random_reads(rng, 30000, 1200, 250, 250, err=0.001)[:2]
    return _random_reads(rng, 40000, 6000, 50, 50, err=0.002)[:2]


@pytest.mark.parametrize("row_bytes,slice_,kind,seed", [
    (8, 0, "fixed50", 1), (8, 0, "varlen", 2), (8, 0, "ties", 3),
    (4, 5000, "fixed50", 4), (4, 5000, "ties", 5), (8, 5000, "varlen", 6),
    (8, 20000, "long250", 7), (8, 4000, "fixed50", 8), (4, 4000, "varlen", 9)])
def test_layout_index_matches_oracle(wctx, row_bytes, slice_, kind, seed):
    codes, off = _reads(kind, seed)
    N = int(off[-1])
    wctx.set_layout(row_bytes, slice_)
    wctx.upload(codes, off, 15)
    wctx.build_index()
    rb, slices = wctx.layout()
    assert rb == row_bytes
    assert slices == 1 if slice_ == 0 or N <= slice_ else slices >= (N + slice_ - 1) // slice_
    _check_against_oracle(wctx, codes, off, 15)


def test_slice_smaller_than_one_bucket_range_is_refused(wctx):
    # 70-code genome: one 6-code prefix range holds far more than 50 suffixes
    codes, off = _reads("ties", 10)
    wctx.set_layout(8, 50)
    wctx.upload(codes, off, 15)
    with pytest.raises(pq.PqError, match="bucket range"):
        wctx.build_index()


@pytest.mark.parametrize("row_bytes,slice_,lmin,lmax,t,seed", [
    (8, 0, 100, 100, 63, 21), (8, 4000, 100, 100, 31, 22), (8, 3000, 20, 250, 31, 23),
    (4, 3000, 60, 60, 17, 24)])
def test_layout_search_and_learned_match_oracle(wctx, row_bytes, slice_, lmin, lmax, t, seed):
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 12000, 2500, lmin, lmax, err=0.003)
    wctx.set_layout(row_bytes, slice_)
    wctx.upload(codes, off, t)
    wctx.build_index()
    want = orc.search_explicit(codes, off, t)
    wctx.find_overlaps(pq.SEARCH_PLAIN)
    _same_graph(wctx.graph(), want, "plain")
    # sampled keys every 7 rows
    wctx.learned_train(samples=1 << 14, batch=1024, epochs=1, sample_stride=7)
    D = min(63, t)
    for d in (0, 1):
        keys, first, rows = wctx.learned_table(d)
        wk, wf, wr = _host_table(codes, off, d, D, 7)
        assert [(int(a) << 64) | int(b) for a, b in keys.tolist()] == wk, d
        assert first.tolist() == wf and rows.tolist() == wr, d
    for mode in (pq.SEARCH_LEARNED, pq.SEARCH_LEARNED | pq.SEARCH_GRAPH,
                 pq.SEARCH_PLAIN | pq.SEARCH_GRAPH):
        wctx.find_overlaps(mode)
        _same_graph(wctx.graph(), want, f"mode {mode:#x}")
    assert wctx.learned_stats()["fallback"] == 0
    # every candidate-search kernel, plain and seeded, on these rows
    for kernel in (1, 2, 3, 4):
        wctx.set_find_chains(kernel)
        for mode in (pq.SEARCH_PLAIN, pq.SEARCH_LEARNED):
            wctx.find_overlaps(mode)
            _same_graph(wctx.graph(), want, f"kernel {kernel} mode {mode:#x}")
    wctx.set_find_chains(0)


def test_layout_edge_cases(wctx):
    for codes, off, t in [
        (np.array([1, 2, 3, 4] * 10, np.uint8), np.array([0, 40], np.uint64), 12),
        _identical_copies(),
        _identical_copies(40),
    ]:
        wctx.set_layout(8, 100)
        wctx.upload(codes, off, t)
        wctx.build_index()
        _check_against_oracle(wctx, codes, off, t)
        wctx.find_overlaps(pq.SEARCH_PLAIN)
        _same_graph(wctx.graph(), orc.search_explicit(codes, off, t), "wide edge case")


def test_repeat_families_wide_sliced(wctx, ctx):
    # the large workload's generator over a 1 Mbp genome (200k reads of 150,
    # 30M suffixes per direction): repeat families give many buckets too large
    # for one sort, refined level by level; 64-bit rows, slices of 6M suffixes
    cfg = pq.synth_preset(5)
    cfg.genome_len = 1_000_000
    codes, off = pq.synth_reads(cfg)
    wctx.set_layout(8, 6_000_000)
    wctx.upload(codes, off, 63)
    wctx.build_index()
    rb, slices = wctx.layout()
    assert rb == 8 and slices >= 5
    Cc, sfx, pfx = orc.counts(codes, off)
    C2, s2, p2 = wctx.counts()
    assert np.array_equal(C2, Cc) and np.array_equal(s2, sfx) and np.array_equal(p2, pfx)
    ctx.upload(codes, off, 63)   # default layout: 32-bit rows, one slice
    ctx.build_index()
    n = len(off) - 1
    for d in (0, 1):
        dt = orc.dir_text(codes, off, d)
        sa = wctx.sa(d)
        assert orc.check_sa(dt, off, sa) == -1, d
        assert np.array_equal(sa, ctx.sa(d)), d
        bo, ri = orc.bwt(dt, off, sa)
        assert np.array_equal(wctx.bwt(d), bo), d
        assert np.array_equal(wctx.reads_index(d), ri), d
        assert np.array_equal(np.sort(ri), np.arange(n, dtype=np.uint64))
    # the overlap graph of both layouts against the explicit-string oracle
    want = orc.search_explicit(codes, off, 63)
    for c, where in ((wctx, "64-bit rows, sliced"), (ctx, "32-bit rows")):
        c.find_overlaps(pq.SEARCH_PLAIN)
        _same_graph(c.graph(), want, f"repeat families, {where}")


@pytest.mark.parametrize("row_bytes", [8, 4])
def test_slices_made_in_one_pass(wctx, row_bytes):
    # 16 M suffixes per direction: 4 radix passes (even), so every slice's pairs
    # come from one pass over the text (the path of texts above 2^31 suffixes);
    # the suffix arrays, BWT and reads_index against the oracle's checks
    rng = np.random.default_rng(61 + row_bytes)
    codes, off, _ = _random_reads(rng, 2_000_000, 160_000, 100, 100, err=0.003)
    N = int(off[-1])
    wctx.set_layout(row_bytes, 5_000_000)
    wctx.upload(codes, off, 31)
    wctx.build_index()
    rb, slices = wctx.layout()
    assert rb == row_bytes and slices >= (N + 4_999_999) // 5_000_000
    n = len(off) - 1
    for d in (0, 1):
        dt = orc.dir_text(codes, off, d)
        sa = wctx.sa(d)
        assert orc.check_sa(dt, off, sa) == -1, d
        bo, ri = orc.bwt(dt, off, sa)
        assert np.array_equal(wctx.bwt(d), bo), d
        assert np.array_equal(wctx.reads_index(d), ri), d
        assert np.array_equal(np.sort(ri), np.arange(n, dtype=np.uint64)), d
    wctx.set_layout(0, 0)
