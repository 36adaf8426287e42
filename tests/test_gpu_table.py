"""Exact part of the learned index: the keys of every stride-th suffix-array row
(first D codes, codes past the read's end 0, bit 0 set for suffixes of at least
D codes) against a host computation from the oracle's suffix array, identical on
repeated trainings, for several strides."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu


def _host_table(codes, off, d, D, stride=8):
    """Keys (Python ints), rows and rows-per-sample of every stride-th SA row."""
    dt = orc.dir_text(codes, off, d).astype(np.uint64) - np.uint64(1)
    sa = orc.suffix_array(orc.dir_text(codes, off, d), off).astype(np.int64)
    N = len(sa)
    ends = np.asarray(off[1:], np.int64)
    rows = np.arange(0, N, stride)
    g = sa[rows]
    rem = ends[np.searchsorted(ends, g, side="right")] - g
    hi = np.zeros(len(g), np.uint64)
    lo = np.zeros(len(g), np.uint64)
    for j in range(D):   # code j lands at bits 127-2j..126-2j of the 128-bit key
        c = np.where(rem > j, dt[np.minimum(g + j, N - 1)], np.uint64(0)).astype(np.uint64)
        if j < 32:
            hi |= c << np.uint64(62 - 2 * j)
        else:
            lo |= c << np.uint64(62 - 2 * (j - 32))
    lo |= (rem >= D).astype(np.uint64)
    keys = [(int(a) << 64) | int(b) for a, b in zip(hi.tolist(), lo.tolist())]
    counts = np.minimum(stride, N - rows)
    return keys, rows.tolist(), counts.tolist()


@pytest.mark.parametrize("L,t,seed,varlen,stride", [(100, 63, 12, False, 8),
                                                    (100, 31, 11, False, 8),
                                                    (60, 17, 3, False, 5),
                                                    (0, 31, 5, True, 1),
                                                    (0, 63, 6, True, 16)])
def test_learned_table_matches_host(ctx, L, t, seed, varlen, stride):
    rng = np.random.default_rng(seed)
    if varlen:
        codes, off, _ = _random_reads(rng, 6000, 600, 20, 120, err=0.003)
    else:
        codes, off, _ = _random_reads(rng, 20000, int(20 * 20000 / L), L, L, err=0.003)
    ctx.upload(codes, off, t)
    ctx.build_index()
    D = min(63, t)
    want = [_host_table(codes, off, d, D, stride) for d in (0, 1)]
    for rep in range(3):
        ctx.learned_train(samples=1 << 14, batch=1024, epochs=1, sample_stride=stride)
        for d in (0, 1):
            keys, first, rows = ctx.learned_table(d)
            got = [(int(a) << 64) | int(b) for a, b in keys.tolist()]
            wk, wf, wr = want[d]
            assert len(got) == len(wk), (rep, d, len(got), len(wk))
            assert got == wk, (rep, d)
            assert first.tolist() == wf and rows.tolist() == wr, (rep, d)
            # the keys are non-decreasing along the suffix array
            assert all(a <= b for a, b in zip(got, got[1:]))
        # the seeded search gives the plain search's graph at this stride
        ctx.find_overlaps(pq.SEARCH_PLAIN)
        want_g = ctx.graph()
        ctx.find_overlaps(pq.SEARCH_LEARNED)
        got_g = ctx.graph()
        for k in want_g:
            assert np.array_equal(got_g[k], want_g[k]), (rep, k)


@pytest.mark.parametrize("L,t,seed,varlen,stride,slice_", [(100, 63, 12, False, 8, 0),
                                                           (100, 31, 11, False, 8, 40000),
                                                           (0, 31, 5, True, 1, 0),
                                                           (0, 63, 6, True, 16, 30000)])
def test_build_written_keys_match_host(ctx, L, t, seed, varlen, stride, slice_):
    # keys written by the index build's run sorts (pq_set_learned_keys), the rows
    # no first-level sort placed filled in after it; sliced sorts (row offsets per
    # slice) and 64-bit rows included
    rng = np.random.default_rng(seed)
    if varlen:
        codes, off, _ = _random_reads(rng, 6000, 600, 20, 120, err=0.003)
    else:
        codes, off, _ = _random_reads(rng, 20000, int(20 * 20000 / L), L, L, err=0.003)
    D = min(63, t)
    want = [_host_table(codes, off, d, D, stride) for d in (0, 1)]
    try:
        ctx.set_layout(8 if slice_ else 0, slice_)
        ctx.upload(codes, off, t)
        ctx.set_learned_keys(stride)
        ctx.build_index()
        ctx.learned_train(samples=1 << 14, batch=1024, epochs=1, sample_stride=stride)
        for d in (0, 1):
            keys, _, _ = ctx.learned_table(d)
            got = [(int(a) << 64) | int(b) for a, b in keys.tolist()]
            assert got == want[d][0], d
        ctx.find_overlaps(pq.SEARCH_PLAIN)
        want_g = ctx.graph()
        ctx.find_overlaps(pq.SEARCH_LEARNED)
        got_g = ctx.graph()
        for k in want_g:
            assert np.array_equal(got_g[k], want_g[k]), k
    finally:
        ctx.set_learned_keys(0)
        ctx.set_layout(0, 0)


def test_learned_keys_stride_must_be_a_power_of_two(ctx):
    with pytest.raises(pq.PqError):
        ctx.set_learned_keys(5)
    ctx.set_learned_keys(0)
