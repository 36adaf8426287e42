"""Exact part of the learned index: the table of distinct depth-D prefixes
(key, first suffix-array row, row count) against a host computation from the
oracle's suffix array, and identical on repeated builds."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu


def _host_table(codes, off, d, D):
    """Distinct D-code prefixes of the suffixes with >= D codes, in SA order, as
    (keys as Python ints, first rows, row counts)."""
    dt = orc.dir_text(codes, off, d).astype(np.uint64) - np.uint64(1)
    sa = orc.suffix_array(orc.dir_text(codes, off, d), off).astype(np.int64)
    ends = np.asarray(off[1:], np.int64)
    rem = ends[np.searchsorted(ends, sa, side="right")] - sa
    idx = np.nonzero(rem >= D)[0]
    g = sa[idx]
    hi = np.zeros(len(g), np.uint64)
    lo = np.zeros(len(g), np.uint64)
    for j in range(D):   # code j lands at bits 127-2j..126-2j of the 128-bit key
        c = dt[g + j]
        if j < 32:
            hi |= c << np.uint64(62 - 2 * j)
        else:
            lo |= c << np.uint64(62 - 2 * (j - 32))
    lo |= np.uint64(1)
    new = np.ones(len(g), bool)
    new[1:] = (hi[1:] != hi[:-1]) | (lo[1:] != lo[:-1])
    starts = np.nonzero(new)[0]
    first = idx[starts]
    last = np.append(starts[1:], len(g)) - 1
    rows = idx[last] - first + 1
    # a run of equal keys is contiguous in the SA
    assert np.array_equal(last - starts, rows - 1)
    keys = [(int(a) << 64) | int(b) for a, b in zip(hi[starts].tolist(), lo[starts].tolist())]
    return keys, first.tolist(), rows.tolist()


@pytest.mark.parametrize("L,t,seed,varlen", [(100, 63, 12, False), (100, 31, 11, False),
                                             (60, 17, 3, False), (0, 31, 5, True)])
def test_learned_table_matches_host(ctx, L, t, seed, varlen):
    rng = np.random.default_rng(seed)
    if varlen:
        codes, off, _ = _random_reads(rng, 6000, 600, 20, 120, err=0.003)
    else:
        codes, off, _ = _random_reads(rng, 20000, int(20 * 20000 / L), L, L, err=0.003)
    ctx.upload(codes, off, t)
    ctx.build_index()
    D = min(63, t)
    want = [_host_table(codes, off, d, D) for d in (0, 1)]
    for rep in range(3):
        ctx.learned_train(samples=1 << 14, batch=1024, epochs=1)
        for d in (0, 1):
            keys, first, rows = ctx.learned_table(d)
            got = [(int(a) << 64) | int(b) for a, b in keys.tolist()]
            wk, wf, wr = want[d]
            assert len(got) == len(wk), (rep, d, len(got), len(wk))
            assert got == wk, (rep, d)
            assert first.tolist() == wf and rows.tolist() == wr, (rep, d)
