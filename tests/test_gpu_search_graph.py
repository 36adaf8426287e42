"""Search run as one captured HIP graph: every replay gives the eager search's
graph, with eager work (other searches, index rebuilds, readbacks) between the
replays, and the graph is recaptured when its launch arguments change (other
reads, another extraction shape, a retrained learned index)."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_oracle import _random_reads
from test_gpu_search import _identical_copies, _same_graph

pytestmark = pytest.mark.gpu

GRAPH = pq.SEARCH_GRAPH


def _eager(ctx, mode):
    ctx.find_overlaps(mode)
    return ctx.graph()


@pytest.mark.parametrize("lmin,lmax,t,seed", [(100, 100, 31, 41), (20, 250, 31, 42),
                                              (100, 100, 63, 43)])
def test_plain_graph_replays_match_eager(ctx, lmin, lmax, t, seed):
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 20000, 3000, lmin, lmax, err=0.003)
    ctx.upload(codes, off, t)
    ctx.build_index()
    want = _eager(ctx, pq.SEARCH_PLAIN)
    if lmin != lmax:
        _same_graph(want, orc.search_explicit(codes, off, t), "eager uneven")
    for i in range(4):
        ctx.find_overlaps(pq.SEARCH_PLAIN | GRAPH)
        _same_graph(ctx.graph(), want, f"replay {i}")
        # eager work on the same stream and scratch buffers between replays
        if i == 1:
            ctx.build_index()
        _same_graph(_eager(ctx, pq.SEARCH_PLAIN), want, f"eager after replay {i}")


def test_graph_recaptured_for_new_arguments(ctx):
    rng = np.random.default_rng(44)
    cases = [_random_reads(rng, 10000, 1500, 80, 80, err=0.003)[:2] + (31,),
             _identical_copies(), _identical_copies(40),
             _random_reads(rng, 30000, 5000, 30, 150, err=0.003)[:2] + (25,)]
    for codes, off, t in cases:
        ctx.upload(codes, off, t)
        ctx.build_index()
        want = orc.search_explicit(codes, off, t)
        for shape in (0, 1, 2):
            ctx.set_tuning(shape)
            for r in range(2):
                ctx.find_overlaps(pq.SEARCH_PLAIN | GRAPH)
                _same_graph(ctx.graph(), want, f"n={len(off) - 1} t={t} shape {shape} rep {r}")
    ctx.set_tuning(0)


def test_learned_graph_replays_match_eager(ctx):
    rng = np.random.default_rng(45)
    codes, off, _ = _random_reads(rng, 20000, 4000, 100, 100, err=0.003)
    ctx.upload(codes, off, 63)
    ctx.build_index()
    want = _eager(ctx, pq.SEARCH_PLAIN)
    for seed in (1, 2):
        # each training gives a new model: the graph is captured again
        ctx.learned_train(samples=1 << 16, batch=1024, epochs=2, seed=seed)
        for i in range(3):
            ctx.find_overlaps(pq.SEARCH_LEARNED | GRAPH)
            _same_graph(ctx.graph(), want, f"learned replay {i}")
            st = ctx.learned_stats()
            assert st["lookups"][0] > 0 and st["fallback"] == 0
            _same_graph(_eager(ctx, pq.SEARCH_LEARNED), want, f"learned eager {i}")
            _same_graph(_eager(ctx, pq.SEARCH_PLAIN), want, f"plain eager {i}")


def test_graph_mode_rejects_unknown_modes(ctx):
    rng = np.random.default_rng(46)
    codes, off, _ = _random_reads(rng, 2000, 200, 50, 50)
    ctx.upload(codes, off, 20)
    ctx.build_index()
    with pytest.raises(pq.PqError):
        ctx.find_overlaps(7 | GRAPH)
    with pytest.raises(pq.PqError):
        ctx.find_overlaps(pq.SEARCH_LEARNED | GRAPH)   # not trained for these reads
