"""Learned index: device training kernels against the CPU oracle's gradients,
learned-seeded search identical to the plain FM-index search, and what the
trained network gains over an untrained one and over the plain search."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from test_oracle import _random_reads
from test_gpu_search import _identical_copies, _same_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,h,n,seed", [(21, 64, 4096, 1), (21, 64, 100, 2), (7, 13, 333, 3),
                                        (1, 1, 64, 4)])
def test_mlp_gradients_match_oracle(ctx, k, h, n, seed):
    rng = np.random.default_rng(seed)
    P = h * k + 2 * h + 1
    params = np.concatenate([rng.uniform(-1, 1, h * k) / np.sqrt(k), rng.uniform(-0.3, 0.3, h),
                             rng.uniform(-1, 1, h) / np.sqrt(h), [0.5]]).astype(np.float32)
    assert params.size == P
    xs = rng.integers(0, 1 << (2 * k), n, dtype=np.uint64)
    ys = rng.uniform(0, 1, n).astype(np.float32)
    g, loss = ctx.mlp_gradients(k, h, params, xs, ys)
    want_loss, want = orc.mlp_loss_grad(k, h, params.astype(np.float64), xs,
                                        ys.astype(np.float64))
    # FP32 kernels against the double oracle: sums of n terms of O(1)
    tol = 2e-5 * max(1.0, np.abs(want).max())
    assert np.abs(g - want).max() <= tol, (np.abs(g - want).max(), tol)
    assert abs(loss - want_loss) <= 1e-5 * max(1.0, want_loss)


@pytest.mark.parametrize("t", [63, 17])
def test_mlp_forward_mfma_matches_oracle(ctx, t):
    # the network as the search evaluates it (matrix cores, 64 keys per wave) and
    # in scalar form, against the oracle's double forward pass on the trained
    # parameters, at prefix values of the trained range and at random ones
    rng = np.random.default_rng(t)
    codes, off, _ = _random_reads(rng, 20000, 4000, 100, 100, err=0.003)
    ctx.upload(codes, off, t)
    ctx.build_index()
    ctx.learned_train(**SMALL)
    st = ctx.learned_stats()
    k = st["k"]
    for d in (0, 1):
        params = ctx.learned_params(d)
        h = (params.size - 1) // (k + 2)
        keys = ctx.learned_table(d)[0]
        # the k-code prefixes of table keys (first code in the top bits of the
        # high half), and random values
        top = keys[:, 0] >> np.uint64(64 - 2 * k)
        xs = np.concatenate([top[rng.integers(0, len(top), 3000)],
                             rng.integers(0, 1 << (2 * k), 1000, dtype=np.uint64)])
        want = orc.mlp_forward(k, h, params.astype(np.float64), xs)
        # FP32 over 21 inputs and 64 hidden units: sums of O(1) terms
        scale = 1.0 + np.abs(params).max() * (k + h)
        for mfma in (True, False):
            got = ctx.mlp_eval(d, xs, mfma)
            err = np.abs(got.astype(np.float64) - want).max()
            assert err <= 1e-6 * scale, (d, mfma, err)
        # a planted error in the oracle's output bias shows beyond the tolerance
        bad = params.astype(np.float64)
        bad[-1] += 1e-3
        assert np.abs(ctx.mlp_eval(d, xs, True) - orc.mlp_forward(k, h, bad, xs)).min() > \
            1e-6 * scale


def _search(ctx, codes, off, t, mode, **cfg):
    ctx.upload(codes, off, t)
    ctx.build_index()
    if mode == pq.SEARCH_LEARNED:
        ctx.learned_train(**cfg)
    ctx.find_overlaps(mode)
    return ctx.graph(), ctx.timing()


SMALL = dict(samples=1 << 16, batch=1024, epochs=3)


@pytest.mark.parametrize("L,t,seed", [(100, 31, 11), (100, 63, 12), (50, 31, 13), (60, 17, 14)])
def test_learned_search_equals_plain(ctx, L, t, seed):
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 20000, int(20 * 20000 / L), L, L, err=0.003)
    plain, tp = _search(ctx, codes, off, t, pq.SEARCH_PLAIN)
    learned, tl = _search(ctx, codes, off, t, pq.SEARCH_LEARNED, **SMALL)
    _same_graph(learned, plain, f"learned L={L} t={t}")
    st = ctx.learned_stats()
    assert st["depth"] == min(63, t) and st["k"] == min(21, st["depth"])
    assert st["lookups"][0] > 0 and st["lookups"][1] > 0 and st["fallback"] == 0
    # the seeded search evaluates fewer rank positions than the plain one
    assert tl["find_positions"] < tp["find_positions"]


def test_seed_threshold_does_not_change_the_graph(ctx):
    # the seed kernel finishes chains of up to 0, 8 or 32 rows at the seed depth;
    # the search kernel, verifying by its own threshold, does the rest
    rng = np.random.default_rng(17)
    codes, off, _ = _random_reads(rng, 20000, 4000, 100, 100, err=0.003)
    want, _ = _search(ctx, codes, off, 31, pq.SEARCH_PLAIN)
    ctx.learned_train(**SMALL)
    try:
        for seed_rows in (0, 8, 32):
            for verify in (0, 32):
                ctx.set_seed_verify(seed_rows)
                ctx.set_verify(verify)
                ctx.find_overlaps(pq.SEARCH_LEARNED)
                _same_graph(ctx.graph(), want, f"seed {seed_rows} verify {verify}")
        with pytest.raises(pq.PqError):
            ctx.set_seed_verify(33)
    finally:
        ctx.set_seed_verify(32)
        ctx.set_verify(32)


@pytest.mark.parametrize("seed", [21, 22])
def test_learned_search_uneven_lengths(ctx, seed):
    # reads of 20..250 codes: reads not longer than the seed depth keep the
    # plain start; the edge set must equal the explicit-string oracle's
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 8000, 1500, 20, 250, err=0.003)
    g, _ = _search(ctx, codes, off, 31, pq.SEARCH_LEARNED, **SMALL)
    _same_graph(g, orc.search_explicit(codes, off, 31), "learned uneven")


def test_learned_search_edge_cases(ctx):
    for codes, off, t in [
        (np.array([1, 2, 3, 4] * 10, np.uint8), np.array([0, 40], np.uint64), 12),
        _identical_copies(),
        _identical_copies(40),
    ]:
        g, _ = _search(ctx, codes, off, t, pq.SEARCH_LEARNED, samples=1024, batch=64, epochs=1)
        _same_graph(g, orc.search_explicit(codes, off, t), "learned edge case")


def test_learned_search_repeats(ctx):
    # a repeat-rich genome (deep intervals at the seed depth)
    cfg = pq.synth_preset(2)
    cfg.genome_len = 200_000
    cfg.n_copies = 20
    codes, off = pq.synth_reads(cfg)
    plain, _ = _search(ctx, codes, off, 63, pq.SEARCH_PLAIN)
    learned, _ = _search(ctx, codes, off, 63, pq.SEARCH_LEARNED, **SMALL)
    _same_graph(learned, plain, "learned repeats")


def test_trained_network_gains_over_untrained(ctx):
    # same reads, same leaves and correction search: the fitted MLP routes keys
    # to leaves whose pieces predict the table entry closely; the untrained one
    # (random parameters) does not; Adam alone already gains over it
    rng = np.random.default_rng(31)
    codes, off, _ = _random_reads(rng, 200000, 40000, 100, 100, err=0.0)
    ctx.upload(codes, off, 31)
    ctx.build_index()
    ctx.find_overlaps(pq.SEARCH_PLAIN)
    plain = ctx.graph()
    res = {}
    for name, cfg in [("untrained", dict(epochs=0, closed_form=0)),
                      ("adam", dict(epochs=4, closed_form=0, lr=1e-2)),
                      ("fitted", dict())]:
        ctx.learned_train(samples=1 << 18, batch=4096, **cfg)
        ctx.find_overlaps(pq.SEARCH_LEARNED)
        _same_graph(ctx.graph(), plain, name)
        st = ctx.learned_stats()
        assert st["fallback"] == 0
        res[name] = (np.mean(st["mean_window"]), sum(st["probes"]) / sum(st["lookups"]),
                     max(st["train_loss"]))
    print("mean window, probes per lookup, loss:", res)
    assert res["adam"][2] < res["untrained"][2] / 10
    assert res["adam"][0] < res["untrained"][0]
    assert res["fitted"][2] < res["adam"][2]
    assert res["fitted"][0] * 4 < res["untrained"][0]
    assert res["fitted"][1] < res["untrained"][1]
