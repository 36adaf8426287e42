"""Device overlap search (candidate intervals + irreducible edges) against the
reference fixtures, the oracle's restatement of string_graph.c, and the
explicit-string oracle for reads of unequal length."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from conftest import golden_cases, load_golden
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu
CASES = sorted(golden_cases())


def _same_graph(g, want, where=""):
    for k in ("overlaps_B", "overlaps_F", "num_B", "num_F", "edge_off"):
        a, b = np.asarray(g[k]).astype(np.int64), np.asarray(want[k]).astype(np.int64)
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0][:5]
            raise AssertionError(f"{where} {k} differs at {bad.tolist()}: {a[bad]} vs {b[bad]}")
    for k in ("dest", "len", "type"):
        a, b = np.asarray(g[k]).astype(np.int64), np.asarray(want[k]).astype(np.int64)
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0][:5]
            raise AssertionError(f"{where} edge {k} differs at {bad.tolist()}")


@pytest.mark.parametrize("name", CASES)
def test_graph_matches_reference_fixture(ctx, name):
    G = load_golden(name, "u32")
    meta = golden_cases()[name]
    n = int(G["n"])
    off = np.arange(n + 1, dtype=np.uint64) * np.uint64(meta["L"])
    ctx.upload(G["reads"], off, meta["t"])
    ctx.build_index()
    ctx.find_overlaps()
    g = ctx.graph()
    nf = G["node_fields"]
    assert np.array_equal(g["num_B"], nf[:, 0]) and np.array_equal(g["num_F"], nf[:, 1])
    assert np.array_equal(g["edge_off"], G["edge_off"])
    # node records computed on the device: fields num_B, num_F, B_type, F_type,
    # status, marked of the reference's nodes, and its average overlap
    rec, len_sum = ctx.graph_nodes()
    for col, k in ((0, "num_B"), (1, "num_F"), (4, "B_type"), (5, "F_type"), (6, "status"),
                   (7, "marked")):
        assert np.array_equal(rec[k].astype(np.int64), nf[:, col].astype(np.int64)), k
    assert len_sum / len(g["dest"]) == float(G["average_overlap"])
    E = np.stack([g["dest"], g["len"], g["type"]], 1).astype(np.uint64)
    assert np.array_equal(E, G["edges"])
    # the one-call export into page-locked memory: the same node records, and the
    # edges in the edge_t layout of both index widths (dest, overlap_len, type)
    nb = ctx.host_buffer(n * 8)
    for ib, dt in ((4, np.dtype([("dest", "<u4"), ("len", "<u2"), ("type", "u1"), ("pad", "u1")])),
                   (8, np.dtype([("dest", "<u8"), ("len", "<u2"), ("type", "u1"),
                                 ("pad", "V5")]))):
        eb = ctx.host_buffer(len(g["dest"]) * dt.itemsize)
        assert ctx.graph_export(nb, eb, ib) == len_sum
        assert np.array_equal(nb.array.view(ctx.NODE_DTYPE), rec)
        got = eb.array.view(dt)
        assert np.array_equal(got["dest"].astype(np.uint64), G["edges"][:, 0])
        assert np.array_equal(got["len"].astype(np.uint64), G["edges"][:, 1])
        assert np.array_equal(got["type"].astype(np.uint64), G["edges"][:, 2])
        eb.free()
    nb.free()
    ex = orc.search_explicit(G["reads"], off, meta["t"])
    _same_graph(g, ex, name)


@pytest.mark.parametrize("L,t,seed", [(50, 31, 1), (100, 63, 2), (80, 25, 3), (150, 63, 4)])
def test_graph_matches_oracle_fixed_length(ctx, L, t, seed):
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 5000, int(25 * 5000 / L), L, L, err=0.004)
    ctx.upload(codes, off, t)
    ctx.build_index()
    ctx.find_overlaps()
    g = ctx.graph()
    ix = orc.build_index(codes, off, t)
    _same_graph(g, orc.search_fm(ix), f"L={L}")


@pytest.mark.parametrize("seed", [5, 6])
def test_graph_matches_oracle_uneven_lengths(ctx, seed):
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 8000, 1500, 36, 250, err=0.003)
    ctx.upload(codes, off, 31)
    ctx.build_index()
    ctx.find_overlaps()
    _same_graph(ctx.graph(), orc.search_explicit(codes, off, 31), "uneven")


def test_graph_with_many_candidates(ctx):
    # a short genome sampled deeply: hundreds of candidates per read side, so the
    # extraction spills past its LDS capacity into the global-memory path
    rng = np.random.default_rng(9)
    codes, off, _ = _random_reads(rng, 400, 4000, 40, 40, err=0.0)
    ctx.upload(codes, off, 21)
    ctx.build_index()
    ctx.find_overlaps()
    g = ctx.graph()
    ex = orc.search_explicit(codes, off, 21)
    assert max(ex["overlaps_F"].max(), ex["overlaps_B"].max()) > 128
    _same_graph(g, ex, "deep")


@pytest.mark.parametrize("G,n,lmin,lmax,err,seed", [(2000, 6000, 60, 120, 0.01, 31),
                                                    (600, 5000, 40, 90, 0.005, 32)])
def test_graph_with_many_candidates_and_errors(ctx, G, n, lmin, lmax, err, seed):
    # deep coverage with errors: hundreds of candidates per read side whose
    # extension keys are prefixes of one another or equal; the one-wave tiers
    # sort the items and read the emission rules off adjacent common prefixes
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, G, n, lmin, lmax, err=err)
    ctx.upload(codes, off, 21)
    ctx.build_index()
    ex = orc.search_explicit(codes, off, 21)
    assert max(ex["overlaps_F"].max(), ex["overlaps_B"].max()) > 128
    for shape in range(3):
        ctx.set_tuning(shape)
        ctx.find_overlaps()
        _same_graph(ctx.graph(), ex, f"deep with errors, shape {shape}")
    ctx.set_tuning(0)


def _identical_copies(copies=3):
    # copies of one read overlapped by another read: the copies form one
    # candidate group whose members are emitted in read order (more than 16 of
    # them take the shared edge pool instead of the per-read slot)
    g = np.random.default_rng(3).integers(1, 5, 30).astype(np.uint8)
    reads = [g[0:16], g[0:16], g[8:24]] + [g[0:16]] * (copies - 2) + [g[14:30]]
    off = np.zeros(len(reads) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return np.concatenate(reads), off, 6


def test_edge_buffer_exhaustion_recovers(ctx):
    # read sides with more than 16 edges share one pool; made too small, the
    # search counts what it needed, runs again with the pool at that size and
    # gives the oracle's graph (eager and graph mode)
    codes, off, t = _identical_copies(40)
    ctx.upload(codes, off, t)
    ctx.build_index()
    want = orc.search_explicit(codes, off, t)
    assert max(want["num_B"].max(), want["num_F"].max()) > 16
    try:
        for mode in (pq.SEARCH_PLAIN, pq.SEARCH_PLAIN | pq.SEARCH_GRAPH):
            ctx.set_edge_capacity(8)
            ctx.find_overlaps(mode)
            _same_graph(ctx.graph(), want, f"edge pool of 8, mode {mode:#x}")
    finally:
        ctx.set_edge_capacity(0)
    for mode in (pq.SEARCH_PLAIN, pq.SEARCH_PLAIN | pq.SEARCH_GRAPH):
        ctx.find_overlaps(mode)
        _same_graph(ctx.graph(), want, f"default edge pool, mode {mode:#x}")


def _check_every_shape_and_kernel(ctx, want, where):
    for shape in range(3):
        ctx.set_tuning(shape)
        for kernel in (1, 2, 3, 4):
            ctx.set_find_chains(kernel)
            for mode in (pq.SEARCH_PLAIN, pq.SEARCH_PLAIN | pq.SEARCH_GRAPH):
                ctx.find_overlaps(mode)
                _same_graph(ctx.graph(), want, f"{where}: shape {shape}, kernel {kernel}, "
                                                f"mode {mode:#x}")
    ctx.set_tuning(0)
    ctx.set_find_chains(0)


def test_global_tier_more_than_512_items(ctx):
    # 600 copies of one read overlap a second read by 8 codes: that read's B side
    # holds 600 items, beyond the one-wave LDS tier (512), so the global-memory
    # tier sorts it; all 600 are emitted (more than 16: the shared pool)
    codes, off, t = _identical_copies(600)
    ctx.upload(codes, off, t)
    ctx.build_index()
    want = orc.search_explicit(codes, off, t)
    assert want["overlaps_B"].max() > 512
    _check_every_shape_and_kernel(ctx, want, "600 items")


def _one_oversized_side(copies_fb=300, copies_ff=100, L=200, t=21, seed=11):
    """One read n of L codes and, for every overlap o in [t, L), copies of the read
    that starts o codes before n's end and ends one code after it (FB candidates
    of n), and of its reverse complement (FF candidates); the extra code is one
    of two alternatives.  n's F side holds (L - t) * (copies_fb + copies_ff) items;
    every other read has n as its only candidate (they all end at the same place,
    so none overlaps another), which keeps the oracle quick."""
    rng = np.random.default_rng(seed)
    g = rng.integers(1, 5, L).astype(np.uint8)
    reads = [g]
    for o in range(t, L):
        for c in range(copies_fb + copies_ff):
            m = np.concatenate([g[L - o:], np.array([1 if c % 3 else 4], np.uint8)])
            reads.append(m if c < copies_fb else (5 - m[::-1]).astype(np.uint8))
    order = rng.permutation(len(reads) - 1) + 1
    reads = [reads[0]] + [reads[i] for i in order]
    off = np.zeros(len(reads) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return np.concatenate(reads), off, t


def test_read_side_beyond_global_tier(ctx):
    # more than 65536 candidate reads on one side (the global tier's limit) but
    # a few hundred edges: the side is sorted by the whole GPU and gives the
    # oracle's graph (graph mode meets it and runs the search eagerly)
    codes, off, t = _one_oversized_side()
    want = orc.search_explicit(codes, off, t)
    assert want["overlaps_F"][0] > 65536
    assert 16 < want["num_F"][0] < 65535
    ctx.upload(codes, off, t)
    ctx.build_index()
    # the side overflows every tier whatever the first tier's shape: the search
    # kernels eagerly, and graph mode once
    for kernel in (1, 2, 3, 4):
        ctx.set_find_chains(kernel)
        ctx.find_overlaps()
        _same_graph(ctx.graph(), want, f"oversized side, kernel {kernel}")
    ctx.set_find_chains(0)
    ctx.find_overlaps(pq.SEARCH_PLAIN | pq.SEARCH_GRAPH)
    _same_graph(ctx.graph(), want, "oversized side, graph mode")


def test_graph_edge_cases(ctx):
    # one read; reads shorter than t; identical reads (no duplicate removal here)
    for codes, off, t in [
        (np.array([1, 2, 3, 4] * 10, np.uint8), np.array([0, 40], np.uint64), 12),
        (np.array([1, 2, 3, 4, 1, 2, 3, 4, 1, 2], np.uint8), np.array([0, 5, 10], np.uint64), 12),
        _identical_copies(),
        _identical_copies(40),
    ]:
        ctx.upload(codes, off, t)
        ctx.build_index()
        ctx.find_overlaps()
        _same_graph(ctx.graph(), orc.search_explicit(codes, off, t), "edge case")


def test_every_extraction_shape_gives_the_same_graph(ctx):
    # the autotuner's first-tier shapes (8, 16, 32 lanes per read) only change
    # where reads are processed; reads with more candidates than a shape holds
    # fall through to the wider tiers
    rng = np.random.default_rng(17)
    codes, off, _ = _random_reads(rng, 3000, 2000, 30, 120, err=0.002)
    ctx.upload(codes, off, 21)
    ctx.build_index()
    want = orc.search_explicit(codes, off, 21)
    assert max(want["overlaps_F"].max(), want["overlaps_B"].max()) > 16
    for shape in range(3):
        ctx.set_tuning(shape)
        ctx.find_overlaps()
        _same_graph(ctx.graph(), want, f"shape {shape}")
    for kernel in (0, 1, 2, 3, 4):
        ctx.set_find_chains(kernel)
        ctx.find_overlaps()
        _same_graph(ctx.graph(), want, f"search kernel {kernel}")
    t = ctx.autotune()
    assert 0 <= t["shape"] < 3 and len(t["extract_ms"]) == 3
    assert t["find_chains"] in (1, 2, 3, 4) and min(t["find_ms"]) > 0
    _same_graph(ctx.graph(), want, "autotuned")
    ctx.set_tuning(0)
    ctx.set_find_chains(0)


@pytest.mark.parametrize("lmin,lmax,t,seed", [(36, 250, 31, 41), (100, 100, 63, 42)])
def test_search_work_counts_match_oracle(ctx, lmin, lmax, t, seed):
    # the rank positions, occurrence lines and verified rows the search kernels
    # report (the roofline's algorithmic bytes) equal the oracle's count of the
    # same chains, for each search kernel, from depth 1 and from the learned
    # seeds' depth, with verification of narrow intervals and without (the
    # learned seeds' narrow intervals: the rows their scan reads)
    rng = np.random.default_rng(seed)
    codes, off, _ = _random_reads(rng, 6000, 2000, lmin, lmax, err=0.003)
    ctx.upload(codes, off, t)
    ctx.build_index()
    stride = 5   # every 5th suffix-array row sampled by the learned index
    ctx.learned_train(sample_stride=stride)
    D = ctx.learned_stats()["depth"]
    for vmax in (0, 32, 5):
        ctx.set_verify(vmax)
        ctx.set_seed_verify(vmax)   # the seed kernel's scan by the same threshold
        v = min(vmax, int(np.diff(off.astype(np.int64)).max()) - t)
        want = orc.chain_work(codes, off, 0, t, v)
        want_l = orc.chain_work(codes, off, D, t, v, stride)
        assert want_l[0] + want_l[2] < want[0] + want[2]
        for kernel in (1, 2, 3, 4):
            ctx.set_find_chains(kernel)
            for mode, w in ((pq.SEARCH_PLAIN, want), (pq.SEARCH_LEARNED, want_l)):
                ctx.find_overlaps(mode)
                tm = ctx.timing()
                assert (tm["find_positions"], tm["find_lines"],
                        tm["verify_rows"] + tm["seed_rows"]) == w, \
                    (vmax, kernel, mode)
    ctx.set_find_chains(0)
    ctx.set_verify(32)
    ctx.set_seed_verify(32)


@pytest.mark.parametrize("vmax", [0, 1, 7, 32])
def test_verification_gives_the_search_graph(ctx, vmax):
    # chains finished on the text give the backward search's candidates exactly:
    # uneven lengths, repeats with errors, identical reads, short t
    rng = np.random.default_rng(50 + vmax)
    cases = [_random_reads(rng, 3000, 2500, 36, 250, err=0.004)[:2] + (31,),
             _random_reads(rng, 800, 3000, 40, 90, err=0.01)[:2] + (21,),
             _random_reads(rng, 20000, 4000, 60, 60, err=0.003)[:2] + (17,),
             _identical_copies(40)]
    ctx.set_verify(vmax)
    ctx.set_seed_verify(vmax)
    try:
        for codes, off, t in cases:
            want = orc.search_explicit(codes, off, t)
            ctx.upload(codes, off, t)
            ctx.build_index()
            for kernel in (1, 2, 3, 4):
                ctx.set_find_chains(kernel)
                ctx.find_overlaps()
                _same_graph(ctx.graph(), want, f"vmax {vmax} kernel {kernel}")
            ctx.learned_train(samples=1 << 14, batch=1024, epochs=1)
            ctx.find_overlaps(pq.SEARCH_LEARNED)
            _same_graph(ctx.graph(), want, f"vmax {vmax} learned")
    finally:
        ctx.set_find_chains(0)
        ctx.set_verify(32)
        ctx.set_seed_verify(32)
