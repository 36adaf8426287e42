"""Device index build (SA, BWT, occurrence table, reads_index, C) against the
reference fixtures and the oracle; size-independent properties at scale."""
import json
import os

import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from conftest import golden_cases, load_golden
from test_oracle import _random_reads

pytestmark = pytest.mark.gpu
CASES = sorted(golden_cases())
with open(os.path.join(os.path.dirname(__file__), "golden", "slots.json")) as _f:
    SLOTS = json.load(_f)


def _check_against_oracle(ctx, codes, off, t):
    ix = orc.build_index(codes, off, t, with_shortcut=False)
    Cc, sfx, pfx = ctx.counts()
    assert np.array_equal(Cc, ix["C"]) and np.array_equal(sfx, ix["suffix_count"])
    assert np.array_equal(pfx, ix["prefix_count"])
    for d in (0, 1):
        sa = ctx.sa(d)
        assert np.array_equal(sa, ix[d]["sa"]), ("sa", d, orc.check_sa(ix[d]["text"], off, sa))
        assert np.array_equal(ctx.bwt(d), ix[d]["bwt"]), ("bwt", d)
        assert np.array_equal(ctx.reads_index(d), ix[d]["reads_index"]), ("reads_index", d)
        bm, sm = ctx.occ_ref(d, 8)
        assert np.array_equal(bm, ix[d]["bitmap"]) and np.array_equal(sm, ix[d]["sample"])
        assert ctx.slot_sizes(d) == orc.slot_sizes(ix[d]["text"], off), ("slots", d)
        lens = np.diff(off.astype(np.int64))
        if len(lens) and lens.min() == lens.max():
            # interval shortcuts at the depths the reference's sort records them
            L = int(lens[0])
            want = orc.shortcut(ix[d]["text"], len(lens), L, t, ix[d]["sa"], ix["C"])
            got = ctx.shortcut(d)
            for k, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a.astype(np.uint64), b.astype(np.uint64)), ("shortcut", d, k)
    return ix


@pytest.mark.parametrize("name", CASES)
def test_index_matches_reference_fixture(ctx, name):
    G = load_golden(name, "u32")
    meta = golden_cases()[name]
    n = int(G["n"])
    off = np.arange(n + 1, dtype=np.uint64) * np.uint64(meta["L"])
    ctx.upload(G["reads"], off, meta["t"])
    ctx.build_index()
    Cc, sfx, pfx = ctx.counts()
    assert np.array_equal(Cc, G["C"]) and np.array_equal(sfx, G["suffix_count"])
    assert np.array_equal(pfx, G["prefix_count"])
    for d in (0, 1):
        bm, sm = ctx.occ_ref(d, 4)
        assert np.array_equal(bm, G[f"bitmap{d}"]), d
        assert np.array_equal(sm.astype(np.uint64), G[f"sample{d}"]), d
        assert np.array_equal(ctx.reads_index(d), G[f"reads_index{d}"]), d
        # first-triplet slot sizes as the reference's construct_index sets them
        size, offs, mx = ctx.slot_sizes(d)
        tag = "fw" if d == 0 else "bw"
        assert (size, offs, mx) == (SLOTS[name]["u32"][f"size_{tag}"],
                                    SLOTS[name]["u32"][f"offset_{tag}"],
                                    SLOTS[name]["u32"][f"max_size_{tag}"]), d
        # interval shortcuts: the reference's own start / size / depth per read
        st, sz, dp = ctx.shortcut(d)
        assert np.array_equal(st, G[f"istart{d}"]) and np.array_equal(sz, G[f"iend{d}"])
        assert np.array_equal(dp.astype(np.uint64), G[f"idepth{d}"]), d


@pytest.mark.parametrize("kind,seed", [("fixed20", 1), ("fixed50", 2), ("fixed101", 3),
                                       ("varlen", 4), ("varlen", 5), ("ties", 6),
                                       ("long250", 7), ("deep", 8)])
def test_index_matches_oracle(ctx, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "ties":          # tiny genome, many identical suffixes
        codes, off, _ = _random_reads(rng, 70, 3000, 30, 30)
    elif kind == "deep":        # ~130x coverage: bucket runs of about 65..512 suffixes
        codes, off, _ = _random_reads(rng, 3000, 8000, 50, 50, err=0.002)
    elif kind == "varlen":
        codes, off, _ = _random_reads(rng, 20000, 3000, 36, 250, err=0.003)
    elif kind == "long250":
        codes, off, _ = _random_reads(rng, 30000, 1500, 250, 250, err=0.001)
    else:
        L = int(kind[5:])
        codes, off, _ = _random_reads(rng, 40000, 6000, L, L, err=0.002)
    ctx.upload(codes, off, 15)
    ctx.build_index()
    _check_against_oracle(ctx, codes, off, 15)


def test_index_properties_config1_scale(ctx):
    # the config-1 workload in full: 400k reads of 50 bp, 20M suffixes per direction
    cfg = pq.synth_preset(1)
    codes, off = pq.synth_reads(cfg)
    ctx.upload(codes, off, 31)
    ctx.build_index()
    Cc, sfx, pfx = orc.counts(codes, off)
    C2, s2, p2 = ctx.counts()
    assert np.array_equal(C2, Cc) and np.array_equal(s2, sfx) and np.array_equal(p2, pfx)
    n = len(off) - 1
    for d in (0, 1):
        dt = orc.dir_text(codes, off, d)
        sa = ctx.sa(d)
        assert orc.check_sa(dt, off, sa) == -1, d
        b = ctx.bwt(d)
        bo, ri = orc.bwt(dt, off, sa)
        assert np.array_equal(b, bo)
        assert np.array_equal(ctx.reads_index(d), ri)
        assert np.array_equal(np.sort(ri), np.arange(n, dtype=np.uint64))
