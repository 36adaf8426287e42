"""The oracle pinned: against the reference's own outputs (golden fixtures and,
where the reference programs are built, fresh runs), against brute force on tiny
inputs, and against its second, independent restatement of the search."""
import bisect
import os

import numpy as np
import pytest

import pq_oracle as orc
from conftest import golden_cases, load_golden

CASES = sorted(golden_cases())


def _off(n, L):
    return np.arange(n + 1, dtype=np.uint64) * np.uint64(L)


@pytest.fixture(scope="module")
def built():
    cache = {}

    def get(name, width):
        if (name, width) not in cache:
            G = load_golden(name, width)
            meta = golden_cases()[name]
            n = int(G["n"])
            off = _off(n, meta["L"])
            ix = orc.build_index(G["reads"], off, meta["t"])
            cache[(name, width)] = (G, meta, off, ix)
        return cache[(name, width)]
    return get


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", CASES)
def test_index_matches_reference(built, name, width):
    G, meta, off, ix = built(name, width)
    assert np.array_equal(ix["C"], G["C"])
    assert np.array_equal(ix["suffix_count"], G["suffix_count"])
    assert np.array_equal(ix["prefix_count"], G["prefix_count"])
    for d in (0, 1):
        assert np.array_equal(ix[d]["bitmap"], G[f"bitmap{d}"]), d
        assert np.array_equal(ix[d]["sample"], G[f"sample{d}"]), d
        assert np.array_equal(ix[d]["reads_index"], G[f"reads_index{d}"]), d
        st, en, dp = ix[d]["shortcut"]
        assert np.array_equal(st, G[f"istart{d}"]) and np.array_equal(en, G[f"iend{d}"])
        assert np.array_equal(dp, G[f"idepth{d}"])


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", CASES)
def test_search_matches_reference(built, name, width):
    G, meta, off, ix = built(name, width)
    fm = orc.search_fm(ix)
    ex = orc.search_explicit(G["reads"], off, meta["t"])
    nf = G["node_fields"]
    for g in (fm, ex):
        assert np.array_equal(g["num_B"], nf[:, 0]) and np.array_equal(g["num_F"], nf[:, 1])
        assert np.array_equal(g["edge_off"], G["edge_off"])
        E = np.stack([g["dest"], g["len"], g["type"]], 1).astype(np.uint64)
        assert np.array_equal(E, G["edges"])
    assert np.array_equal(fm["overlaps_B"], ex["overlaps_B"])
    assert np.array_equal(fm["overlaps_F"], ex["overlaps_F"])
    # SUSPECT marks follow from the candidate counts (string_graph.c:843-854)
    a, b = fm["overlaps_B"].astype(float), fm["overlaps_F"].astype(float)
    mx = np.maximum(a, b)
    ratio = np.where(mx == 0, 0.0, np.abs(a - b) / np.where(mx == 0, 1, mx))
    assert np.array_equal((ratio > 0.6).astype(np.uint64) * 4, nf[:, 7])
    tot = int(G["edge_off"][-1])
    if tot:
        assert np.isclose(G["edges"][:, 1].sum() / tot, float(G["average_overlap"]), rtol=0,
                          atol=0)


# ---- brute force on tiny inputs ------------------------------------------------------

def _random_reads(rng, G, n, lmin, lmax, err=0.0):
    genome = rng.integers(1, 5, G).astype(np.uint8)
    reads = []
    for _ in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        s = int(rng.integers(0, G - L + 1))
        r = genome[s:s + L].copy()
        if rng.random() < 0.5:
            r = (5 - r[::-1]).astype(np.uint8)
        flip = rng.random(L) < err
        r[flip] = ((r[flip] - 1 + rng.integers(1, 4, flip.sum())) % 4 + 1).astype(np.uint8)
        reads.append(r)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return np.concatenate(reads).astype(np.uint8), off, reads


@pytest.mark.parametrize("seed", range(6))
def test_sa_bruteforce(seed):
    rng = np.random.default_rng(seed)
    varlen = seed % 2 == 1
    codes, off, reads = _random_reads(rng, 60, 25, 8, 20 if varlen else 8)
    for d in (0, 1):
        dt = orc.dir_text(codes, off, d)
        sa = orc.suffix_array(dt, off)
        keys = []
        for r in range(len(reads)):
            a, b = int(off[r]), int(off[r + 1])
            for g in range(a, b):
                keys.append((list(dt[g:b]), g))
        want = [g for _, g in sorted(keys)]
        assert sa.tolist() == want
        assert orc.check_sa(dt, off, sa) == -1
        # a planted mistake: swapping two suffixes must be caught
        bad = sa.copy()
        bad[[3, 4]] = bad[[4, 3]]
        assert orc.check_sa(dt, off, bad) != -1
        b, ri = orc.bwt(dt, off, sa)
        starts = set(off[:-1].tolist())
        for i, g in enumerate(sa.tolist()):
            assert b[i] == (0 if g in starts else dt[g - 1])
        assert sorted(ri.tolist()) == list(range(len(reads)))


def _dfs_graph(reads, t):
    """The reference's depth-first extraction, simulated on explicit read sets."""
    def rc(s):
        return [5 - c for c in reversed(s)]
    out = []
    for q, s in enumerate(reads):
        s = list(s)
        node = {}
        for side, query in (("B", rc(s)), ("F", s)):
            Lq = len(query)
            items = []
            for o in range(t, Lq):
                for m, M in enumerate(reads):
                    M = list(M)
                    if o >= len(M):
                        continue
                    R = rc(M)
                    plain_hit = M[:o] == query[Lq - o:]
                    rc_hit = R[:o] == query[Lq - o:]
                    if side == "F":
                        if plain_hit:
                            items.append((0, o, m, M[o:], 1))
                        if rc_hit:
                            items.append((1, o, m, R[o:], 2))
                    else:
                        if rc_hit:
                            items.append((0, o, m, R[o:], 0))
                        if plain_hit:
                            items.append((1, o, m, M[o:], 3))
            edges = []

            def dfs(S, level):
                if level > 0:
                    done = [it for it in S if len(it[3]) == level]
                    if done:
                        rank, o = min((it[0], it[1]) for it in done)
                        grp = sorted(it[2] for it in done if it[0] == rank and it[1] == o)
                        ty = [it[4] for it in done if it[0] == rank][0]
                        edges.extend((m, o, ty) for m in grp if m != q)
                        return
                for a in (1, 2, 3, 4):
                    sub = [it for it in S if len(it[3]) > level and it[3][level] == a]
                    if sub:
                        dfs(sub, level + 1)
            dfs(items, 0)
            node[side] = (len(items), edges)
        out.append(node)
    return out


@pytest.mark.parametrize("seed", range(8))
def test_search_bruteforce(seed):
    rng = np.random.default_rng(100 + seed)
    varlen = seed % 2 == 1
    t = 6
    codes, off, reads = _random_reads(rng, 90, 40, 10, 24 if varlen else 10,
                                      err=0.02 if seed % 4 >= 2 else 0.0)
    ex = orc.search_explicit(codes, off, t)
    want = _dfs_graph(reads, t)
    for q, node in enumerate(want):
        a, b = int(ex["edge_off"][q]), int(ex["edge_off"][q + 1])
        got = list(zip(ex["dest"][a:b].tolist(), ex["len"][a:b].tolist(),
                       ex["type"][a:b].tolist()))
        assert got == node["B"][1] + node["F"][1], q
        assert ex["overlaps_B"][q] == node["B"][0] and ex["overlaps_F"][q] == node["F"][0]
    if not varlen:
        ix = orc.build_index(codes, off, t)
        fm = orc.search_fm(ix)
        for k in ("edge_off", "dest", "len", "type", "overlaps_B", "overlaps_F"):
            assert np.array_equal(fm[k], ex[k]), k


# ---- fresh runs of the reference (where its programs are built) ---------------------------

@pytest.mark.skipif(not orc.ref_available(), reason="reference programs not built")
@pytest.mark.parametrize("seed", range(3))
def test_against_live_reference(tmp_path, seed):
    import pasqual_amd as pq
    c = pq.SynthCfg()
    c.genome_len, c.len_min, c.len_max, c.coverage = 3000 + 1000 * seed, 40, 40, 15
    c.sub_rate, c.seed = 0.004 * seed, 900 + seed
    fa = tmp_path / "r.fa"
    pq.synth_fasta(c, fa)
    t = 21
    orc.run_ref_dump(fa, 40, t, tmp_path / "d.bin")
    R = orc.read_ref_dump(tmp_path / "d.bin")
    off = _off(R["n"], 40)
    ix = orc.build_index(R["reads"], off, t)
    for d in (0, 1):
        assert np.array_equal(ix[d]["bitmap"], R[d]["bitmap"])
        assert np.array_equal(ix[d]["sample"], R[d]["sample"])
        assert np.array_equal(ix[d]["reads_index"], R[d]["reads_index"])
        assert np.array_equal(ix[d]["shortcut"][2], R[d]["idepth"])
    fm = orc.search_fm(ix)
    for i, nd in enumerate(R["nodes"]):
        a, b = int(fm["edge_off"][i]), int(fm["edge_off"][i + 1])
        E = np.stack([fm["dest"][a:b], fm["len"][a:b], fm["type"][a:b]], 1).astype(np.uint64)
        assert np.array_equal(E, nd["edges"]), i


@pytest.mark.parametrize("k,h,seed", [(21, 64, 1), (5, 7, 2), (1, 1, 3)])
def test_mlp_gradient_oracle_central_differences(k, h, seed):
    # the learned-index MLP oracle: its analytic gradient against central
    # differences of its own loss (ReLU kinks avoided: pre-activations kept away
    # from zero by the draw below)
    rng = np.random.default_rng(seed)
    P = h * k + 2 * h + 1
    params = rng.uniform(-1, 1, P)
    xs = rng.integers(0, 1 << (2 * k), 48, dtype=np.uint64)
    ys = rng.uniform(0, 1, 48)
    loss, g = orc.mlp_loss_grad(k, h, params, xs, ys)
    eps = 1e-6
    num = np.zeros(P)
    for q in range(P):
        a = params.copy()
        a[q] += eps
        b = params.copy()
        b[q] -= eps
        num[q] = (orc.mlp_loss_grad(k, h, a, xs, ys, False)[0] -
                  orc.mlp_loss_grad(k, h, b, xs, ys, False)[0]) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7), np.abs(g - num).max()
    # closed form of the loss for a one-unit net: out = b2 + w2 relu(b1 + W1 f)
    if k == 1 and h == 1:
        f = (xs & 3).astype(np.float64) * 2 / 3 - 1
        out = params[3] + params[2] * np.maximum(0, params[1] + params[0] * f)
        assert np.isclose(loss, np.mean((out - ys) ** 2), rtol=1e-12)


@pytest.mark.parametrize("k,h,seed", [(21, 64, 5), (3, 4, 6)])
def test_mlp_forward_oracle_matches_its_loss(k, h, seed):
    # the forward oracle against the loss oracle (pinned by central differences
    # and a closed form above): for one sample, (out - 0)^2 - (out - 1)^2 = 2 out - 1
    rng = np.random.default_rng(seed)
    P = h * k + 2 * h + 1
    params = rng.uniform(-1, 1, P)
    xs = rng.integers(0, 1 << (2 * k), 32, dtype=np.uint64)
    out = orc.mlp_forward(k, h, params, xs)
    for i in range(len(xs)):
        l0 = orc.mlp_loss_grad(k, h, params, xs[i:i + 1], np.zeros(1), False)[0]
        l1 = orc.mlp_loss_grad(k, h, params, xs[i:i + 1], np.ones(1), False)[0]
        assert np.isclose(out[i], (l0 - l1 + 1) / 2, rtol=1e-12, atol=1e-12)
    # the output bias shifts every output by itself
    p2 = params.copy()
    p2[-1] += 0.25
    assert np.allclose(orc.mlp_forward(k, h, p2, xs) - out, 0.25)


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", CASES)
def test_slot_sizes_match_reference(name, width):
    # index_t's first-triplet slot sizes as the reference's construct_index
    # reports them (tests/golden/slots.json, written by make_slot_golden.py)
    import json
    want = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "slots.json")))
    want = want[name][width]
    G = load_golden(name, width)
    L = golden_cases()[name]["L"]
    n = int(G["n"])
    off = _off(n, L)
    for d, tag in ((0, "fw"), (1, "bw")):
        size, offs, mx = orc.slot_sizes(orc.dir_text(G["reads"], off, d), off)
        assert size == want[f"size_{tag}"] and offs == want[f"offset_{tag}"], (d, size)
        assert mx == want[f"max_size_{tag}"] and sum(size) == n * L


def _seed_scan_walk(sufs, pat, st, vmax):
    # rows the learned seed kernel's scan reads, walked over the sorted suffixes
    # themselves: samples are rows 0, st, 2 st, ...; the first sample whose
    # D-code prefix is >= the pattern, and whether it holds the pattern; the scan
    # goes down (and up, from a sample holding the pattern) until a row leaves
    # the pattern, or the previous sample's row is reached
    D, N = len(pat), len(sufs)
    samples = list(range(0, N, st))
    jl = next((j for j, r in enumerate(samples) if sufs[r][:D] >= pat), len(samples))
    m1 = jl < len(samples) and sufs[samples[jl]][:D] == pat
    if m1:
        jh = next((j for j in range(jl, len(samples)) if sufs[samples[j]][:D] > pat), len(samples))
        if (jh - jl - 1) * st + 1 > vmax:
            return None
    floor = samples[jl - 1] + 1 if jl > 0 else 0
    top = samples[jl] if jl < len(samples) else N
    rows = 0
    if m1:
        rows, r = 1, top - 1
        while r >= floor:
            rows += 1
            if sufs[r][:D] != pat:
                break
            r -= 1
        r = top + 1
        while r < N:
            rows += 1
            if sufs[r][:D] != pat:
                break
            r += 1
    else:
        r = top - 1
        while r >= floor:
            rows += 1
            if sufs[r][:D] < pat:
                break
            r -= 1
    return rows


def _chain_work_bruteforce(codes, off, seed_depth=0, t=0, vmax=0, stride=0):
    # the chains' work from explicit suffix strings: the interval of a pattern is
    # [number of suffixes sorting before it, + number having it as a prefix) in the
    # order end of read first, codes 1<2<3<4 (the suffix-array order)
    n = len(off) - 1
    texts = [orc.dir_text(codes, off, d) for d in (0, 1)]
    sufs = []
    for d in (0, 1):
        s = []
        for r in range(n):
            a, b = int(off[r]), int(off[r + 1])
            for g in range(a, b):
                s.append(tuple(int(x) for x in texts[d][g:b]))
        sufs.append(sorted(s))

    def interval(d, pat):
        lo = bisect.bisect_left(sufs[d], pat)
        k = sum(1 for x in sufs[d] if x[:len(pat)] == pat)
        return lo, k

    pos = lines = rows = 0
    for r in range(n):
        s = [int(x) for x in codes[int(off[r]):int(off[r + 1])]]
        L = len(s)
        for ch in range(4):
            d = 0 if ch in (0, 3) else 1
            minsize = 2 if ch in (0, 2) else 1

            def sym(o):
                x = s[L - o] if ch < 2 else s[o - 1]
                return 5 - x if ch in (1, 3) else x
            seeded = 1 <= seed_depth < L
            o = seed_depth if seeded else 1
            while o < L:
                pat = tuple(sym(j) for j in range(o, 0, -1))   # prepended codes first
                lo, k = interval(d, pat)
                if seeded and o == seed_depth and stride and vmax and o <= t:
                    scanned = _seed_scan_walk(sufs[d], pat, stride, vmax)
                    if scanned is not None:
                        rows += scanned
                        if minsize <= k <= vmax:
                            break
                if k < minsize:
                    break
                if k <= vmax and o <= t:
                    rows += k
                    break
                pos += 2
                lines += 1 if lo >> 7 == (lo + k) >> 7 else 2
                o += 1
    return pos, lines, rows


@pytest.mark.parametrize("seed,varlen,depth,t,vmax,stride",
                         [(0, False, 0, 0, 0, 0), (1, True, 0, 0, 0, 0), (2, False, 5, 0, 0, 0),
                          (3, True, 7, 0, 0, 0), (4, False, 0, 0, 0, 0), (5, False, 0, 6, 4, 0),
                          (6, True, 0, 9, 8, 0), (7, True, 5, 8, 32, 0), (8, True, 5, 8, 32, 8),
                          (9, False, 4, 6, 5, 3), (10, True, 6, 9, 7, 1), (11, True, 6, 6, 2, 2)])
def test_chain_work_matches_bruteforce(seed, varlen, depth, t, vmax, stride):
    # the oracle's count of the device chains' rank positions, occurrence lines
    # and verified rows (FM walk with per-block ranks) equals the count from
    # explicit sorted suffixes; the repeat-rich genome makes intervals span
    # 128-row lines; with a sampling stride, the seeded chains' scan is walked
    # row by row over the sorted suffixes
    rng = np.random.default_rng(100 + seed)
    codes, off, _ = _random_reads(rng, 90, 60, 10, 30 if varlen else 10, err=0.02)
    want = _chain_work_bruteforce(codes, off, depth, t, vmax, stride)
    assert orc.chain_work(codes, off, depth, t, vmax, stride) == want
    if vmax:
        assert want[2] > 0
    else:
        assert want[1] > want[0] // 2      # some steps span two lines


def test_chain_work_plant_is_caught():
    # a read whose last code changes moves the count: the count follows the reads
    rng = np.random.default_rng(7)
    codes, off, _ = _random_reads(rng, 90, 60, 12, 12, err=0.0)
    a = orc.chain_work(codes, off)
    c2 = codes.copy()
    c2[11] = 5 - c2[11]
    assert orc.chain_work(c2, off) != a
    # verification takes over from the backward steps
    v = orc.chain_work(codes, off, 0, 11, 8)
    assert v[2] > 0 and v[0] < a[0]
