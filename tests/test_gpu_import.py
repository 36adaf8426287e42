"""Reads import with duplicate removal, the filter on the device: the kept reads,
counts and error lines equal the reference's (golden fixtures written by the
reference's own import, under six Bloom-filter settings and both index widths)
and the oracle's (malformed and boundary files, a large duplicate-heavy file,
tiny filters whose false positives chain through the whole record list)."""
import numpy as np
import pytest

import pasqual_amd as pq
import pq_oracle as orc
from conftest import (golden_cases, golden_fasta, import_cases, import_expected,
                      import_records, load_golden)
from test_oracle_import import codes, edge_files

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_import_matches_reference(ctx, tmp_path, name, width):
    G = load_golden(name, width)
    fa = golden_fasta(name, tmp_path)
    imp = pq.import_fasta(fa, golden_cases()[name]["L"], index_bits=32 if width == "u32" else 64,
                          ctx=ctx)
    assert imp["n"] == int(G["n"])
    assert np.array_equal(imp["reads"], G["reads"])
    assert imp["C"] == G["C"].tolist()
    assert imp["suffix_count"] == G["suffix_count"].tolist()
    assert imp["prefix_count"] == G["prefix_count"].tolist()


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", sorted(import_cases()))
def test_import_filter_settings_match_reference(ctx, tmp_path, name, width):
    fa, recs = import_records(name, tmp_path)
    meta = import_cases()[name]
    L = meta["L"]
    for m, nf in meta["settings"]:
        mask, want = import_expected(name, recs, width, m, nf, L)
        imp = pq.import_fasta(fa, L, m, nf, 32 if width == "u32" else 64, ctx=ctx)
        assert imp["n"] == int(mask.sum()), (m, nf)
        assert np.array_equal(imp["reads"], want), (m, nf)
        assert imp["removed"] == len(recs) - imp["n"], (m, nf)


@pytest.mark.parametrize("case", edge_files(), ids=lambda c: c[0])
def test_import_edge_files(ctx, tmp_path, case):
    name, content, L, (kind, want) = case
    fa = tmp_path / f"{name}.fa"
    fa.write_bytes(content.encode())
    if kind == "bad":
        with pytest.raises(pq.PqError, match=f"PQ_ERR_FORMAT line {want}"):
            pq.import_fasta(fa, L, ctx=ctx)
    else:
        imp = pq.import_fasta(fa, L, ctx=ctx)
        assert np.array_equal(imp["reads"], codes(want))
        assert imp["removed"] == content.count(">") - len(want)
    with pytest.raises(pq.PqError, match="PQ_ERR_IO"):
        pq.import_fasta(tmp_path / "missing.fa", L, ctx=ctx)


def _records(rng, nrec, L, pool, n_frac=0.03):
    """Records drawn from a small pool of reads, half of them reverse-complemented,
    some with N: many exact and reverse-complement duplicates."""
    base = rng.integers(1, 5, (pool, L), dtype=np.uint8)
    pick = base[rng.integers(0, pool, nrec)]
    flip = rng.random(nrec) < 0.5
    pick[flip] = (5 - pick[flip])[:, ::-1]
    ok = (rng.random(nrec) >= n_frac).astype(np.uint8)
    return np.ascontiguousarray(pick).reshape(-1), ok


@pytest.mark.parametrize("width", [32, 64])
def test_dedup_matches_oracle_under_collisions(ctx, width):
    rng = np.random.default_rng(width)
    passes = []
    for nrec, L, pool, bits, nf in [(5000, 30, 1500, 64, 1), (5000, 30, 1500, 7, 2),
                                    (20000, 50, 8000, 20000, 3), (20000, 17, 20000, 5000, 5),
                                    (3000, 100, 100, 3000 * 36, 2), (1, 10, 1, 1, 4),
                                    (4000, 24, 4000, 4001, 4)]:
        cod, ok = _records(rng, nrec, L, pool)
        want = orc.dedup(cod, ok, L, bits, nf, width)
        got, p = ctx.dedup_reads(cod, ok, L, bits, nf, width)
        assert np.array_equal(got, want), (nrec, L, bits, nf)
        passes.append(p)
    # tiny filters make verdicts depend on long chains of earlier verdicts
    assert max(passes) >= 3, passes


def test_dedup_zero_bit_filter(ctx):
    cod = np.ones(40, np.uint8)
    keep, _ = ctx.dedup_reads(cod, np.zeros(4, np.uint8), 10, 0)
    assert not keep.any()                      # every record holds N: the filter is never probed
    with pytest.raises(pq.PqError, match="PQ_ERR_LIMIT"):
        ctx.dedup_reads(cod, np.array([0, 1, 0, 0], np.uint8), 10, 0)


@pytest.mark.parametrize("width", ["u32", "u64"])
def test_large_duplicate_heavy_file_matches_oracle(ctx, tmp_path, width):
    # 200k records over 60k distinct reads (both strands, lower case, N): the
    # file is cut by every host thread and the filter settles over many passes
    rng = np.random.default_rng(7 if width == "u32" else 8)
    L, nrec = 100, 200000
    cod, ok = _records(rng, nrec, L, 60000, n_frac=0.01)
    letters = np.frombuffer(b"NACGT", np.uint8)[cod.reshape(nrec, L)]
    letters[ok == 0, rng.integers(0, L)] = ord("N")
    low = rng.random(nrec) < 0.1
    letters[low] = letters[low] | 0x20
    fa = tmp_path / "big.fa"
    with open(fa, "wb") as f:
        for i in range(nrec):
            f.write(b">r%d\n" % i + letters[i].tobytes() + b"\n")
    bits = 32 if width == "u32" else 64
    for m, nf in [(36, 2), (3, 3)]:
        want = orc.import_fasta(fa, L, m, nf, bits)
        imp = pq.import_fasta(fa, L, m, nf, bits, ctx=ctx)
        assert imp["n"] == want["n"] and imp["removed"] == want["removed"], (m, nf)
        assert np.array_equal(imp["reads"], want["reads"]), (m, nf)
        assert imp["C"] == want["C"] and imp["prefix_count"] == want["prefix_count"]
        assert imp["suffix_count"] == want["suffix_count"]
