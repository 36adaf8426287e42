"""The oracle's reads import pinned: against the reference's own import of the
golden reads files (tests/golden/*.npz), of the duplicate-heavy files under six
Bloom-filter settings and both index widths (import_*.npz), and, where the
reference programs are built, against fresh reference runs on malformed and
boundary files (line pieces of the 1000-byte fgets buffer, NUL bytes, CRLF)."""
import os
import re
import subprocess

import numpy as np
import pytest

import pq_oracle as orc
from conftest import (golden_cases, golden_fasta, import_cases, import_expected,
                      import_records, load_golden)


def edge_files():
    """(name, content, read length, expected): expected is ("ok", kept codes as
    strings) or ("bad", line the reference reports)."""
    long_tail = ">" + "x" * 998 + "ACGTACGTAC\n"     # 1010 bytes: pieces of 999 and 11
    return [
        ("bad_len", ">a\nACGTACGTAC\n>b\nTTGCA\n", 10, ("bad", 4)),
        ("bad_char", ">a\nACGTACGXAC\n", 10, ("bad", 2)),
        ("no_header", "ACGTACGTAC\n", 10, ("bad", 2)),
        ("with_n", ">a\nACGTNCGTAC\n>b\nACGTACGTAC\n>c\nGTACGTACGT\n", 10,
         ("ok", ["ACGTACGTAC"])),
        ("no_final_newline", ">a\nACGTACGTAC\n>b\nACGTACGTAA", 10, ("bad", 4)),
        ("header_without_sequence", ">a\nACGTACGTAC\n>b\n", 10, ("bad", 4)),
        ("empty", "", 10, ("ok", [])),
        ("blank_line", ">a\nACGTACGTAC\n\n>b\nACGTACGTAA\n", 10, ("bad", 4)),
        ("long_header", ">" + "x" * 1498 + "\n>b\nACGTACGTAC\n", 10, ("bad", 2)),
        ("long_header_tail_is_a_read", long_tail + ">b\nTTTTTTTTTA\n", 10,
         ("ok", ["ACGTACGTAC", "TTTTTTTTTA"])),
        ("read_of_998", ">a\n" + "ACGT" * 249 + "AC\n", 998, ("ok", ["ACGT" * 249 + "AC"])),
        ("read_of_999", ">a\n" + "ACGT" * 249 + "ACG\n", 999, ("bad", 2)),
        ("nul_in_sequence", ">a\nACGTACGTACX\0junk\n>b\nGGGGGGGGGA\n", 10,
         ("ok", ["ACGTACGTAC", "GGGGGGGGGA"])),
        ("crlf", ">a\r\nACGTACGTAC\r\n", 10, ("bad", 2)),
        ("lower_and_rc", ">a\nacgtacgtac\n>b\nGTACGTACGT\n>c\nAAAACCCCGG\n", 10,
         ("ok", ["ACGTACGTAC", "AAAACCCCGG"])),
    ]


def codes(strings):
    lut = {"A": 1, "C": 2, "G": 3, "T": 4}
    return np.array([lut[c] for s in strings for c in s], np.uint8)


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_oracle_import_matches_reference(tmp_path, name, width):
    G = load_golden(name, width)
    fa = golden_fasta(name, tmp_path)
    r = orc.import_fasta(fa, golden_cases()[name]["L"], index_bits=32 if width == "u32" else 64)
    assert r["rc"] == 0
    assert r["n"] == int(G["n"])
    assert np.array_equal(r["reads"], G["reads"])
    assert r["C"] == G["C"].tolist()
    assert r["suffix_count"] == G["suffix_count"].tolist()
    assert r["prefix_count"] == G["prefix_count"].tolist()


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", sorted(import_cases()))
def test_oracle_import_matches_reference_filter_settings(tmp_path, name, width):
    fa, recs = import_records(name, tmp_path)
    meta = import_cases()[name]
    L = meta["L"]
    for m, nf in meta["settings"]:
        mask, want = import_expected(name, recs, width, m, nf, L)
        r = orc.import_fasta(fa, L, m, nf, 32 if width == "u32" else 64)
        assert r["rc"] == 0
        assert r["n"] == int(mask.sum()), (m, nf)
        assert np.array_equal(r["reads"], want), (m, nf)
        assert r["removed"] == len(recs) - r["n"]
        # the filter alone, over the parsed records, gives the same verdicts
        ok = np.array(["N" not in s.upper() for s in recs], np.uint8)
        lut = np.zeros(256, np.uint8)
        lut[np.frombuffer(b"ACGTacgt", np.uint8)] = [1, 2, 3, 4, 1, 2, 3, 4]
        cod = lut[np.frombuffer("".join(recs).encode(), np.uint8)]
        keep = orc.dedup(cod, ok, L, len(recs) * m, nf, 32 if width == "u32" else 64)
        assert np.array_equal(keep.astype(bool), mask), (m, nf)


@pytest.mark.parametrize("case", edge_files(), ids=lambda c: c[0])
def test_oracle_import_edge_files(tmp_path, case):
    name, content, L, (kind, want) = case
    fa = tmp_path / f"{name}.fa"
    fa.write_bytes(content.encode())
    r = orc.import_fasta(fa, L)
    if kind == "bad":
        assert r["rc"] == -3 and r["bad_line"] == want
    else:
        assert r["rc"] == 0 and np.array_equal(r["reads"], codes(want))
    assert orc.import_fasta(tmp_path / "missing.fa", L)["rc"] == -2


@pytest.mark.skipif(not orc.ref_available("u32"), reason="reference programs not built")
@pytest.mark.parametrize("case", [c for c in edge_files() if c[3] != ("ok", [])],
                         ids=lambda c: c[0])
def test_edge_files_against_live_reference(tmp_path, case):
    # the reference's own construct_index (oracle/_ref/ref_dump_u32) on the same file
    name, content, L, (kind, want) = case
    fa = tmp_path / f"{name}.fa"
    fa.write_bytes(content.encode())
    out = tmp_path / "dump.bin"
    exe = os.path.join(orc.REFDIR, "ref_dump_u32")
    p = subprocess.run([exe, str(fa), str(L), str(max(1, L // 2)), str(out), "nograph"],
                       capture_output=True, text=True, timeout=60)
    m = re.search(r"incorrect reads file at line (\d+)", p.stderr)
    if kind == "bad":
        assert m and int(m.group(1)) == want, p.stderr
    else:
        # the import report (index.c:375-386); the dump only when the reference's
        # suffix sort survives the tiny read set (it faults on a single read)
        rep = re.search(r"has imported (\d+) reads \(removed (\d+) duplicate", p.stdout)
        assert not m and rep, p.stdout + p.stderr
        assert int(rep.group(1)) == len(want)
        assert int(rep.group(2)) == content.count(">") - len(want)
        if out.exists():
            assert np.array_equal(orc.read_ref_dump(out)["reads"], codes(want))
