"""Host-side checks that need no GPU: the library loads and exports its C ABI,
reads import and duplicate removal match the reference, the workload
generator is deterministic."""
import os
import re

import numpy as np
import pytest

import pasqual_amd as pq
from conftest import ROOT, golden_cases, golden_fasta, load_golden

CASES = sorted(golden_cases())


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(pq_\w+|construct_\w+|destroy_graph|free_index)\s*\(", txt))
    if "time_importreads" in txt:
        names.add("time_importreads")
    return names


@pytest.mark.parametrize("width", ["u32", "u64"])
def test_library_exports_every_declared_symbol(width):
    L = pq.lib(width)
    declared = _declared("pasqual_amd.h") | _declared("pasqual_refabi.h")
    assert declared >= set(pq.REF_SYMBOLS)
    assert set(pq.PQ_SYMBOLS) <= declared
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.parametrize("width", ["u32", "u64"])
def test_library_keeps_reference_host_symbols_private(width):
    # the reference's bloom.c / common.c are linked beside the library in the
    # drop-in build: the library must not export their names
    L = pq.lib(width)
    for s in ("bloom_filter", "hash0", "nt_table", "ntc_table", "malloc_align"):
        assert not hasattr(L, s), s


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", CASES)
def test_import_matches_reference(tmp_path, name, width):
    G = load_golden(name, width)
    meta = golden_cases()[name]
    fa = golden_fasta(name, tmp_path)
    imp = pq.import_fasta(fa, meta["L"], index_bits=32 if width == "u32" else 64)
    assert imp["n"] == int(G["n"])
    assert np.array_equal(imp["reads"], G["reads"])
    assert imp["C"] == G["C"].tolist()
    assert imp["suffix_count"] == G["suffix_count"].tolist()
    assert imp["prefix_count"] == G["prefix_count"].tolist()


def test_import_errors(tmp_path):
    good = tmp_path / "g.fa"
    good.write_text(">a\nACGTACGTAC\n>b\nTTGCAACGTA\n")
    assert pq.import_fasta(good, 10)["n"] == 2
    bad_len = tmp_path / "l.fa"
    bad_len.write_text(">a\nACGTACGTAC\n>b\nTTGCA\n")
    with pytest.raises(pq.PqError, match="PQ_ERR_FORMAT line 4"):
        pq.import_fasta(bad_len, 10)
    bad_char = tmp_path / "c.fa"
    bad_char.write_text(">a\nACGTACGXAC\n")
    with pytest.raises(pq.PqError, match="PQ_ERR_FORMAT line 2"):
        pq.import_fasta(bad_char, 10)
    no_header = tmp_path / "h.fa"
    no_header.write_text("ACGTACGTAC\n")
    with pytest.raises(pq.PqError, match="PQ_ERR_FORMAT"):
        pq.import_fasta(no_header, 10)
    with pytest.raises(pq.PqError, match="PQ_ERR_IO"):
        pq.import_fasta(tmp_path / "missing.fa", 10)
    with_n = tmp_path / "n.fa"
    with_n.write_text(">a\nACGTNCGTAC\n>b\nACGTACGTAC\n>c\nGTACGTACGT\n")
    r = pq.import_fasta(with_n, 10)
    # the N read is dropped, and the reverse complement of read b is a duplicate
    assert r["n"] == 1 and r["removed"] == 2


def test_synth_presets_and_determinism():
    c1 = pq.synth_preset(1)
    assert pq.lib().pq_synth_num_reads(__import__("ctypes").byref(c1)) == 400000
    c2 = pq.synth_preset(2)
    assert pq.lib().pq_synth_num_reads(__import__("ctypes").byref(c2)) == 2000000
    c3 = pq.synth_preset(3)
    assert pq.lib().pq_synth_num_reads(__import__("ctypes").byref(c3)) == 10489511
    small = pq.SynthCfg()
    small.genome_len, small.len_min, small.len_max, small.coverage, small.seed = 5000, 36, 90, 5, 3
    a, oa = pq.synth_reads(small)
    b, ob = pq.synth_reads(small)
    assert np.array_equal(a, b) and np.array_equal(oa, ob)
    lens = np.diff(oa.astype(np.int64))
    assert lens.min() >= 36 and lens.max() <= 90 and set(np.unique(a)) <= {1, 2, 3, 4}
    small.seed = 4
    c, _ = pq.synth_reads(small)
    assert not np.array_equal(a[:1000], c[:1000])
