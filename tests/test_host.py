"""Host-side checks that need no GPU: the library loads and exports its C ABI,
keeps the reference's host names private, and the workload generator is
deterministic.  (The reads import runs its duplicate filter on the device:
tests/test_gpu_import.py.)"""
import os
import re

import numpy as np
import pytest

import pasqual_amd as pq
from conftest import ROOT


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


def test_build_byte_model_matches_its_worked_values():
    # the prefix-doubling byte model the bench reports the suffix-array build
    # against: 78.9 B per suffix for 50 bp reads and a 32-bit index (2 rounds),
    # 235.3 B for 150 bp reads and a 64-bit index (4 rounds), exactly 78.9375 and
    # 235.25; a dropped term, a wrong width or an off-by-one round count moves
    # these
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "bench", os.path.join(os.path.dirname(__file__), "..", "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    r, b = bench.build_model_bytes(50, 4)
    assert r == 2 and b == 78.9375
    r, b = bench.build_model_bytes(150, 8)
    assert r == 4 and b == 235.25
    # 29 codes fit the initial key: one round
    assert bench.build_model_bytes(29, 4)[0] == 1 and bench.build_model_bytes(30, 4)[0] == 2
