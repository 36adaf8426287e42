"""The library as a drop-in for index.o + sufsort.o + string_graph.o: the
reference's unchanged CLI, bloom filter and assembly linked against it produce
the reference's contigs.fa; the reference's own string_graph.c run on the
library's construct_index produces the reference's graph."""
import argparse
import os
import subprocess

import numpy as np
import pytest

import pq_oracle as orc
from conftest import golden_cases, golden_fasta, load_golden

pytestmark = pytest.mark.gpu
CASES = sorted(golden_cases())


def _exe(name):
    p = os.path.join(orc.REFDIR, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not built (make -C oracle dropin)")
    return p


@pytest.mark.parametrize("width", ["u32", "u64"])
@pytest.mark.parametrize("name", CASES)
def test_contigs_match_reference(tmp_path, name, width):
    exe = _exe(f"pasqual_hip_{width}")
    G = load_golden(name, width)
    meta = golden_cases()[name]
    fa = golden_fasta(name, tmp_path)
    r = subprocess.run([exe, "--singlereads", fa, "-l", str(meta["L"]), "-t", str(meta["t"])],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert r.returncode == 0, r.stderr
    assert f"has imported {int(G['n'])} reads" in r.stdout
    contigs = (tmp_path / "contigs.fa").read_bytes()
    assert contigs == G["contigs"].tobytes()


@pytest.mark.parametrize("name", CASES)
def test_reference_search_runs_on_device_index(tmp_path, name):
    exe = _exe("refsg_hip_u32")
    G = load_golden(name, "u32")
    meta = golden_cases()[name]
    fa = golden_fasta(name, tmp_path)
    out = tmp_path / "d.bin"
    r = subprocess.run([exe, fa, str(meta["L"]), str(meta["t"]), str(out)], capture_output=True,
                       text=True, timeout=120, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert r.returncode == 0, r.stderr
    R = orc.read_ref_dump(out)
    for d in (0, 1):
        assert np.array_equal(R[d]["bitmap"], G[f"bitmap{d}"])
        assert np.array_equal(R[d]["sample"], G[f"sample{d}"])
        assert np.array_equal(R[d]["reads_index"], G[f"reads_index{d}"])
    edges = [nd["edges"] for nd in R["nodes"]]
    E = np.concatenate(edges) if sum(len(e) for e in edges) else np.zeros((0, 3), np.uint64)
    assert np.array_equal(E, G["edges"])
    assert R["average_overlap"] == float(G["average_overlap"])


def test_config5_generator_matches_reference(tmp_path):
    # the large workload's generator (repeat families of 1-10 kbp at 0.5-3%
    # divergence, 150 bp reads, t=63) over a 300 kbp genome, 64-bit index:
    # the reference CLI and the same CLI on this library import the same reads
    # and write the same contigs.fa (one thread each: the reference's assembly
    # has unsynchronised updates); the reference's own index and overlap graph
    # (ref_graph_u64) and this library's, from the same file, are byte-identical
    import full_parity
    ref_exe, hip_exe = _exe("pasqual_omp_u64"), _exe("pasqual_hip_u64")
    _exe("ref_graph_u64")
    import pasqual_amd as pq
    cfg = pq.synth_preset(5)
    cfg.genome_len = 300_000
    fa = str(tmp_path / "c5.fa")
    pq.synth_fasta(cfg, fa)
    outs = []
    for exe in (ref_exe, hip_exe):
        wd = tmp_path / os.path.basename(exe)
        wd.mkdir()
        r = subprocess.run([exe, "--singlereads", fa, "-l", "150", "-t", "63"], cwd=wd,
                           capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        assert r.returncode == 0, r.stderr
        n = int(r.stdout.split("has imported ")[1].split()[0])
        outs.append((n, (wd / "contigs.fa").read_bytes()))
    assert outs[0][0] == outs[1][0] and outs[0][0] > 50000
    assert outs[0][1] == outs[1][1] and len(outs[0][1]) > 1000
    a = argparse.Namespace(config=5, genome_len=300_000, tmp=str(tmp_path))
    res = full_parity.compare(full_parity.side_ref(a), full_parity.side_lib(a))
    assert res["equal"]["all_equal"], res["equal"]
