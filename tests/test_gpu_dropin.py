"""The library as a drop-in for index.o + sufsort.o + string_graph.o: the
reference's unchanged CLI, bloom filter and assembly linked against it produce
the reference's contigs.fa; the reference's own string_graph.c run on the
library's construct_index produces the reference's graph."""
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
