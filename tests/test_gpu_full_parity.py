"""Full-size overlap-graph parity of configs 1 and 2 against the reference's own
graph: the committed records (profiles/config<N>_parity.json, written by
oracle/full_parity.py from the reference's construct_index +
construct_string_graph at the workload's full size) hold the reference's read
count, edge count, overlap-length sum and the digests of its node records and
edge_t records; the device library, run here on the same generated reads file,
must give the same read set and the same graph in both search modes, exported
after the search and during it."""
import argparse
import json
import os

import pytest

import full_parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("config", [1, 2])
def test_full_size_graph_equals_reference_record(config, tmp_path):
    with open(os.path.join(ROOT, "profiles", f"config{config}_parity.json")) as f:
        ref = json.load(f)["reference"]
    lib = full_parity.side_lib(argparse.Namespace(config=config, genome_len=0, tmp=str(tmp_path)))
    assert lib["read_sets_equal"]
    assert lib["reads"] == ref["reads"]
    for name in ("plain", "learned"):
        g = lib[name]
        for k in ("edges", "overlap_len_sum", "nodes_digest", "edges_digest"):
            assert g[k] == ref[k], f"config {config} {name} {k}"
        assert g["export_during_search_equal"], f"config {config} {name}"
