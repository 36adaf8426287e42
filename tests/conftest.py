import gzip
import json
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def golden_cases():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        return json.load(f)


def load_golden(name, width="u32"):
    """Reference outputs of one fixture (written by tests/golden/make_golden.py)."""
    z = np.load(os.path.join(GOLDEN, f"{name}_{width}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def golden_fasta(name, tmpdir):
    """Decompress the fixture's reads file into tmpdir and return its path."""
    src = os.path.join(GOLDEN, f"{name}.fa.gz")
    dst = os.path.join(str(tmpdir), f"{name}.fa")
    with gzip.open(src, "rb") as fi, open(dst, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return dst


def import_cases():
    """Duplicate-removal fixtures (written by tests/golden/make_import_golden.py)."""
    with open(os.path.join(GOLDEN, "import_cases.json")) as f:
        return json.load(f)


def import_records(name, tmpdir):
    """Reads file of a duplicate-removal fixture: (path, sequences in file order)."""
    path = golden_fasta(name, tmpdir)
    with open(path) as f:
        lines = f.read().split("\n")
    return path, lines[1::2][:import_cases()[name]["records"]]


def import_expected(name, records, width, bloom_m, n_hashfunc, L):
    """Codes of the records the reference keeps under one filter setting."""
    z = np.load(os.path.join(GOLDEN, f"import_{name}.npz"), allow_pickle=False)
    mask = np.unpackbits(z[f"{width}_m{bloom_m}_nf{n_hashfunc}"])[:len(records)].astype(bool)
    lut = np.zeros(256, np.uint8)
    for i, c in enumerate("ACGT"):
        lut[ord(c)] = lut[ord(c.lower())] = i + 1
    kept = [lut[np.frombuffer(s.encode(), np.uint8)] for s, k in zip(records, mask) if k]
    return mask, (np.concatenate(kept) if kept else np.zeros(0, np.uint8))


def golden_edges(G, i):
    a, b = int(G["edge_off"][i]), int(G["edge_off"][i + 1])
    return G["edges"][a:b]


@pytest.fixture(scope="session")
def ctx():
    import pasqual_amd as pq
    c = pq.Context()
    yield c
    c.close()
