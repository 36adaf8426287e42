"""Write the duplicate-removal fixtures of tests/golden/ from the reference itself.

Two seeded reads files full of duplicates (exact copies, reverse-complement
copies, lower-case copies, reads with N), imported by the reference's own
construct_index (oracle/_ref/ref_dump_*, the reference's index.c and bloom.c)
for both index widths and several Bloom-filter settings, down to one bit per
record, where false positives chain through the file.  Stored per setting: the
mask of records the reference keeps (packed bits), recovered from its kept
reads in input order (among identical records only the first can be kept, so
the first match is the kept one).  Run from the repository root after
`make -C oracle`:

    python tests/golden/make_import_golden.py
"""
import gzip
import json
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pq_oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# name: read length, records, seed
CASES = {"dupA": dict(L=24, records=4000, seed=101), "dupB": dict(L=150, records=1500, seed=202)}
# (bloom_m, n_hashfunc)
SETTINGS = [(36, 2), (1, 1), (2, 5), (4, 3), (1, 4), (8, 5)]


def records_of(L, count, seed):
    """Sequences (strings) of the reads file: new random reads, copies of an
    earlier read as it is, reverse-complemented or in lower case, and reads
    with an N."""
    rng = np.random.default_rng(seed)
    comp = str.maketrans("ACGTacgt", "TGCAtgca")
    out = []
    for i in range(count):
        u = rng.random()
        if i < 10 or u < 0.45:
            s = "".join(rng.choice(list("ACGT"), L))
        else:
            src = out[int(rng.integers(0, len(out)))]
            if u < 0.65:
                s = src
            elif u < 0.82:
                s = src.translate(comp)[::-1]
            elif u < 0.90:
                s = src.lower()
            else:
                j = int(rng.integers(0, L))
                s = src[:j] + ("N" if rng.random() < 0.5 else "n") + src[j + 1:]
        out.append(s)
    return out


def code_of(s):
    return np.frombuffer(s.upper().translate(str.maketrans("ACGT", "\x01\x02\x03\x04")).encode(),
                         np.uint8)


def keep_mask(recs, reads, L):
    kept = reads.reshape(-1, L)
    mask = np.zeros(len(recs), bool)
    k = 0
    for i, s in enumerate(recs):
        if k < len(kept) and "N" not in s.upper() and np.array_equal(code_of(s), kept[k]):
            mask[i] = True
            k += 1
    assert k == len(kept), "kept reads are not a subsequence of the records"
    return mask


def main():
    tmp = tempfile.mkdtemp()
    meta = {}
    for name, p in CASES.items():
        L = p["L"]
        recs = records_of(L, p["records"], p["seed"])
        fa = os.path.join(tmp, name + ".fa")
        with open(fa, "w") as f:
            for i, s in enumerate(recs):
                f.write(f">r{i}\n{s}\n")
        with open(fa, "rb") as fi, gzip.GzipFile(os.path.join(HERE, name + ".fa.gz"), "wb",
                                                  mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        out = {}
        for width in ("u32", "u64"):
            for m, nf in SETTINGS:
                dump = os.path.join(tmp, f"{name}_{width}_{m}_{nf}.bin")
                orc.run_ref_dump(fa, L, L // 2, dump, width=width, graph=False, bloom_m=m,
                                 n_hashfunc=nf)
                R = orc.read_ref_dump(dump)
                mask = keep_mask(recs, R["reads"], L)
                out[f"{width}_m{m}_nf{nf}"] = np.packbits(mask)
                print(name, width, m, nf, "kept", int(mask.sum()), "of", len(recs))
        np.savez_compressed(os.path.join(HERE, f"import_{name}.npz"), **out)
        meta[name] = {"L": L, "records": p["records"],
                      "settings": [list(x) for x in SETTINGS]}
    with open(os.path.join(HERE, "import_cases.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
