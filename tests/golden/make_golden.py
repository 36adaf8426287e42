"""Write the golden fixtures of tests/golden/ from the reference itself.

For every case: a seeded reads file (gzip'ed FASTA), and for each index width
the reference's index_t arrays and string-graph nodes (oracle/_ref/ref_dump_*,
one OpenMP thread so the BWT tie order is the canonical one), plus the
reference's contigs.fa at one thread.  Run from the repository root after
`make -C oracle`:

    python tests/golden/make_golden.py
"""
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pasqual_amd as pq  # noqa: E402
import pq_oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# name: genome, read length, coverage, seed, t, substitutions, repeat units
CASES = {
    "tiny": dict(G=400, L=20, cov=18, seed=7, t=11),
    "c1s": dict(G=8000, L=50, cov=20, seed=11, t=31),
    "c2s": dict(G=12000, L=100, cov=25, seed=13, t=63, sub=0.005, units=4, copies=8,
                ulen=500, div=0.01),
}


def cfg_of(p):
    c = pq.SynthCfg()
    c.genome_len = p["G"]
    c.len_min = c.len_max = p["L"]
    c.coverage = p["cov"]
    c.seed = p["seed"]
    c.min_overlap = p["t"]
    c.sub_rate = p.get("sub", 0.0)
    c.n_units = p.get("units", 0)
    c.n_copies = p.get("copies", 0)
    c.unit_len = p.get("ulen", 0)
    c.copy_div = p.get("div", 0.0)
    return c


def main():
    tmp = tempfile.mkdtemp()
    meta = {}
    for name, p in CASES.items():
        fa = os.path.join(tmp, name + ".fa")
        pq.synth_fasta(cfg_of(p), fa)
        with open(fa, "rb") as fi, gzip.GzipFile(os.path.join(HERE, name + ".fa.gz"), "wb",
                                                  mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        meta[name] = {"L": p["L"], "t": p["t"]}
        for width in ("u32", "u64"):
            dump = os.path.join(tmp, f"{name}_{width}.bin")
            orc.run_ref_dump(fa, p["L"], p["t"], dump, width=width)
            R = orc.read_ref_dump(dump)
            out = {"n": np.uint64(R["n"]), "reads": R["reads"], "C": R["C"],
                   "suffix_count": R["suffix_count"], "prefix_count": R["prefix_count"],
                   "average_overlap": np.float64(R["average_overlap"])}
            for d in (0, 1):
                for k in ("bitmap", "sample", "reads_index", "istart", "iend", "idepth"):
                    out[f"{k}{d}"] = R[d][k]
            nodes = R["nodes"]
            fields = ("num_B", "num_F", "max_B", "max_F", "B_type", "F_type", "status",
                      "marked", "has_path_q")
            out["node_fields"] = np.array([[nd[f] for f in fields] for nd in nodes], np.uint64)
            off = np.zeros(len(nodes) + 1, np.uint64)
            off[1:] = np.cumsum([len(nd["edges"]) for nd in nodes])
            out["edge_off"] = off
            out["edges"] = (np.concatenate([nd["edges"] for nd in nodes])
                            if off[-1] else np.zeros((0, 3), np.uint64))
            # contigs at one thread
            wd = os.path.join(tmp, f"run_{name}_{width}")
            os.makedirs(wd)
            stdout, contigs = orc.run_reference(fa, p["L"], p["t"], wd, width=width, threads=1)
            out["contigs"] = np.frombuffer(contigs, np.uint8)
            np.savez_compressed(os.path.join(HERE, f"{name}_{width}.npz"), **out)
            print(name, width, "reads", R["n"], "edges", int(off[-1]), "contigs bytes",
                  len(contigs))
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
