"""Test infrastructure only: overlap-graph parity with the reference at a
workload's full size, where a full dump of both graphs is impractical.

    python oracle/full_parity.py ref  --config 5 --out ref.json [--genome-len G]
    python oracle/full_parity.py lib  --config 5 --out lib.json [--genome-len G]
    python oracle/full_parity.py compare ref.json lib.json

`ref` writes the workload's reads file (pq_synth_fasta) and runs the reference's
own construct_index + construct_string_graph on it (oracle/_ref/ref_graph_u64 or
_u32, built from the reference sources by oracle/Makefile) with one OpenMP
thread per usable core; the overlap graph is thread-count invariant.  `lib`
runs the device library on the same reads twice: imported from the same file
(pq_import_fasta, the drop-in's path) and generated in memory through the
device duplicate filter (the bench's path), then builds, searches (plain and
learned) and exports the graph.  Both sides digest the graph in the same byte
layout: 8-byte node records (num_B, num_F, B_type, F_type, status, marked) and
edge_t records with zeroed padding, in node order, B list then F list.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

INDEX_BITS = {1: 32, 2: 32, 3: 64, 4: 64, 5: 64}


def log(msg):
    print(msg, flush=True)


def file_digest(path):
    import xxhash
    h = xxhash.xxh3_128()
    with open(path, "rb") as f:
        while True:
            b = f.read(1 << 26)
            if not b:
                break
            h.update(b)
    return "xxh3_128:" + h.hexdigest()


def workload(pq, config, genome_len):
    cfg = pq.synth_preset(config)
    if genome_len:
        cfg.genome_len = genome_len
    return cfg


def side_ref(args):
    import pasqual_amd as pq
    cfg = workload(pq, args.config, args.genome_len)
    L, t, bits = cfg.len_min, cfg.min_overlap, INDEX_BITS[args.config]
    exe = os.path.join(HERE, "_ref", f"ref_graph_u{bits}")
    threads = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir=args.tmp) as wd:
        fa = os.path.join(wd, "reads.fa")
        t0 = time.time()
        pq.synth_fasta(cfg, fa)
        log(f"reads file written ({os.path.getsize(fa)} bytes, {time.time() - t0:.1f} s)")
        nodes, edges = os.path.join(wd, "nodes.bin"), os.path.join(wd, "edges.bin")
        t0 = time.time()
        r = subprocess.run([exe, fa, str(L), str(t), nodes, edges], cwd=wd, capture_output=True,
                           text=True, env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        wall = time.time() - t0
        if r.returncode != 0:
            raise SystemExit("reference failed: " + r.stderr[-2000:])
        log(r.stdout)
        line = [x for x in r.stdout.splitlines() if x.startswith("ref_graph ")][-1].split()
        out = {"side": "reference ref_graph_u%d (construct_index + construct_string_graph of "
                       "the reference sources)" % bits,
               "config": args.config, "genome_len": cfg.genome_len, "threads": threads,
               "reads": int(line[2]), "edges": int(line[4]), "overlap_len_sum": int(line[6]),
               "nodes_digest": file_digest(nodes), "edges_digest": file_digest(edges),
               "wall_s": round(wall, 1), "stdout": r.stdout[-1500:]}
    return out


def side_lib(args):
    import numpy as np
    import pasqual_amd as pq
    import bench
    cfg = workload(pq, args.config, args.genome_len)
    L, t, bits = cfg.len_min, cfg.min_overlap, INDEX_BITS[args.config]
    rec = 8 if bits == 32 else 16
    out = {"side": "device library", "config": args.config, "genome_len": cfg.genome_len}
    with tempfile.TemporaryDirectory(dir=args.tmp) as wd:
        fa = os.path.join(wd, "reads.fa")
        pq.synth_fasta(cfg, fa)
        log("reads file written")
        with pq.Context(0) as ctx:
            imp = pq.import_fasta(fa, L, index_bits=bits, ctx=ctx)
            os.remove(fa)
            n = imp["n"]
            log(f"imported {n} reads")
            # the bench's read set: generated reads through the device filter
            codes, off = pq.synth_reads(cfg)
            mem_codes, _, info = bench.dedup_in_memory(pq, ctx, codes, off, bits)
            del codes
            out["reads"] = n
            out["reads_in_memory_filter"] = info["kept"]
            out["read_sets_equal"] = bool(np.array_equal(mem_codes, imp["reads"]))
            del mem_codes
            ctx.upload(imp["reads"], np.arange(n + 1, dtype=np.uint64) * np.uint64(L), t)
            ctx.build_index()
            log("index built")
            for name, mode in (("plain", pq.SEARCH_PLAIN), ("learned", pq.SEARCH_LEARNED)):
                if mode == pq.SEARCH_LEARNED:
                    ctx.learned_train()
                ctx.find_overlaps(mode)
                _, E = ctx.graph_size()
                nb, eb = ctx.host_buffer(n * 8), ctx.host_buffer(max(1, E) * rec)
                s = ctx.graph_export(nb, eb, bits // 8)
                out[name] = {"edges": E, "overlap_len_sum": s,
                             "nodes_digest": bench._digest_buf(nb.array),
                             "edges_digest": bench._digest_buf(eb.array[:E * rec])}
                # the same graph handed over batch by batch during the search (the
                # bench's path), into fresh buffers
                nb2, eb2 = ctx.host_buffer(n * 8), ctx.host_buffer(max(1, E) * rec)
                E2, s2 = ctx.find_overlaps_export(nb2, eb2, bits // 8, mode)
                out[name]["export_during_search_equal"] = bool(
                    E2 == E and s2 == s and
                    bench._digest_buf(nb2.array) == out[name]["nodes_digest"] and
                    bench._digest_buf(eb2.array[:E * rec]) == out[name]["edges_digest"])
                for b in (nb, eb, nb2, eb2):
                    b.free()
                log(f"{name} search exported: {E} edges")
    return out


def compare(a, b):
    ref, lib = (a, b) if "reference" in a["side"] else (b, a)
    res = {"reads_equal": ref["reads"] == lib["reads"]}
    for name in ("plain", "learned"):
        g = lib[name]
        res[name] = {k: ref[k] == g[k] for k in ("edges", "overlap_len_sum", "nodes_digest",
                                                 "edges_digest")}
    res["all_equal"] = res["reads_equal"] and lib["read_sets_equal"] and all(
        all(v.values()) for k, v in res.items() if isinstance(v, dict)) and all(
        lib[name].get("export_during_search_equal", True) for name in ("plain", "learned"))
    return {"reference": ref, "library": lib, "equal": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["ref", "lib", "compare"])
    ap.add_argument("files", nargs="*")
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--genome-len", type=int, default=0)
    ap.add_argument("--out")
    ap.add_argument("--tmp", default=None)
    args = ap.parse_args()
    if args.what == "compare":
        a, b = (json.load(open(f)) for f in args.files)
        res = compare(a, b)
    else:
        res = side_ref(args) if args.what == "ref" else side_lib(args)
    txt = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")
    print(txt, flush=True)


if __name__ == "__main__":
    main()
