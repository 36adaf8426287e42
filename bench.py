"""Benchmark of the hot path on one MI355X: suffix-array + FM-index build of
both directions and the FM-index overlap search, plain and seeded by the
learned index, on a seeded synthetic workload.

Default workload (the largest single-GPU one): 200 Mbp genome with 5% repeat
families of 1-10 kbp at 0.5-3% divergence, 150 bp reads at 30x with 0.3%
substitutions (40 M reads, 6 G suffixes per direction: 64-bit device rows,
suffix sort in slices), t = 63.  The reads pass the reference's duplicate
filter first (index.c:145-405, bloom.c:91-135: Bloom filter of 36 bits per
record, 2 hashes, 64-bit index arithmetic), resolved on the device, so the
read set is the one the reference's U64 build imports from the same reads.
The headline search is the one seeded by the learned index.

--config 2: 5 Mbp genome with 100 interspersed copies of ten 1 kbp repeat
units at 1% divergence, 100 bp reads at 40x with 0.5% substitutions, t = 63,
32-bit index, reads through the same import.  --config 3 / 4: 50 Mbp genome,
read lengths uniform in 36..250 at 30x, 0.3% substitutions, t = 31 (4 = the
same reads with the learned index as the headline search).

A step = one index build (SA, BWT occurrence lines, reads_index for the
forward and the backward text) + one plain overlap search + (where the learned
search is the headline) the learned index trained on the freshly built
suffix arrays + one search seeded by it.  A search ends with the node records
and the edges (the reference's edge_t layout) in page-locked host memory: the
reference's overlap-graph timer (string_graph.c:737-1044) ends with the
graph in host memory too.  The reads are resident in HBM before the timed
region.  Prints one JSON line; with the default workload it also carries the
numbers of config 2 under "config2".

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 5]
                    [--cpu-baseline full|off] [--learned 1|0] [--extra-config 2|0]
"""
import argparse
import json
import math
import os
import re
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LINE_BYTES = 64
BLOOM_M, N_HASHFUNC = 36, 2          # the reference CLI's defaults (test.c:143-144)

WORKLOADS = {
    1: "1 Mbp uniform genome, 50 bp error-free reads at 20x, t=31, U32 index",
    2: "5 Mbp genome + 100 copies of ten 1 kbp repeat units (1% div), 100 bp reads at 40x, "
       "0.5% substitutions, t=63, U32 index",
    3: "50 Mbp uniform genome, read lengths U{36..250} at 30x, 0.3% substitutions, t=31, "
       "U64 index (device rows 32-bit: text < 2^32 codes)",
    4: "50 Mbp uniform genome, read lengths U{36..250} at 30x, 0.3% substitutions, t=31, "
       "U64 index (device rows 32-bit), learned index k=21, 64 hidden, 4096 leaves",
    5: "200 Mbp genome (uniform + 5% repeat families of 1-10 kbp at 0.5-3% divergence), "
       "150 bp reads at 30x, 0.3% substitutions, t=63, U64 index (device rows 64-bit), "
       "learned index k=21, 64 hidden, 4096 leaves",
}
# configurations whose headline search is the learned one
LEARNED_HEADLINE = (4, 5)
# index width of the reference build each configuration runs with
INDEX_BITS = {1: 32, 2: 32, 3: 64, 4: 64, 5: 64}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """Model name of the host CPU and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "usable_cores": len(os.sched_getaffinity(0))}


def usable_cores():
    return len(os.sched_getaffinity(0))


def _digest(path):
    import hashlib
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def _reference_run(exe, fasta, L, t, threads, timeout):
    """One run of a reference-driver program; (stdout timers, reads, contigs digest)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
    with tempfile.TemporaryDirectory() as wd:
        t0 = time.time()
        r = subprocess.run([exe, "--singlereads", fasta, "-l", str(L), "-t", str(t)], cwd=wd,
                           capture_output=True, text=True, env=env, timeout=timeout)
        wall = time.time() - t0
        digest = _digest(os.path.join(wd, "contigs.fa"))
    if r.returncode != 0:
        log("reference run failed: " + r.stderr[-500:])
        return None
    takes = [float(x) for x in re.findall(r"takes ([0-9.]+) ms", r.stdout)]
    n = int(re.search(r"has imported (\d+) reads", r.stdout).group(1))
    return {"takes": takes, "n": n, "digest": digest, "wall": wall}


def cpu_baseline_reference(fasta, L, t, n_expected, exe_name="pasqual_omp_u32", runs=3):
    """The reference OpenMP program (oracle/_ref) on the same reads file, timed
    by its own timers (index.c:643-680, string_graph.c:737-1044): the median of
    `runs` runs with one OpenMP thread per usable core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pq_oracle as orc
    exe = os.path.join(orc.REFDIR, exe_name)
    if not os.path.exists(exe):
        return None
    threads = usable_cores()
    res = []
    for _ in range(runs):
        r = _reference_run(exe, fasta, L, t, threads, 900)
        if r is None:
            return None
        res.append(r)
    n = res[0]["n"]
    digests = [r["digest"] for r in res]
    # takes: import, forward sort, backward sort, overlap graph, assembly
    graph = statistics.median(r["takes"][3] for r in res)
    build = statistics.median(r["takes"][1] + r["takes"][2] for r in res)
    return {"value": round(n / (graph / 1e3), 1), "unit": "reads/s", "cores": threads,
            "import_ms": statistics.median(r["takes"][0] for r in res),
            "kind": "reference",
            "sample": f"the full workload ({n} reads), reference {exe_name}, "
                      f"OMP_NUM_THREADS={threads} (usable cores), median of {runs} runs, "
                      "overlap-graph timer (string_graph.c:737-1044)", "reads": n,
            "graph_ms_runs": [r["takes"][3] for r in res],
            "sa_build_gsuffix_per_s": round(2 * n * L / (build / 1e3) / 1e9, 5),
            "wall_s": round(sum(r["wall"] for r in res), 1),
            "reads_match_gpu": None if n_expected is None else n == n_expected,
            # the reference's OpenMP assembly is not always the same from run to
            # run (unsynchronised passes, assemble.c:396-469): the digest most of
            # its runs gave, and every run's
            "contigs_sha256": max(digests, key=digests.count),
            "contigs_sha256_runs": digests,
            "contigs_same_every_run": len(set(digests)) == 1}


def dropin_run(fasta, L, t, width="u32"):
    """The reference's own driver, assembler and host files linked against this
    library (oracle/_ref/pasqual_hip_u32, built by `make -C oracle dropin`) on
    the same reads: its 'Building overlap graph' timer covers the device search,
    the copy of the edges and the host node build, like the reference's
    (string_graph.c:737-1044)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pq_oracle as orc
    exe = os.path.join(orc.REFDIR, f"pasqual_hip_{width}")
    if not os.path.exists(exe):
        return None
    threads = usable_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), PQ_TIMING="1")
    with tempfile.TemporaryDirectory() as wd:
        r = subprocess.run([exe, "--singlereads", fasta, "-l", str(L), "-t", str(t)], cwd=wd,
                           capture_output=True, text=True, env=env, timeout=600)
        digest = _digest(os.path.join(wd, "contigs.fa"))
    if r.returncode != 0:
        log("drop-in run failed: " + r.stderr[-500:])
        return None
    takes = [float(x) for x in re.findall(r"takes ([0-9.]+) ms", r.stdout)]
    n = int(re.search(r"has imported (\d+) reads", r.stdout).group(1))
    # phases inside the graph timer, printed by the library under PQ_TIMING=1
    phases = {k.replace(" ", "_") + "_ms": float(v)
              for k, v in re.findall(r"\[pq\] ([a-z ]+) ([0-9.]+) ms", r.stderr)}
    return {"reads": n, "import_ms": takes[0], "fw_index_ms": takes[1], "bw_index_ms": takes[2],
            "graph_ms": takes[3], "graph_phases": phases, "assembly_ms": takes[4],
            "graph_reads_per_s": round(n / (takes[3] / 1e3), 1), "host_threads": threads,
            "contigs_sha256": digest}


def cpu_baseline_port(pq, config, t):
    """Reads of uneven length are rejected by the reference's import, so the
    baseline is the oracle's single-threaded restatement (explicit-string overlap
    search and qsort suffix arrays) on a bounded sample: the same generator and
    coverage over a 0.5 Mbp genome."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pq_oracle as orc
    cfg = pq.synth_preset(config)
    cfg.genome_len = 500_000
    codes, off = pq.synth_reads(cfg)
    n = len(off) - 1
    t0 = time.perf_counter()
    orc.search_explicit(codes, off, t)
    ts = time.perf_counter() - t0
    t0 = time.perf_counter()
    for d in (0, 1):
        orc.suffix_array(orc.dir_text(codes, off, d), off)
    tb = time.perf_counter() - t0
    return {"value": round(n / ts, 1), "unit": "reads/s", "cores": 1, "kind": "port",
            "sample": f"{n} reads of the same generator over a 0.5 Mbp genome (same coverage "
                      "and lengths); oracle orc_search_explicit and orc_sa, one thread",
            "sa_build_gsuffix_per_s": round(2 * int(off[-1]) / tb / 1e9, 5),
            "wall_s": round(ts + tb, 1)}


def cpu_baseline_reference_sample(pq, config, t, genome_len=5_000_000):
    """The reference OpenMP program (pasqual_omp_u64) on a bounded sample of a
    workload too large for it: the same generator over a smaller genome."""
    cfg = pq.synth_preset(config)
    cfg.genome_len = genome_len
    L = cfg.len_min
    wd = tempfile.mkdtemp(prefix="pqbase_")
    fasta = os.path.join(wd, "sample.fa")
    pq.synth_fasta(cfg, fasta)
    base = cpu_baseline_reference(fasta, L, t, None, exe_name="pasqual_omp_u64")
    os.remove(fasta)
    os.rmdir(wd)
    if base:
        base["sample"] = (f"{base['reads']} reads of the same generator over a "
                          f"{genome_len // 1_000_000} Mbp genome (same repeat families, coverage, "
                          "length and errors), reference pasqual_omp_u64, "
                          f"OMP_NUM_THREADS={base['cores']} (usable cores), median of 3 runs, "
                          "overlap-graph timer (string_graph.c:737-1044)")
    return base


def dedup_in_memory(pq, ctx, codes, off, index_bits):
    """The reads import's duplicate filter (index.c:327-333, bloom.c:91-135) over
    generated fixed-length reads, on the device: the reads the reference keeps
    when it imports the same reads from a 2-line FASTA (filter of 36 bits per
    record, 2 hashes, hashes in the index width)."""
    import numpy as np
    n = len(off) - 1
    L = int(off[1] - off[0])
    t0 = time.perf_counter()
    keep, passes = ctx.dedup_reads(codes, np.ones(n, np.uint8), L, n * BLOOM_M, N_HASHFUNC,
                                   index_bits)
    kept = np.flatnonzero(keep)
    out = codes.reshape(n, L)[kept].reshape(-1)
    ms = (time.perf_counter() - t0) * 1e3
    return out, np.arange(len(kept) + 1, dtype=np.uint64) * np.uint64(L), {
        "records": n, "kept": int(len(kept)), "passes": passes, "ms": round(ms, 1),
        "filter_bits": n * BLOOM_M, "index_bits": index_bits}


def load_workload(pq, config, np):
    """(codes, off, t, L or None, FASTA path or None, import record, read set note)."""
    cfg = pq.synth_preset(config)
    t = cfg.min_overlap
    if config == 5:
        # 40 M reads: the 6.8 GB reads file is not written; the generated reads
        # pass the same duplicate filter on the device (dedup_in_memory)
        codes, off = pq.synth_reads(cfg)
        with pq.Context(0) as c:
            codes, off, info = dedup_in_memory(pq, c, codes, off, INDEX_BITS[config])
        return codes, off, t, None, None, info, \
            "after the reference's duplicate filter (Bloom filter resolved on the device)"
    if cfg.len_min == cfg.len_max:
        # fixed-length reads go through the reference's import (duplicate removal)
        L = cfg.len_min
        fasta = os.path.join(tempfile.mkdtemp(prefix="pqbench_"), "reads.fa")
        pq.synth_fasta(cfg, fasta)
        imp = pq.import_fasta(fasta, L, index_bits=INDEX_BITS[config])
        n = imp["n"]
        off = np.arange(n + 1, dtype=np.uint64) * np.uint64(L)
        return imp["reads"], off, t, L, fasta, \
            {"records": n + imp["removed"], "kept": n, "ms": round(imp["ms"], 1)}, \
            "after duplicate removal (reads import from the FASTA file)"
    codes, off = pq.synth_reads(cfg)
    return codes, off, t, None, None, None, "as generated (uneven lengths: no duplicate removal)"


def build_model_bytes(lmax, W):
    """(rounds, bytes per suffix) of the prefix-doubling byte model of a suffix-array
    build: R = 1 + ceil(log2(lmax / 29)) rounds from a 64-bit initial key of 29
    two-bit codes and a clamped length, each round one key + index pass (read and
    write, 2 * (8 + W)) and a rank read, gather and write (3 * W), then the BWT
    gather and the occurrence write (W + 2 + (40 + 5 W) / 64); W = index width."""
    rounds = 1 + max(0, math.ceil(math.log2(lmax / 29)))
    return rounds, rounds * (2 * (8 + W) + 3 * W) + (W + 2 + (40 + 5 * W) / 64)


def _digest_buf(arr):
    """Digest of a host buffer (xxh3-128 when xxhash is importable, else SHA-256)."""
    try:
        import xxhash
        return "xxh3_128:" + xxhash.xxh3_128(memoryview(arr)).hexdigest()
    except ImportError:
        import hashlib
        return "sha256:" + hashlib.sha256(memoryview(arr)).hexdigest()


def _mean(xs):
    return sum(xs) / len(xs) if xs else 0.0


VERIFY_ROW_BYTES = 8 + LINE_BYTES   # a suffix-array row and the text line left of it


def search_roofline(positions, lines, rows, kernel_ms, launches, copy_gbs, read_gbs, gather,
                    kname, n, steps):
    """Roofline of the candidate-search phase (search kernel + verification) from
    its live HIP-event times: algorithmic bytes = 64 B per distinct occurrence
    line a backward-search step needs (one line when the rows l and u+1 share a
    line, else two) + 72 B per row finished on the text (its suffix-array entry
    and the 64-byte text line left of it); the kernels count both, the same
    counts as the oracle's restatement of the chains; against the HBM peak; also
    the 64 B per rank position form and the measured ceilings."""
    avg_launch_ms = kernel_ms / max(1, launches)
    alg = (lines * LINE_BYTES + rows * VERIFY_ROW_BYTES) / max(1, launches)
    pos_bytes = positions / max(1, launches) * LINE_BYTES
    achieved = alg / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms else 0.0
    best = max(gather["lane"], gather["quad"]) if gather else None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kname, "alg_bytes_per_launch": int(alg),
            "alg_bytes_basis": "64 B x distinct occurrence lines per backward-search step + "
                               "72 B x rows verified on the text",
            "lines_per_read": round(lines / (n * steps), 1),
            "verified_rows_per_read": round(rows / (n * steps), 1),
            "rank_positions_per_read": round(positions / (n * steps), 1),
            "position_bytes_per_launch": int(pos_bytes),
            "position_bytes_gbs": round(pos_bytes / (avg_launch_ms / 1e3) / 1e9, 1)
            if avg_launch_ms else None,
            "avg_launch_ms": round(avg_launch_ms, 3),
            "launches_per_search": round(launches / steps, 2),
            "copy_ceiling": copy_gbs, "read_ceiling": read_gbs,
            "frac_of_read_ceiling": round(achieved / read_gbs, 4) if read_gbs else None,
            "scattered_line_ceilings": gather,
            "frac_of_scattered_ceiling": round(achieved / best, 4) if best else None}


def reference_record_match(config, edges, digest):
    """Whether this run's plain graph (node and edge_t record digests, sum of overlap
    lengths) equals the committed full-size record of the reference's own graph for
    this configuration (profiles/config<N>_parity.json, written by
    oracle/full_parity.py from the reference's construct_string_graph), or None
    when there is no record or the digests are of another kind."""
    path = os.path.join(ROOT, "profiles", f"config{config}_parity.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        ref = json.load(f)["reference"]
    if not str(ref.get("edges_digest", "")).split(":")[0] == digest[1].split(":")[0]:
        return None
    return (ref["edges"] == edges and ref["nodes_digest"] == digest[0] and
            ref["edges_digest"] == digest[1] and ref["overlap_len_sum"] == digest[2])


def extract_roofline(items, edges, extract_ms, steps):
    """The extraction phase (all tiers of one search) against the HBM peak: 64 B per
    candidate item (its extension key, a scattered read of the text) + 8 B per item
    of candidate record and item bookkeeping + 16 B per edge written."""
    per_search = items / max(1, steps)
    alg = per_search * (LINE_BYTES + 8) + edges * 16
    gbs = alg / (extract_ms / 1e3) / 1e9 if extract_ms else 0.0
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "items_per_search": int(per_search),
            "alg_bytes_per_search": int(alg), "ms_per_search": round(extract_ms, 3),
            "alg_bytes_basis": "64 B per item (its key from the text) + 8 B per item "
                               "(candidate record) + 16 B per edge"}


def committed_traffic(config, kname, field="traffic_bytes"):
    """Bytes past the L2 per launch of kernel kname on this configuration (to HBM,
    or served by the memory-side cache) from the committed counter summary
    (profiles/config<N>_traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for the plain copy kernel), or None.  field
    "traffic_by_request_size": reads counted by request size instead."""
    tpath = os.path.join(ROOT, "profiles", f"config{config}_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        kernels = json.load(f)["kernels"]
    tk = next((v for k, v in kernels.items() if k == kname or k.startswith(kname + "<")), None)
    if not tk:
        return None
    # the mean over the step's dispatches, as the launch time it is set against
    return int(tk[field]) if tk.get(field) is not None else None


def run_workload(pq, np, config, args, cpu_baseline):
    t0 = time.time()
    codes, off, t, L, fasta, imp, how = load_workload(pq, config, np)
    n = len(off) - 1
    N = int(off[-1])
    log(f"workload config {config}: {n} reads {how}, N={N} per direction, "
        f"prepared in {time.time() - t0:.1f}s")
    headline_learned = config in LEARNED_HEADLINE
    learned_on = bool(args.learned) or headline_learned
    index_bytes = INDEX_BITS[config] // 8
    rec_bytes = 8 if index_bytes == 4 else 16

    ctx = pq.Context(0)
    ctx.upload(codes, off, t)
    del codes
    if learned_on:
        # the builds write the learned index's sampled keys as they sort
        ctx.set_learned_keys(int(ctx.learned_defaults()["sample_stride"]))
    for _ in range(args.warmup):
        ctx.build_index()
        ctx.find_overlaps()
    if args.warmup == 0:
        ctx.build_index()
    ctx.sync()
    # untimed: the extraction shape and search kernel for this read set (identical
    # results), for the plain search and, after a training, the learned one
    tune = ctx.autotune()
    tune_l = None
    if learned_on:
        ctx.learned_train()
        tune_l = ctx.autotune(pq.SEARCH_LEARNED)
    log(f"warm-up done; plain tuning {tune}; learned tuning {tune_l}")
    _, E = ctx.graph_size()
    nodes_buf = ctx.host_buffer(max(1, n) * 8)
    edges_buf = ctx.host_buffer(max(1, E) * rec_bytes)

    def use(tn):
        ctx.set_tuning(tn["shape"])
        ctx.set_find_chains(tn["find_chains"])
        ctx.set_verify(tn["verify_rows"])
        if tn is tune_l:
            ctx.set_seed_verify(tn["seed_rows"])

    def export():
        nonlocal edges_buf
        _, e = ctx.graph_size()
        if e * rec_bytes > edges_buf.nbytes:
            edges_buf.free()
            edges_buf = ctx.host_buffer(e * rec_bytes)
        return ctx.graph_export(nodes_buf, edges_buf, index_bytes)

    def search_export(mode):
        """One search whose node and edge_t records reach page-locked host memory
        batch by batch while later batches are searched (pq_find_overlaps_export);
        an edges buffer too small is grown and the result exported after the
        search.  Returns the overlap-length sum."""
        try:
            return ctx.find_overlaps_export(nodes_buf, edges_buf, index_bytes, mode)[1]
        except pq.PqError:
            _, e = ctx.graph_size()
            if e * rec_bytes <= edges_buf.nbytes:
                raise
            return export()

    def kernels_ms(tm):
        """Event-timed device phases of the last search: seeding, candidate search
        and verification, extraction."""
        return tm["learned_ms"] + tm["find_ms"] + tm["extract_ms"]

    P = {"build": [], "search": [], "search_dev": [], "export": [], "kernel": [], "extract": [],
         "launches": 0, "positions": 0, "lines": 0, "rows": 0, "items": 0}
    Lr = {"train": [], "search": [], "search_dev": [], "export": [], "kernel": [], "seed": [],
          "seed_kernel": [], "extract": [], "launches": 0, "positions": 0, "lines": 0, "rows": 0,
          "seed_rows": 0, "probes": 0, "items": 0, "same": True}
    digests = {}
    t_start = time.perf_counter()
    for step in range(args.steps):
        a = time.perf_counter()
        ctx.build_index()
        ctx.sync()
        b = time.perf_counter()
        use(tune)
        len_sum = search_export(pq.SEARCH_PLAIN)
        d = time.perf_counter()
        tm = ctx.timing()
        P["build"].append((b - a) * 1e3)
        P["search"].append((d - b) * 1e3)
        P["search_dev"].append(kernels_ms(tm))
        P["export"].append((d - b) * 1e3 - kernels_ms(tm))
        P["kernel"].append(tm["find_kernel_ms"])
        P["launches"] += tm["find_launches"]
        P["positions"] += tm["find_positions"]
        P["lines"] += tm["find_lines"]
        P["rows"] += tm["verify_rows"]
        P["extract"].append(tm["extract_ms"])
        P["items"] += tm["extract_items"]
        if step == 0:
            _, E = ctx.graph_size()
            digests["plain"] = (_digest_buf(nodes_buf.array), _digest_buf(
                edges_buf.array[:E * rec_bytes]), len_sum)
        if learned_on:
            e0 = time.perf_counter()
            ctx.learned_train()
            e1 = time.perf_counter()
            use(tune_l)
            ls = search_export(pq.SEARCH_LEARNED)
            e3 = time.perf_counter()
            tl = ctx.timing()
            Lr["train"].append((e1 - e0) * 1e3)
            Lr["search"].append((e3 - e1) * 1e3)
            Lr["search_dev"].append(kernels_ms(tl))
            Lr["export"].append((e3 - e1) * 1e3 - kernels_ms(tl))
            Lr["kernel"].append(tl["find_kernel_ms"])
            Lr["seed"].append(tl["learned_ms"])
            Lr["seed_kernel"].append(tl["seed_kernel_ms"])
            Lr["items"] += tl["extract_items"]
            Lr["launches"] += tl["find_launches"]
            Lr["positions"] += tl["find_positions"]
            Lr["lines"] += tl["find_lines"]
            Lr["rows"] += tl["verify_rows"]
            Lr["seed_rows"] += tl["seed_rows"]
            Lr["extract"].append(tl["extract_ms"])
            Lr["probes"] += sum(ctx.learned_stats()["probes"])
            if step == 0:
                _, E = ctx.graph_size()
                got = (_digest_buf(nodes_buf.array), _digest_buf(
                    edges_buf.array[:E * rec_bytes]), ls)
                Lr["same"] = got == digests["plain"]
    ctx.sync()
    total_s = time.perf_counter() - t_start
    _, edges = ctx.graph_size()
    log("timed steps done")
    st = ctx.learned_stats() if learned_on else None

    # the same searches as one HIP graph each (captured by the first, untimed call)
    reps = min(args.steps, 5)
    graph = {}
    for name, mode, tn in (("plain", pq.SEARCH_PLAIN, tune),
                           ("learned", pq.SEARCH_LEARNED, tune_l)):
        if tn is None:
            continue
        use(tn)
        ctx.find_overlaps(mode | pq.SEARCH_GRAPH)
        ms = []
        for _ in range(reps):
            a = time.perf_counter()
            ctx.find_overlaps(mode | pq.SEARCH_GRAPH)
            ms.append((time.perf_counter() - a) * 1e3)
        export()
        _, E = ctx.graph_size()
        graph[f"{name}_search_device_ms"] = round(_mean(ms), 3)
        graph[f"{name}_edges_identical"] = (
            (_digest_buf(nodes_buf.array), _digest_buf(edges_buf.array[:E * rec_bytes]))
            == digests["plain"][:2])
    log("graph-mode searches done")

    # measured streaming ceilings of this GPU, reported beside the published HBM
    # peak (untimed, best of 3): a copy (1 GiB read + 1 GiB written) and a read of
    # 1 GiB; scattered 64-byte line reads over a buffer the size of both indexes'
    # occurrence lines (whole lines per lane, independent and in dependent
    # chains, and four lanes per line)
    copy_gbs = read_gbs = gather = None
    try:
        copy_gbs = round(2 * (1 << 30) / min(ctx.probe_copy(1 << 30) for _ in range(3)) / 1e6, 1)
        read_gbs = round((1 << 30) / min(ctx.probe_read(1 << 30) for _ in range(3)) / 1e6, 1)
        occ_bytes = 2 * ((N + 127) // 128) * LINE_BYTES
        nbytes = (1 << 25) * LINE_BYTES
        gather = {name: round(nbytes / min(ctx.probe_gather(occ_bytes, mode) for _ in range(3))
                              / 1e6, 1)
                  for name, mode in (("lane", 0), ("lane_chained", 1), ("quad", 2))}
        gather["footprint_bytes"] = occ_bytes
        log(f"scattered line reads over {occ_bytes / 1e9:.2f} GB: {gather}")
    except pq.PqError as e:
        log(f"ceilings not measured: {e}")
    if args.probe_copy:
        pms = ctx.probe_copy(args.probe_copy << 20)
        log(f"copy probe: {args.probe_copy} MiB in {pms:.3f} ms")
    row_bytes, slices = ctx.layout()
    lens = np.diff(off)
    read_len = int(lens[0]) if int(lens.min()) == int(lens.max()) else \
        f"{int(lens.min())}..{int(lens.max())}"
    kname = {0: "k_find_list", 1: "k_find1", 2: "k_find2", 3: "k_find1q", 4: "k_find_list"}
    avg_build = _mean(P["build"])
    avg_search = _mean(P["search"])
    gsuffix_per_s = 2 * N / (avg_build / 1e3) / 1e9
    # the suffix-array build against the HBM peak by the prefix-doubling byte model;
    # W = the width of the device rows this build used
    W = row_bytes
    rounds, b_suffix = build_model_bytes(int(lens.max()), W)
    build_gbs = gsuffix_per_s * b_suffix
    build_roofline = {"bound": "hbm", "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                      "model_bytes_per_suffix": round(b_suffix, 2), "rounds": rounds,
                      "index_width": W}
    plain_roof = search_roofline(P["positions"], P["lines"], P["rows"], sum(P["kernel"]),
                                 P["launches"],
                                 copy_gbs, read_gbs, gather, kname[tune["find_chains"]], n,
                                 args.steps)
    plain_roof["traffic"] = committed_traffic(config, plain_roof["kernel"])
    plain_roof["traffic_by_request_size"] = committed_traffic(config, plain_roof["kernel"],
                                                              "traffic_by_request_size")
    res = {
        "config": config, "n": n, "N": N, "t": t, "read_len": read_len, "row_bytes": row_bytes,
        "slices": slices, "how": how, "import": imp, "steps": args.steps,
        "ms_per_step": total_s * 1e3 / args.steps,
        "plain_search_reads_per_s": round(n / (avg_search / 1e3), 1),
        "plain_search_ms": round(avg_search, 3),
        "plain_search_kernels_ms": round(_mean(P["search_dev"]), 3),
        "plain_search_tail_ms": round(_mean(P["export"]), 3),
        "sa_build_gsuffix_per_s": round(gsuffix_per_s, 3),
        "build_ms": round(avg_build, 3),
        "edges": edges, "edge_record_bytes": rec_bytes,
        "edges_digest": digests["plain"][1], "nodes_digest": digests["plain"][0],
        "overlap_len_sum": digests["plain"][2],
        "matches_reference_record": reference_record_match(config, edges, digests["plain"]),
        "extraction_shape": {k: tune[k] for k in ("group_lanes", "items")},
        "find_kernel": tune["find_chains"], "verify_rows": tune["verify_rows"],
        "plain_roofline": plain_roof, "build_roofline": build_roofline, "graph_mode": graph,
        "plain_phases_ms": {plain_roof["kernel"] + " + k_verify": round(sum(P["kernel"]) / args.steps, 3),
                            "extraction": round(_mean(P["extract"]), 3)},
        "extract_roofline": extract_roofline(P["items"], edges, _mean(P["extract"]), args.steps),
    }
    if learned_on:
        lsearch = _mean(Lr["search"])
        lookups = sum(st["lookups"])
        lroof = search_roofline(Lr["positions"], Lr["lines"], Lr["rows"], sum(Lr["kernel"]),
                                Lr["launches"], copy_gbs, read_gbs, gather,
                                kname[tune_l["find_chains"]], n, args.steps)
        lroof["traffic"] = committed_traffic(config, lroof["kernel"] + "_seeded")
        lroof["traffic_by_request_size"] = committed_traffic(
            config, lroof["kernel"] + "_seeded", "traffic_by_request_size")
        # the seed kernel: model, correction probes (sampled keys and suffix-array
        # rows), and the rows it finishes on the text
        seed_ms = sum(Lr["seed"])                 # prediction + seed kernel
        seed_kernel = sum(Lr["seed_kernel"])      # the seed kernel alone
        seed_launches = Lr["launches"]   # one per batch, as the search kernel
        seed_alg = (Lr["probes"] * LINE_BYTES + Lr["seed_rows"] * VERIFY_ROW_BYTES) / \
            max(1, seed_launches)
        seed_avg = seed_kernel / max(1, seed_launches)
        seed_gbs = seed_alg / (seed_avg / 1e3) / 1e9 if seed_avg else 0.0
        seed_roof = {"bound": "hbm", "achieved": round(seed_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(seed_gbs / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic(config, "k_learned_seed"),
                     "traffic_by_request_size": committed_traffic(config, "k_learned_seed",
                                                                  "traffic_by_request_size"),
                     "kernel": "k_learned_seed", "alg_bytes_per_launch": int(seed_alg),
                     "alg_bytes_basis": "64 B per correction probe (a sampled key, or a "
                                        "suffix-array row and its text) + 72 B per row "
                                        "finished on the text",
                     "probes_per_read": round(Lr["probes"] / (n * args.steps), 2),
                     "verified_rows_per_read": round(Lr["seed_rows"] / (n * args.steps), 1),
                     "avg_launch_ms": round(seed_avg, 3),
                     "launches_per_search": round(seed_launches / args.steps, 2),
                     "copy_ceiling": copy_gbs, "read_ceiling": read_gbs}
        train = _mean(Lr["train"])
        res["learned"] = {
            "reads_per_s": round(n / (lsearch / 1e3), 1), "search_ms": round(lsearch, 3),
            "search_kernels_ms": round(_mean(Lr["search_dev"]), 3),
            "search_tail_ms": round(_mean(Lr["export"]), 3),
            "seed_ms": round(_mean(Lr["seed"]), 3),
            "seed_kernel_ms": round(_mean(Lr["seed_kernel"]), 3),
            "find_kernel_ms": round(_mean(Lr["kernel"]), 3),
            "train_ms": round(train, 1),
            "train_ms_in_library": round(st["train_ms"], 1),
            "seed_depth": st["depth"], "model_k": st["k"], "model_segments": st["segments"],
            "table_size": list(st["table_size"]),
            "mean_window": [round(x, 2) for x in st["mean_window"]],
            "max_window": list(st["max_window"]),
            "probes_per_lookup": round(sum(st["probes"]) / max(1, lookups), 2),
            "absent_chains": st["absent"],
            "fit_loss": [float(f"{x:.3g}") for x in st["fit_loss"]],
            "train_loss": [float(f"{x:.3g}") for x in st["train_loss"]],
            "edges_identical_to_plain": bool(Lr["same"]),
            "extraction_shape": {k: tune_l[k] for k in ("group_lanes", "items")},
            "find_kernel": tune_l["find_chains"], "verify_rows": tune_l["verify_rows"],
            "seed_rows": tune_l["seed_rows"],
            "extract_ms": round(_mean(Lr["extract"]), 3),
            "phases_ms": {"k_learned_predict": round((seed_ms - seed_kernel) / args.steps, 3),
                          "k_learned_seed": round(seed_kernel / args.steps, 3),
                          lroof["kernel"] + " + k_verify": round(sum(Lr["kernel"]) / args.steps, 3),
                          "extraction": round(_mean(Lr["extract"]), 3)},
            "roofline": lroof, "seed_roofline": seed_roof,
            "extract_roofline": extract_roofline(Lr["items"], edges, _mean(Lr["extract"]),
                                                 args.steps),
        }
        res["learned_vs_plain_search"] = round(avg_search / lsearch, 3)
        # build-once / search-once pipeline: the learned search pays for its table
        # and its training on every freshly built index
        res["learned_end_to_end"] = {
            "plain_ms": round(avg_search, 3), "learned_ms": round(train + lsearch, 3),
            "train_ms": round(train, 1), "search_ms": round(lsearch, 3),
            "learned_vs_plain": round(avg_search / (train + lsearch), 3)}
    nodes_buf.free()
    edges_buf.free()
    ctx.close()

    base = dropin = None
    if cpu_baseline:
        log(f"timing the CPU baseline of config {config} ...")
        if L is not None:
            base = cpu_baseline_reference(fasta, L, t, n,
                                          exe_name=f"pasqual_omp_u{INDEX_BITS[config]}")
            log("timing the reference driver on this library ...")
            dropin = dropin_run(fasta, L, t, width=f"u{INDEX_BITS[config]}")
            if dropin and base:
                # equal to what the reference wrote in its own runs (the reference
                # runs of this line: how many of them it equals)
                dropin["contigs_identical_to_reference"] = (
                    dropin["contigs_sha256"] is not None and
                    dropin["contigs_sha256"] in base["contigs_sha256_runs"])
                dropin["reference_runs_equal"] = sum(
                    1 for x in base["contigs_sha256_runs"] if x == dropin["contigs_sha256"])
        elif config == 5:
            base = cpu_baseline_reference_sample(pq, config, t)
        else:
            base = cpu_baseline_port(pq, config, t)
        if base:
            base["host"] = host_cpu()
            if imp and base.get("import_ms"):
                imp["reference_ms"] = base["import_ms"]
    if fasta:
        os.remove(fasta)
        os.rmdir(os.path.dirname(fasta))
    res["cpu_baseline"] = base
    res["dropin"] = dropin
    return res


def summary(res):
    """The numbers of one configuration carried as extra keys of the line."""
    keys = ("plain_search_reads_per_s", "plain_search_ms", "plain_search_kernels_ms",
            "sa_build_gsuffix_per_s", "build_ms", "edges", "edges_digest", "find_kernel",
            "extraction_shape", "learned_vs_plain_search", "learned_end_to_end", "dropin",
            "import", "graph_mode")
    out = {"workload": f"config {res['config']}: {WORKLOADS[res['config']]}", "reads": res["n"],
           "suffixes_per_direction": res["N"], "ms_per_step": round(res["ms_per_step"], 3)}
    out.update({k: res[k] for k in keys if k in res})
    out["search_roofline"] = res["plain_roofline"]
    out["build_roofline"] = res["build_roofline"]
    if "learned" in res:
        out["learned_reads_per_s"] = res["learned"]["reads_per_s"]
        out["learned_search_ms"] = res["learned"]["search_ms"]
        out["learned_edges_identical_to_plain"] = res["learned"]["edges_identical_to_plain"]
    base = res.get("cpu_baseline")
    if base:
        out["cpu_baseline"] = base
        out["speedup_vs_cpu_search"] = round(res["plain_search_reads_per_s"] / base["value"], 1)
        out["speedup_vs_cpu_sa_build"] = round(
            res["sa_build_gsuffix_per_s"] / base["sa_build_gsuffix_per_s"], 1)
    return out


def _kernel_key(name):
    """Kernel name with template arguments, without the parameter list."""
    n = name[5:] if name.startswith("void ") else name
    depth = 0
    for i, ch in enumerate(n):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0 and i > 0:
            return n[:i].strip()
    return n.strip()


def traffic_summary(fetch_csv, write_csv, out_path, probe_mib, commands, md_path=None,
                    title="HBM/fabric traffic per dispatch", sizes_csv=None):
    """Per-dispatch HBM traffic of one step from two rocprofv3 --pmc runs of the
    same bench command (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled
    by what its counter reads for the plain copy kernel of probe_mib MiB in each
    direction.  The step is the last index build and the searches after it; a
    search kernel launched right after the learned seed kernel is listed as
    <name>_seeded<...>.  With sizes_csv (a third run collecting
    TCC_EA0_RDREQ_{32B,64B,128B}_sum) each kernel also gets its read bytes counted
    by request size, which needs no scale: FETCH_SIZE counts a 128-byte request
    as 64 bytes on this GPU (hence the copy kernel's scale of 2), while a kernel
    of scattered 64-byte reads is counted exactly, so the scaled figure
    overstates such kernels' reads up to twice."""
    import csv

    def load(path):
        with open(path) as f:
            rows = list(csv.DictReader(f))
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        return rows

    F, W = load(fetch_csv), load(write_csv)
    # optional third run: read requests by size, one row per (dispatch, counter)
    Q = None
    if sizes_csv:
        Q = {}
        for r in load(sizes_csv):
            Q.setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = float(r["Counter_Value"])

    def read_bytes(did):
        q = Q.get(did, {}) if Q is not None else {}
        return int(32 * q.get("TCC_EA0_RDREQ_32B_sum", 0) + 64 * q.get("TCC_EA0_RDREQ_64B_sum", 0)
                   + 128 * q.get("TCC_EA0_RDREQ_128B_sum", 0))

    names = [_kernel_key(r["Kernel_Name"]) for r in F]
    if len(F) != len(W) or names != [_kernel_key(r["Kernel_Name"]) for r in W]:
        raise SystemExit("the two counter runs dispatched different kernels")
    probe = [i for i, n in enumerate(names) if n.startswith("k_copy_probe")]
    if not probe:
        raise SystemExit("no copy probe in the counter runs (run the bench with --probe-copy)")
    pb = probe_mib << 20
    fs = pb / (float(F[probe[0]]["Counter_Value"]) * 1024)
    ws = pb / (float(W[probe[0]]["Counter_Value"]) * 1024)
    # each build starts with one key histogram per direction: the step is the last build
    kh = [i for i, n in enumerate(names) if n.startswith("k_key_hist")]
    start = kh[-2] if len(kh) >= 2 else 0
    kernels, prev = {}, ""
    for i in range(start, len(F)):
        n = names[i]
        if n.startswith("k_copy_probe") or n.startswith("k_read_probe") or \
                n.startswith("k_gather_probe"):
            continue
        key = n
        if prev.startswith("k_learned_seed") and n.startswith("k_find"):
            key = n.replace("<", "_seeded<", 1) if "<" in n else n + "_seeded"
        e = kernels.setdefault(key, {"fetch_kib": [], "write_kib": [], "traffic_bytes": [],
                                     "read_request_bytes": []})
        fk, wk = float(F[i]["Counter_Value"]), float(W[i]["Counter_Value"])
        e["fetch_kib"].append(round(fk, 1))
        e["write_kib"].append(round(wk, 1))
        e["traffic_bytes"].append(int(fk * 1024 * fs + wk * 1024 * ws))
        if Q is not None:
            e["read_request_bytes"].append(read_bytes(int(F[i]["Dispatch_Id"])))
        prev = n
    for e in kernels.values():
        c, tb, rq = len(e["traffic_bytes"]), e["traffic_bytes"], e.pop("read_request_bytes")
        e.update({"dispatches": c, "fetch_kib": round(sum(e["fetch_kib"]) / c, 1),
                  "write_kib": round(sum(e["write_kib"]) / c, 1), "traffic_bytes": sum(tb) // c,
                  "traffic_min": min(tb), "traffic_max": max(tb)})
        if Q is not None:
            # reads counted exactly by request size, writes by the scaled WRITE_SIZE
            e["traffic_by_request_size"] = int(sum(rq) / c + e["write_kib"] * 1024 * ws)
    res = {"command_fetch": commands[0], "command_write": commands[1],
           "command_read_sizes": commands[2] if Q is not None else None,
           "probe_bytes": pb, "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
           "dispatches": "the last index build of the run and the searches after it; "
                         "per kernel the mean KiB and bytes per dispatch; a search kernel "
                         "launched right after the learned seed kernel is listed as "
                         "<name>_seeded",
           "counts": "FETCH_SIZE and WRITE_SIZE count the bytes that leave the L2 for the "
                     "fabric: served by HBM or by the memory-side cache",
           "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(res, f, indent=0)
        f.write("\n")
    if md_path:
        out = [f"# {title}", "", "Counter runs, each in its own process:", "", "```",
               commands[0], commands[1], "```", "",
               f"Scale: the plain copy kernel `k_copy_probe` moves {pb} bytes each way; "
               f"FETCH_SIZE reads {float(F[probe[0]]['Counter_Value']):.0f} KiB for it (scale "
               f"{fs:.3f}), WRITE_SIZE {float(W[probe[0]]['Counter_Value']):.0f} KiB (scale "
               f"{ws:.3f}).  The counters count what leaves the L2 for the fabric, served "
               "by HBM or by the memory-side cache.  Dispatches: " + res["dispatches"] + ".",
               "", "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled traffic "
               "(GB, mean) | min | max | reads by request size + writes (GB) |",
               "|---|---:|---:|---:|---:|---:|---:|---:|"]
        if Q is not None:
            out[7:7] = [commands[2]]
        top = sorted(kernels.items(), key=lambda x: -x[1]["traffic_bytes"] * x[1]["dispatches"])
        for k, e in top[:20]:
            exact = e.get("traffic_by_request_size")
            out.append(f"| `{k}` | {e['dispatches']} | {e['fetch_kib']:.0f} | "
                       f"{e['write_kib']:.0f} | {e['traffic_bytes'] / 1e9:.2f} | "
                       f"{e['traffic_min'] / 1e9:.2f} | {e['traffic_max'] / 1e9:.2f} | "
                       + (f"{exact / 1e9:.2f} |" if exact is not None else "- |"))
        with open(md_path, "w") as f:
            f.write("\n".join(out) + "\n")


def stats_summary(trace_csv, stats_csv, bench_json, out_md, command, title):
    """Markdown summary of one rocprofv3 --kernel-trace --stats run of the bench: the
    step's kernels (the last index build and the searches after it, from the trace)
    and the whole run's top kernels (from the stats file), beside the bench line's
    own launch times measured with HIP events in the same run."""
    import csv
    with open(trace_csv) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Dispatch_Id"]))
    names = [_kernel_key(r["Kernel_Name"]) for r in rows]
    kh = [i for i, n in enumerate(names) if n.startswith("k_key_hist")]
    start = kh[-2] if len(kh) >= 2 else 0
    step = {}
    for r, n in zip(rows[start:], names[start:]):
        e = step.setdefault(n, [0, 0.0])
        e[0] += 1
        e[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
    with open(stats_csv) as f:
        stats = list(csv.DictReader(f))
    line = {}
    if bench_json and os.path.exists(bench_json):
        with open(bench_json) as f:
            txt = [x for x in f.read().splitlines() if x.startswith("{")]
        line = json.loads(txt[-1]) if txt else {}
    out = [f"# {title}", "", "Command (run on the GPU box from the repository root):", "",
           "```", command, "```", ""]
    if line:
        roof = line.get("roofline", {})
        out += [f"Bench line of the same run: value {line.get('value')} {line.get('unit')}, "
                f"{line.get('ms_per_step')} ms per step; headline kernel "
                f"`{roof.get('kernel')}` averages {roof.get('avg_launch_ms')} ms per launch "
                "(HIP events on the library's stream).", ""]
    out += ["## The step (last index build and the searches after it)", "",
            "| kernel | launches | total ms | average ms |", "|---|---:|---:|---:|"]
    for n, (c, t) in sorted(step.items(), key=lambda x: -x[1][1])[:30]:
        out.append(f"| `{n}` | {c} | {t:.2f} | {t / c:.3f} |")
    out += ["", "## Whole run (rocprofv3 --stats)", "",
            "| kernel | calls | total ms | average ms | % |", "|---|---:|---:|---:|---:|"]
    for r in stats[:25]:
        out.append(f"| `{_kernel_key(r['Name'])}` | {r['Calls']} | "
                   f"{int(r['TotalDurationNs']) / 1e6:.2f} | {float(r['AverageNs']) / 1e6:.3f} | "
                   f"{float(r['Percentage']):.1f} |")
    with open(out_md, "w") as f:
        f.write("\n".join(out) + "\n")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--stats-summary":
        sp = argparse.ArgumentParser(prog="bench.py --stats-summary")
        sp.add_argument("trace_csv")
        sp.add_argument("stats_csv")
        sp.add_argument("out")
        sp.add_argument("--bench-json", default="")
        sp.add_argument("--command", default="")
        sp.add_argument("--title", default="Kernel statistics")
        a = sp.parse_args(sys.argv[2:])
        stats_summary(a.trace_csv, a.stats_csv, a.bench_json, a.out, a.command, a.title)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic-summary":
        tp = argparse.ArgumentParser(prog="bench.py --traffic-summary")
        tp.add_argument("fetch_csv")
        tp.add_argument("write_csv")
        tp.add_argument("out")
        tp.add_argument("--probe-mib", type=int, default=1024)
        tp.add_argument("--command-fetch", default="")
        tp.add_argument("--command-write", default="")
        tp.add_argument("--md", default="")
        tp.add_argument("--sizes-csv", default="")
        tp.add_argument("--command-sizes", default="")
        tp.add_argument("--title", default="HBM/fabric traffic per dispatch")
        a = tp.parse_args(sys.argv[2:])
        traffic_summary(a.fetch_csv, a.write_csv, a.out, a.probe_mib,
                        (a.command_fetch, a.command_write, a.command_sizes), a.md, a.title,
                        a.sizes_csv or None)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=5, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-baseline", default="full", choices=["full", "off"])
    ap.add_argument("--learned", type=int, default=1,
                    help="also time the learned-index seeded search (1) or not (0)")
    ap.add_argument("--extra-config", type=int, default=2, choices=[0, 1, 2, 3, 4],
                    help="also measure this configuration and report it under config<N> "
                         "(0: none; only with the default workload)")
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the overlap search shards by read ranges with no exchange; see the "
              "design notes)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pasqual_amd as pq
    res = run_workload(pq, np, args.config, args, args.cpu_baseline == "full")
    extra = None
    if args.extra_config and args.extra_config != args.config and args.config == 5:
        extra = run_workload(pq, np, args.extra_config, args, args.cpu_baseline == "full")

    headline_learned = args.config in LEARNED_HEADLINE and "learned" in res
    value = res["learned"]["reads_per_s"] if headline_learned else res["plain_search_reads_per_s"]
    # the roofline of the headline search's dominant kernel (by time per search)
    if headline_learned:
        # the seeded search's largest single kernel is the seed kernel (the
        # extraction's time is split over its tiers' kernels)
        roofline = res["learned"]["seed_roofline"]
    else:
        roofline = res["plain_roofline"]
    line = {
        "metric": "overlap-search reads/s + SA build Gsuffix/s, learned vs plain index, 1 GPU",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if res["row_bytes"] == 8 else "u32",
        "data": "synthetic (seeded generator, pasqual_amd/csrc/host/pq_synth.c)",
        "config": {"workload": f"config {args.config}: {WORKLOADS[args.config]}",
                   "reads": res["n"], "read_set": res["how"],
                   "suffixes_per_direction": res["N"], "read_len": res["read_len"],
                   "min_overlap": res["t"], "device_rows": f"{8 * res['row_bytes']}-bit",
                   "sa_slices_per_direction": res["slices"],
                   "search": "learned" if headline_learned else "plain",
                   "step": "index build + plain search" +
                           (" + learned-index training + learned search"
                            if "learned" in res else "") +
                           "; each search ends with node records and edge_t records in "
                           "page-locked host memory (copied batch by batch while later "
                           "batches are searched)"},
        "sa_build_gsuffix_per_s": res["sa_build_gsuffix_per_s"],
        "build_ms": res["build_ms"],
        "plain_search_reads_per_s": res["plain_search_reads_per_s"],
        "plain_search_ms": res["plain_search_ms"],
        "plain_search_kernels_ms": res["plain_search_kernels_ms"],
        "plain_search_tail_ms": res["plain_search_tail_ms"],
        "edges": res["edges"],
        "edges_digest": res["edges_digest"],
        "nodes_digest": res["nodes_digest"],
        "overlap_len_sum": res["overlap_len_sum"],
        "matches_reference_record": res["matches_reference_record"],
        "extraction_shape": res["extraction_shape"],
        "find_kernel": res["find_kernel"], "verify_rows": res["verify_rows"],
        "plain_phases_ms": res["plain_phases_ms"],
        "roofline": roofline,
        "plain_roofline": res["plain_roofline"],
        "build_roofline": res["build_roofline"],
        "cpu_baseline": res["cpu_baseline"],
        "learned": res.get("learned"),
        "learned_vs_plain_search": res.get("learned_vs_plain_search"),
        "learned_end_to_end": res.get("learned_end_to_end"),
        "graph_mode": res["graph_mode"],
        "dropin": res["dropin"],
        "import": res["import"],
    }
    base = res["cpu_baseline"]
    if base:
        line["speedup_vs_cpu_search"] = round(value / base["value"], 1)
        line["speedup_vs_cpu_sa_build"] = round(
            res["sa_build_gsuffix_per_s"] / base["sa_build_gsuffix_per_s"], 1)
    if extra:
        line[f"config{args.extra_config}"] = summary(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
