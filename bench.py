"""Benchmark of the hot path on one MI355X: suffix-array + FM-index build of
both directions and the FM-index overlap search, plain and seeded by the
learned index, on a seeded synthetic workload.

Default workload ("repeats and errors"): 5 Mbp genome with 100 interspersed
copies of ten 1 kbp repeat units at 1% divergence, 100 bp reads at 40x with 0.5%
substitutions, t = 63, 32-bit index.  --config 3 / 4: 50 Mbp genome, read
lengths uniform in 36..250 at 30x, 0.3% substitutions, t = 31 (4 = the same
reads with the learned index as the headline search).  --config 5: 200 Mbp
genome with 5% repeat families, 150 bp reads at 30x (40 M reads, 6 G suffixes
per direction: 64-bit device rows, suffix sort in slices), t = 63, learned index
as the headline search.

A step = one index build (SA, BWT occurrence lines, reads_index for the forward
and the backward text) + one overlap search (candidate intervals, irreducible
edges, edge CSR) over the reads, which are resident in HBM before the timed
region.  The learned variant trains its model once on the built index and then
times the same number of seeded searches.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 2]
                    [--cpu-baseline full|off] [--learned 1|0]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LINE_BYTES = 64

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


def cpu_baseline_reference(fasta, L, t, n_expected, exe_name="pasqual_omp_u32"):
    """The reference OpenMP program (oracle/_ref) on the same reads file, timed
    by its own timers (index.c:643-680, string_graph.c:737-1044)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pq_oracle as orc
    exe = os.path.join(orc.REFDIR, exe_name)
    if not os.path.exists(exe):
        return None
    threads = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
    with tempfile.TemporaryDirectory() as wd:
        t0 = time.time()
        r = subprocess.run([exe, "--singlereads", fasta, "-l", str(L), "-t", str(t)], cwd=wd,
                           capture_output=True, text=True, env=env, timeout=900)
        wall = time.time() - t0
        digest = _digest(os.path.join(wd, "contigs.fa"))
    if r.returncode != 0:
        log("reference run failed: " + r.stderr[-500:])
        return None
    out = r.stdout
    takes = [float(x) for x in re.findall(r"takes ([0-9.]+) ms", out)]
    n = int(re.search(r"has imported (\d+) reads", out).group(1))
    # takes: import, forward sort, backward sort, overlap graph, assembly
    fw, bw, graph = takes[1], takes[2], takes[3]
    return {"value": round(n / (graph / 1e3), 1), "unit": "reads/s", "cores": threads,
            "import_ms": takes[0],
            "kind": "reference",
            "sample": f"the full workload ({n} reads), reference {exe_name}, "
                      f"OMP_NUM_THREADS={threads}, one run, overlap-graph timer "
                      "(string_graph.c:737-1044)", "reads": n,
            "sa_build_gsuffix_per_s": round(2 * n * L / ((fw + bw) / 1e3) / 1e9, 5),
            "wall_s": round(wall, 1),
            "reads_match_gpu": None if n_expected is None else n == n_expected,
            "contigs_sha256": digest}


def _digest(path):
    import hashlib
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def dropin_run(fasta, L, t):
    """The reference's own driver, assembler and host files linked against this
    library (oracle/_ref/pasqual_hip_u32, built by `make -C oracle dropin`) on
    the same reads: its 'Building overlap graph' timer covers the device search,
    the copy of the edges and the host node build, like the reference's
    (string_graph.c:737-1044)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pq_oracle as orc
    exe = os.path.join(orc.REFDIR, "pasqual_hip_u32")
    if not os.path.exists(exe):
        return None
    threads = min(16, os.cpu_count() or 1)
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


def cpu_baseline_reference_sample(pq, config, t):
    """The reference OpenMP program (pasqual_omp_u64) on a bounded sample of a
    workload too large for it: the same generator over a 10 Mbp genome."""
    cfg = pq.synth_preset(config)
    cfg.genome_len = 10_000_000
    L = cfg.len_min
    wd = tempfile.mkdtemp(prefix="pqbase_")
    fasta = os.path.join(wd, "sample.fa")
    pq.synth_fasta(cfg, fasta)
    base = cpu_baseline_reference(fasta, L, t, None, exe_name="pasqual_omp_u64")
    os.remove(fasta)
    if base:
        base["sample"] = (f"{base['reads']} reads of the same generator over a 10 Mbp genome "
                          "(same repeat families, coverage, length and errors), reference "
                          f"pasqual_omp_u64, OMP_NUM_THREADS={base['cores']}, one run, "
                          "overlap-graph timer (string_graph.c:737-1044)")
    return base


# reads import of the fixed-length workloads (file read, records cut on the
# host, duplicate filter on the device), filled by load_workload
IMPORT_INFO = {}


def load_workload(pq, config, np):
    cfg = pq.synth_preset(config)
    t = cfg.min_overlap
    if config == 5:
        # 40 M reads: the generator's reads as they are (the 6.8 GB reads file
        # and its import are not part of the measured path)
        codes, off = pq.synth_reads(cfg)
        return codes, off, t, None, None, "as generated (no duplicate removal)"
    if cfg.len_min == cfg.len_max:
        # fixed-length reads go through the reference's import (duplicate removal)
        L = cfg.len_min
        fasta = os.path.join(tempfile.mkdtemp(prefix="pqbench_"), "reads.fa")
        pq.synth_fasta(cfg, fasta)
        imp = pq.import_fasta(fasta, L, index_bits=32)
        n = imp["n"]
        IMPORT_INFO.update(records=n + imp["removed"], kept=n, ms=round(imp["ms"], 1))
        off = np.arange(n + 1, dtype=np.uint64) * np.uint64(L)
        return imp["reads"], off, t, L, fasta, "after duplicate removal"
    codes, off = pq.synth_reads(cfg)
    return codes, off, t, None, None, "as generated (uneven lengths: no duplicate removal)"


def build_model_bytes(lmax, W):
    """(rounds, bytes per suffix) of the prefix-doubling byte model of a suffix-array
    build: R = 1 + ceil(log2(lmax / 29)) rounds from a 64-bit initial key of 29
    two-bit codes and a clamped length, each round one key + index pass (read and
    write, 2 * (8 + W)) and a rank read, gather and write (3 * W), then the BWT
    gather and the occurrence write (W + 2 + (40 + 5 W) / 64); W = index width."""
    rounds = 1 + max(0, math.ceil(math.log2(lmax / 29)))
    return rounds, rounds * (2 * (8 + W) + 3 * W) + (W + 2 + (40 + 5 * W) / 64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-baseline", default="full", choices=["full", "off"])
    ap.add_argument("--learned", type=int, default=1,
                    help="also time the learned-index seeded search (1) or not (0)")
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the overlap search shards by read ranges with no exchange; see the "
              "design notes)", file=sys.stderr)
        sys.exit(2)
    if args.config in LEARNED_HEADLINE:
        args.learned = 1

    import numpy as np
    import pasqual_amd as pq
    t0 = time.time()
    codes, off, t, L, fasta, how = load_workload(pq, args.config, np)
    n = len(off) - 1
    N = int(off[-1])
    log(f"workload config {args.config}: {n} reads {how}, N={N} per direction, "
        f"prepared in {time.time() - t0:.1f}s")

    ctx = pq.Context(0)
    ctx.upload(codes, off, t)
    for _ in range(args.warmup):
        ctx.build_index()
        ctx.find_overlaps()
    if args.warmup == 0:
        ctx.build_index()
    ctx.sync()
    # untimed: pick the extraction shape for this read set (identical results)
    tune = ctx.autotune()
    log(f"warm-up done; extraction shape {tune}")
    build_ms, search_ms, find_launch_ms = [], [], []
    positions = lines = launches = 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        ctx.build_index()
        ctx.sync()
        b = time.perf_counter()
        ctx.find_overlaps()
        ctx.sync()
        c = time.perf_counter()
        tm = ctx.timing()
        build_ms.append((b - a) * 1e3)
        search_ms.append((c - b) * 1e3)
        find_launch_ms.append(tm["find_kernel_ms"])
        launches += tm["find_launches"]
        positions += tm["find_positions"]
        lines += tm["find_lines"]
    ctx.sync()
    total_s = time.perf_counter() - t_start
    g = ctx.graph()
    edges = int(g["edge_off"][-1])
    log("plain steps done")
    # the same search as one HIP graph (captured by the first, untimed call)
    ctx.find_overlaps(pq.SEARCH_GRAPH)
    graph_ms = []
    for _ in range(args.steps):
        a = time.perf_counter()
        ctx.find_overlaps(pq.SEARCH_GRAPH)
        ctx.sync()
        graph_ms.append((time.perf_counter() - a) * 1e3)
    gg = ctx.graph()
    graph = {"plain_search_ms": round(sum(graph_ms) / len(graph_ms), 3),
             "plain_edges_identical": all(np.array_equal(gg[k], g[k]) for k in g)}
    log("graph-mode steps done")

    # learned-index variant on the same index: train once on the built SA,
    # then the same number of timed searches seeded by it
    learned = None
    if args.learned:
        ctx.build_index()
        ctx.learned_train()
        tune_l = ctx.autotune(pq.SEARCH_LEARNED)
        log(f"learned search tuned: {tune_l}")
        gl = ctx.graph()
        same = all(np.array_equal(gl[k], g[k]) for k in g)
        ctx.sync()
        l_ms, l_seed, l_find, l_pos = [], [], [], 0
        for _ in range(args.steps):
            a = time.perf_counter()
            ctx.find_overlaps(pq.SEARCH_LEARNED)
            ctx.sync()
            l_ms.append((time.perf_counter() - a) * 1e3)
            tm = ctx.timing()
            l_seed.append(tm["learned_ms"])
            l_find.append(tm["find_kernel_ms"])
            l_pos += tm["find_positions"]
        st = ctx.learned_stats()
        ctx.find_overlaps(pq.SEARCH_LEARNED | pq.SEARCH_GRAPH)
        lg_ms = []
        for _ in range(args.steps):
            a = time.perf_counter()
            ctx.find_overlaps(pq.SEARCH_LEARNED | pq.SEARCH_GRAPH)
            ctx.sync()
            lg_ms.append((time.perf_counter() - a) * 1e3)
        gl = ctx.graph()
        graph["learned_search_ms"] = round(sum(lg_ms) / len(lg_ms), 3)
        graph["learned_edges_identical"] = all(np.array_equal(gl[k], g[k]) for k in g)
        avg_l = sum(l_ms) / len(l_ms)
        lookups = sum(st["lookups"])
        learned = {"reads_per_s": round(n / (avg_l / 1e3), 1), "search_ms": round(avg_l, 3),
                   "seed_kernel_ms": round(sum(l_seed) / len(l_seed), 3),
                   "find_kernel_ms": round(sum(l_find) / len(l_find), 3),
                   "rank_positions_per_read": round(l_pos / (n * args.steps), 1),
                   "seed_depth": st["depth"], "model_k": st["k"],
                   "table_size": list(st["table_size"]),
                   "mean_window": [round(x, 2) for x in st["mean_window"]],
                   "max_window": list(st["max_window"]),
                   "probes_per_lookup": round(sum(st["probes"]) / max(1, lookups), 2),
                   "absent_chains": st["absent"], "fallback_chains": st["fallback"],
                   "train_ms": round(st["train_ms"], 1),
                   "fit_loss": [float(f"{x:.3g}") for x in st["fit_loss"]],
                   "train_loss": [float(f"{x:.3g}") for x in st["train_loss"]],
                   "edges_identical_to_plain": bool(same),
                   "extraction_shape": {k: tune_l[k] for k in ("group_lanes", "items")},
                   "find_kernel": tune_l["find_chains"]}
        log("learned steps done")
    if args.probe_copy:
        pms = ctx.probe_copy(args.probe_copy << 20)
        log(f"copy probe: {args.probe_copy} MiB in {pms:.3f} ms "
            f"({2 * (args.probe_copy << 20) / pms / 1e6:.1f} GB/s read+write)")
    # measured streaming ceilings of this GPU, reported beside the published HBM
    # peak (untimed region, best of 3): a copy (1 GiB read + 1 GiB written) and a
    # read of 1 GiB (the candidate search only reads)
    copy_gbs = read_gbs = None
    try:
        copy_gbs = round(2 * (1 << 30) / min(ctx.probe_copy(1 << 30) for _ in range(3)) / 1e6, 1)
        read_gbs = round((1 << 30) / min(ctx.probe_read(1 << 30) for _ in range(3)) / 1e6, 1)
    except pq.PqError as e:
        log(f"streaming ceilings not measured: {e}")
    # scattered 64-byte line reads over a buffer the size of both indexes'
    # occurrence lines (0.5 B per row and direction): independent positions, and
    # chains of dependent loads as a backward search makes them
    # (GB/s of lines, best of 3): whole lines per lane (positions independent, and
    # in dependent chains) and four lanes per line, 16 bytes each
    gather = None
    try:
        occ_bytes = 2 * ((N + 127) // 128) * LINE_BYTES
        nbytes = (1 << 25) * LINE_BYTES
        gather = {name: round(nbytes / min(ctx.probe_gather(occ_bytes, mode) for _ in range(3))
                              / 1e6, 1)
                  for name, mode in (("lane", 0), ("lane_chained", 1), ("quad", 2))}
        gather["footprint_bytes"] = occ_bytes
        log(f"scattered line reads over {occ_bytes / 1e9:.2f} GB: {gather}")
    except pq.PqError as e:
        log(f"scattered-line ceilings not measured: {e}")
    row_bytes, slices = ctx.layout()
    lens = np.diff(off)
    read_len = int(lens[0]) if int(lens.min()) == int(lens.max()) else \
        f"{int(lens.min())}..{int(lens.max())}"
    ms_per_step = total_s * 1e3 / args.steps
    avg_search = sum(search_ms) / len(search_ms)
    avg_build = sum(build_ms) / len(build_ms)
    reads_per_s = n / (avg_search / 1e3)
    gsuffix_per_s = 2 * N / (avg_build / 1e3) / 1e9
    # the suffix-array build against the HBM peak by a prefix-doubling byte
    # model: R = 1 + ceil(log2(Lmax / 29)) rounds from a 64-bit key of 29 codes,
    # each one key + index pass (read and write) plus a rank read, gather and
    # write, then the BWT gather and the occurrence write; W = the index width
    # (4 bytes for 32-bit indexes, 8 for the 64-bit index of the larger
    # workloads, whatever width the device rows have)
    W = 8 if args.config >= 3 else 4
    rounds, b_suffix = build_model_bytes(int(np.diff(off).max()), W)
    build_gbs = gsuffix_per_s * b_suffix
    build_roofline = {"bound": "hbm", "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                      "model_bytes_per_suffix": round(b_suffix, 2), "rounds": rounds,
                      "index_width": W}

    # dominant kernel: the candidate search (bidirectional backward search)
    avg_launch_ms = sum(find_launch_ms) / max(1, launches)
    per_launch_positions = positions / max(1, launches)
    alg_bytes = per_launch_positions * LINE_BYTES
    achieved = alg_bytes / (avg_launch_ms / 1e3) / 1e9
    # HBM/fabric bytes per k_find launch from the committed counter runs of this
    # same workload: FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for
    # the plain copy kernel
    # (the autotuner's choice of search kernel)
    kname = {1: "k_find1", 2: "k_find2", 3: "k_find1q"}[tune["find_chains"]]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"config{args.config}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            kernels = json.load(f)["kernels"]
        tk = next((v for k, v in kernels.items() if k == kname or k.startswith(kname + "<")),
                  None)
        if tk:
            traffic = tk["traffic_bytes"][0]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "alg_bytes_per_launch": int(alg_bytes),
                "rank_positions_per_read": round(positions / (n * args.steps), 1),
                "lines_per_read": round(lines / (n * args.steps), 1),
                "avg_launch_ms": round(avg_launch_ms, 3),
                "launches_per_search": round(launches / args.steps, 2)}
    roofline["copy_ceiling"] = copy_gbs
    roofline["read_ceiling"] = read_gbs
    roofline["frac_of_read_ceiling"] = round(achieved / read_gbs, 4) if read_gbs else None
    # the search's distinct line loads (one or two per step) against the
    # scattered-line ceilings of the same footprint
    line_gbs = lines * LINE_BYTES / (sum(find_launch_ms) / 1e3) / 1e9 if sum(find_launch_ms) else None
    roofline["line_gbs"] = round(line_gbs, 1) if line_gbs else None
    roofline["scattered_line_ceilings"] = gather
    best = max(gather["lane"], gather["quad"]) if gather else None
    roofline["lines_frac_of_scattered_ceiling"] = round(line_gbs / best, 4) \
        if line_gbs and best else None

    base = dropin = None
    if args.cpu_baseline == "full":
        log("timing the CPU baseline ...")
        if L is not None:
            base = cpu_baseline_reference(fasta, L, t, n)
            log("timing the reference driver on this library ...")
            dropin = dropin_run(fasta, L, t)
            if dropin and base:
                dropin["contigs_identical_to_reference"] = (
                    dropin["contigs_sha256"] is not None and
                    dropin["contigs_sha256"] == base["contigs_sha256"])
        elif args.config == 5:
            base = cpu_baseline_reference_sample(pq, args.config, t)
        else:
            base = cpu_baseline_port(pq, args.config, t)
    headline_learned = args.config in LEARNED_HEADLINE
    value = learned["reads_per_s"] if headline_learned else reads_per_s
    line = {
        "metric": "overlap-search reads/s + SA build Gsuffix/s, learned vs plain index, 1 GPU",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if row_bytes == 8 else "u32",
        "data": "synthetic (seeded generator, pasqual_amd/csrc/host/pq_synth.c)",
        "config": {"workload": f"config {args.config}: {WORKLOADS[args.config]}",
                   "reads": n, "suffixes_per_direction": N,
                   "read_len": read_len, "min_overlap": t,
                   "device_rows": f"{8 * row_bytes}-bit", "sa_slices_per_direction": slices,
                   "search": "learned" if headline_learned else "plain"},
        "plain_search_reads_per_s": round(reads_per_s, 1),
        "sa_build_gsuffix_per_s": round(gsuffix_per_s, 3),
        "search_ms": round(avg_search, 3),
        "build_ms": round(avg_build, 3),
        "edges": edges,
        "extraction_shape": {k: tune[k] for k in ("group_lanes", "items")},
        "find_kernel": tune["find_chains"],
        "roofline": roofline,
        "build_roofline": build_roofline,
        "cpu_baseline": base,
        "learned": learned,
        "graph_mode": graph,
        "dropin": dropin,
        "import": dict(IMPORT_INFO) or None,
    }
    if learned:
        line["learned_vs_plain_search"] = round(learned["reads_per_s"] / reads_per_s, 3)
    if base:
        base["host"] = host_cpu()
    if base and IMPORT_INFO and base.get("import_ms"):
        line["import"]["reference_ms"] = base["import_ms"]
    if base:
        line["speedup_vs_cpu_search"] = round(value / base["value"], 1)
        line["speedup_vs_cpu_sa_build"] = round(gsuffix_per_s / base["sa_build_gsuffix_per_s"], 1)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
