"""Python binding of the CPU oracle (test infrastructure only).

Loads oracle/_build/liborc.so (built by `make -C oracle oracle`) and runs the
reference programs of oracle/_ref/ when they are present.  Only tests/, the
package's smoke check and the CPU-baseline leg of the benchmark use this file;
the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "liborc.so")
REFDIR = os.path.join(HERE, "_ref")


class OrcGraph(C.Structure):
    _fields_ = [("n", C.c_uint64), ("overlaps_B", C.POINTER(C.c_int64)),
                ("overlaps_F", C.POINTER(C.c_int64)), ("num_B", C.POINTER(C.c_uint64)),
                ("num_F", C.POINTER(C.c_uint64)), ("edge_off", C.POINTER(C.c_uint64)),
                ("dest", C.POINTER(C.c_uint64)), ("len", C.POINTER(C.c_uint32)),
                ("type", C.POINTER(C.c_uint8)), ("rank_evals", C.c_uint64),
                ("rank_evals_find", C.c_uint64)]


class OrcFm(C.Structure):
    _fields_ = [("n", C.c_uint64), ("N", C.c_uint64), ("L", C.c_uint32), ("t", C.c_uint32),
                ("reads", C.c_void_p), ("C", C.c_uint64 * 6),
                ("bitmap", C.c_void_p * 2), ("sample", C.c_void_p * 2),
                ("reads_index", C.c_void_p * 2), ("istart", C.c_void_p * 2),
                ("iend", C.c_void_p * 2), ("idepth", C.c_void_p * 2)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            subprocess.run(["make", "-C", HERE, "oracle"], check=True, capture_output=True)
        _lib = C.CDLL(LIBPATH)
        P = C.c_void_p
        _lib.orc_dir_text.argtypes = [P, P, C.c_uint64, C.c_int, P]
        _lib.orc_sa.argtypes = [P, P, C.c_uint64, P]
        _lib.orc_bwt.argtypes = [P, P, C.c_uint64, P, P, P]
        _lib.orc_counts.argtypes = [P, P, C.c_uint64, P, P, P]
        _lib.orc_occ_ref.argtypes = [P, C.c_uint64, P, P, P]
        _lib.orc_shortcut.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, P, P, P, P, P]
        _lib.orc_search_fm.argtypes = [C.POINTER(OrcFm), C.POINTER(OrcGraph)]
        _lib.orc_search_explicit.argtypes = [P, P, C.c_uint64, C.c_uint32, C.POINTER(OrcGraph)]
        _lib.orc_graph_free.argtypes = [C.POINTER(OrcGraph)]
        _lib.orc_check_sa.argtypes = [P, P, C.c_uint64, P]
        _lib.orc_check_sa.restype = C.c_int64
        _lib.orc_mlp_loss_grad.argtypes = [C.c_uint32, C.c_uint32, P, P, P, C.c_uint64, P]
        _lib.orc_mlp_loss_grad.restype = C.c_double
        _lib.orc_import_fasta.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_long)]
        _lib.orc_dedup.argtypes = [P, P, C.c_uint64, C.c_int, C.c_uint64, C.c_int, C.c_int, P]
        _lib.orc_slot_sizes.argtypes = [P, P, C.c_uint64, P, P, P]
        _lib.orc_chain_work.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, P]
        _lib.orc_mlp_forward.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_uint64, P]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def dir_text(codes, off, d):
    out = np.empty_like(codes)
    lib().orc_dir_text(_p(codes), _p(off), len(off) - 1, d, _p(out))
    return out


def suffix_array(dtext, off):
    sa = np.empty(max(1, int(off[-1])), np.uint64)
    if lib().orc_sa(_p(dtext), _p(off), len(off) - 1, _p(sa)) != 0:
        raise MemoryError("oracle suffix sort")
    return sa[:int(off[-1])]


def check_sa(dtext, off, sa):
    """-1 if sa is the canonical suffix array of dtext, else the first bad index."""
    sa = np.ascontiguousarray(sa, np.uint64)
    return int(lib().orc_check_sa(_p(dtext), _p(off), len(off) - 1, _p(sa)))


def bwt(dtext, off, sa):
    N, n = int(off[-1]), len(off) - 1
    b = np.empty(max(1, N), np.uint8)
    ri = np.empty(max(1, n), np.uint64)
    lib().orc_bwt(_p(dtext), _p(off), n, _p(sa), _p(b), _p(ri))
    return b[:N], ri[:n]


def import_fasta(path, L, bloom_m=36, n_hashfunc=2, index_bits=32):
    """Sequential reads import (orc_import_fasta): dict like the library's, or
    {"rc": code, "bad_line": line} when it fails."""
    rp, n, rem, bad = C.c_void_p(), C.c_uint64(), C.c_uint64(), C.c_long()
    rc = lib().orc_import_fasta(str(path).encode(), L, bloom_m, n_hashfunc, index_bits,
                                C.byref(rp), C.byref(n), C.byref(rem), C.byref(bad))
    if rc != 0:
        return {"rc": rc, "bad_line": bad.value}
    k = int(n.value)
    reads = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint8)), shape=(max(k * L, 1),))
    reads = reads[:k * L].copy()
    C.CDLL(None).free(rp)
    off = np.arange(k + 1, dtype=np.uint64) * np.uint64(L)
    Cc, sfx, pfx = counts(reads, off) if k else (np.zeros(6, np.uint64),) + (np.zeros(4, np.uint64),) * 2
    return {"rc": 0, "reads": reads, "n": k, "removed": int(rem.value), "C": Cc.tolist(),
            "suffix_count": sfx.tolist(), "prefix_count": pfx.tolist()}


def slot_sizes(dtext, off):
    """(size[5], offset[5], max_size) of a direction text (orc_slot_sizes)."""
    size = np.zeros(5, np.uint64)
    offs = np.zeros(5, np.uint64)
    mx = np.zeros(1, np.uint64)
    lib().orc_slot_sizes(_p(np.ascontiguousarray(dtext, np.uint8)),
                         _p(np.ascontiguousarray(off, np.uint64)), len(off) - 1, _p(size),
                         _p(offs), _p(mx))
    return size.tolist(), offs.tolist(), int(mx[0])


def dedup(codes, ok, L, bloom_bits, n_hashfunc=2, index_bits=32):
    """bloom_filter applied in record order to parsed records (orc_dedup)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    ok = np.ascontiguousarray(ok, dtype=np.uint8)
    keep = np.zeros(max(len(ok), 1), np.uint8)
    lib().orc_dedup(_p(codes), _p(ok), len(ok), L, int(bloom_bits), n_hashfunc, index_bits,
                    _p(keep))
    return keep[:len(ok)]


def counts(codes, off):
    Cc = np.zeros(6, np.uint64)
    s = np.zeros(4, np.uint64)
    p = np.zeros(4, np.uint64)
    lib().orc_counts(_p(codes), _p(off), len(off) - 1, _p(Cc), _p(s), _p(p))
    return Cc, s, p


def occ_ref(b, fix4):
    N = len(b)
    rows = (N + 64) // 64
    bm = np.zeros(5 * rows, np.uint64)
    sm = np.zeros(5 * rows, np.uint64)
    fix = np.ascontiguousarray(fix4, np.uint64)
    lib().orc_occ_ref(_p(b), N, _p(fix), _p(bm), _p(sm))
    return bm.reshape(5, rows), sm


def shortcut(dtext, n, L, t, sa, Cc):
    st = np.zeros(max(1, n), np.uint64)
    en = np.zeros(max(1, n), np.uint64)
    dp = np.zeros(max(1, n), np.uint64)
    Cc = np.ascontiguousarray(Cc, np.uint64)
    if lib().orc_shortcut(_p(dtext), n, L, t, _p(sa), _p(Cc), _p(st), _p(en), _p(dp)) != 0:
        raise MemoryError("oracle shortcut")
    return st[:n], en[:n], dp[:n]


def _graph_to_dict(g):
    n = g.n
    E = int(g.edge_off[n]) if n else 0

    def arr(ptr, k, dt):
        return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt).copy() if k else np.zeros(0, dt)
    out = {"overlaps_B": arr(g.overlaps_B, n, np.int64), "overlaps_F": arr(g.overlaps_F, n, np.int64),
           "num_B": arr(g.num_B, n, np.uint64), "num_F": arr(g.num_F, n, np.uint64),
           "edge_off": arr(g.edge_off, n + 1, np.uint64), "dest": arr(g.dest, E, np.uint64),
           "len": arr(g.len, E, np.uint32), "type": arr(g.type, E, np.uint8),
           "rank_evals": int(g.rank_evals), "rank_evals_find": int(g.rank_evals_find)}
    lib().orc_graph_free(C.byref(g))
    return out


def build_index(codes, off, t, with_shortcut=True):
    """Both directions of the index, in reference layout, for reads codes/off."""
    n = len(off) - 1
    lens = np.diff(off.astype(np.int64))
    fixed = n > 0 and lens.min() == lens.max()
    Cc, sfx, pfx = counts(codes, off)
    out = {"C": Cc, "suffix_count": sfx, "prefix_count": pfx, "n": n, "N": int(off[-1]),
           "codes": codes, "off": off, "t": t, "L": int(lens[0]) if fixed else 0}
    for d in (0, 1):
        dt = dir_text(codes, off, d)
        sa = suffix_array(dt, off)
        b, ri = bwt(dt, off, sa)
        bm, sm = occ_ref(b, sfx if d == 0 else pfx)
        out[d] = {"text": dt, "sa": sa, "bwt": b, "reads_index": ri, "bitmap": bm, "sample": sm}
        if with_shortcut and fixed:
            out[d]["shortcut"] = shortcut(dt, n, int(lens[0]), t, sa, Cc)
    return out


def search_fm(ix, use_shortcut=True):
    """string_graph.c search restated on the reference-layout index (fixed L)."""
    f = OrcFm()
    f.n, f.N, f.L, f.t = ix["n"], ix["N"], ix["L"], ix["t"]
    keep = []

    def p(a):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data
    f.reads = p(ix["codes"])
    for c in range(6):
        f.C[c] = int(ix["C"][c])
    for d in (0, 1):
        f.bitmap[d] = p(ix[d]["bitmap"].astype(np.uint64))
        f.sample[d] = p(ix[d]["sample"].astype(np.uint64))
        f.reads_index[d] = p(ix[d]["reads_index"].astype(np.uint64))
        if use_shortcut and "shortcut" in ix[d]:
            st, en, dp = ix[d]["shortcut"]
        else:
            last = ix["codes"][(np.arange(ix["n"]) * ix["L"] + (ix["L"] - 1 if d == 0 else 0))]
            Cc = np.asarray(ix["C"], np.uint64)
            st, en, dp = Cc[last], Cc[last + 1] - Cc[last], np.ones(ix["n"], np.uint64)
        f.istart[d] = p(np.asarray(st, np.uint64))
        f.iend[d] = p(np.asarray(en, np.uint64))
        f.idepth[d] = p(np.asarray(dp, np.uint64))
    g = OrcGraph()
    if lib().orc_search_fm(C.byref(f), C.byref(g)) != 0:
        lib().orc_graph_free(C.byref(g))
        raise RuntimeError("oracle FM search failed (edge limit or memory)")
    return _graph_to_dict(g)


def search_explicit(codes, off, t):
    g = OrcGraph()
    codes = np.ascontiguousarray(codes, np.uint8)
    off = np.ascontiguousarray(off, np.uint64)
    if lib().orc_search_explicit(_p(codes), _p(off), len(off) - 1, t, C.byref(g)) != 0:
        lib().orc_graph_free(C.byref(g))
        raise RuntimeError("oracle explicit search failed")
    return _graph_to_dict(g)


def chain_work(codes, off, seed_depth=0, t=0, vmax=0, stride=0):
    """(rank positions, occurrence lines, verified rows) of the device's
    candidate-search chains over all reads (orc_chain_work); seed_depth > 0:
    chains of reads longer than it start at that depth (the learned search's
    seeds); vmax > 0: chains of at most vmax rows at depths <= t are verified on
    the text; stride > 0: seeded intervals are scanned as the learned seed
    kernel scans them over suffix-array rows sampled every stride rows."""
    codes = np.ascontiguousarray(codes, np.uint8)
    off = np.ascontiguousarray(off, np.uint64)
    out = np.zeros(3, np.uint64)
    if lib().orc_chain_work(_p(codes), _p(off), len(off) - 1, int(seed_depth), int(t), int(vmax),
                            int(stride), _p(out)) != 0:
        raise RuntimeError("oracle chain work failed")
    return int(out[0]), int(out[1]), int(out[2])


# ---- reference programs (oracle/_ref) ------------------------------------------------

def ref_available(width="u32"):
    return os.path.exists(os.path.join(REFDIR, f"ref_dump_{width}"))


def run_ref_dump(fasta, L, t, out, width="u32", graph=True, threads=1, bloom_m=36,
                 n_hashfunc=2):
    exe = os.path.join(REFDIR, f"ref_dump_{width}")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    args = [exe, str(fasta), str(L), str(t), str(out)] + ([] if graph else ["nograph"])
    args += [f"bloom={bloom_m}", f"nf={n_hashfunc}"]
    r = subprocess.run(args, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"ref_dump failed: {r.stderr}")
    return r.stdout


def read_ref_dump(path):
    raw = np.fromfile(path, dtype=np.uint64)
    pos = 0

    def take(k):
        nonlocal pos
        a = raw[pos:pos + k]
        pos += k
        return a
    magic, W, n, L, t, N = [int(x) for x in take(6)]
    assert magic == 0x31464552514150, "not a reference dump"
    n, N = int(n), int(N)
    nb = (N + 7) // 8
    reads = raw[pos:pos + nb].view(np.uint8)[:N].copy()
    pos += nb
    out = {"W": W, "n": n, "L": L, "t": t, "N": N, "reads": reads}
    out["C"] = take(6).copy()
    out["suffix_count"] = take(4).copy()
    out["prefix_count"] = take(4).copy()
    rows = (N + 64) // 64
    for d in (0, 1):
        dd = {}
        dd["bitmap"] = take(5 * rows).reshape(5, rows).copy()
        dd["sample"] = take(5 * rows).copy()
        dd["reads_index"] = take(n).copy()
        dd["istart"] = take(n).copy()
        dd["iend"] = take(n).copy()
        dd["idepth"] = take(n).copy()
        out[d] = dd
    has_graph = int(take(1)[0])
    if has_graph:
        out["average_overlap"] = take(1).view(np.float64)[0]
        nodes = []
        for _ in range(n):
            hdr = take(9)
            nb_, nf_ = int(hdr[0]), int(hdr[1])
            e = take(3 * (nb_ + nf_)).reshape(-1, 3)
            nodes.append({"num_B": nb_, "num_F": nf_, "max_B": int(hdr[2]), "max_F": int(hdr[3]),
                          "B_type": int(hdr[4]), "F_type": int(hdr[5]), "status": int(hdr[6]),
                          "marked": int(hdr[7]), "has_path_q": int(hdr[8]),
                          "edges": e.copy()})
        out["nodes"] = nodes
    return out


def run_reference(fasta, L, t, workdir, width="u32", threads=None, extra=()):
    """Run the reference CLI in workdir; returns (stdout, contigs.fa bytes)."""
    exe = os.path.join(REFDIR, f"pasqual_omp_{width}")
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run([exe, "--singlereads", str(fasta), "-l", str(L), "-t", str(t), *extra],
                       cwd=workdir, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed: {r.stderr}")
    with open(os.path.join(workdir, "contigs.fa"), "rb") as f:
        return r.stdout, f.read()


def mlp_forward(k, h, params, xs):
    """Outputs of the learned index's MLP in double (orc_mlp_forward)."""
    params = np.ascontiguousarray(params, np.float64)
    xs = np.ascontiguousarray(xs, np.uint64)
    out = np.zeros(max(1, xs.size), np.float64)
    lib().orc_mlp_forward(k, h, _p(params), _p(xs), xs.size, _p(out))
    return out[:xs.size]


def mlp_loss_grad(k, h, params, xs, targets, want_grad=True):
    """Loss and gradient of the learned-index MLP in double (orc_mlp_loss_grad)."""
    params = np.ascontiguousarray(params, np.float64)
    xs = np.ascontiguousarray(xs, np.uint64)
    targets = np.ascontiguousarray(targets, np.float64)
    grad = np.zeros(params.size, np.float64) if want_grad else None
    loss = lib().orc_mlp_loss_grad(k, h, _p(params), _p(xs), _p(targets), xs.size,
                                   _p(grad) if want_grad else None)
    return loss, grad
