"""Host-side Python binding of the MI355X suffix-array / FM-index overlap library.

The library itself is C ABI (include/pasqual_amd.h, include/pasqual_refabi.h);
this module loads the in-tree build (pasqual_amd/lib/libpasqual_amd_u32.so or
_u64.so) through ctypes and wraps the lower pq_* interface with numpy arrays.
There is no fallback: if the library is missing the import of `lib()` raises.

Mirror of the reference's operator path (string_graph.c / index.c):
    ctx = Context()
    ctx.upload(codes, off, min_overlap)     # reads, codes 1..4
    ctx.build_index()                       # construct_index's device part
    ctx.find_overlaps()                     # construct_string_graph's search
    g = ctx.graph()                         # per-read counts + CSR edges
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

PQ_OK = 0
ERRORS = {-1: "PQ_ERR_ARG", -2: "PQ_ERR_IO", -3: "PQ_ERR_FORMAT", -4: "PQ_ERR_HIP",
          -5: "PQ_ERR_NOMEM", -6: "PQ_ERR_STATE", -7: "PQ_ERR_LIMIT"}
SEARCH_PLAIN, SEARCH_LEARNED = 0, 1
SEARCH_GRAPH = 0x100   # or'ed into a mode: search captured once as a HIP graph, then replayed
BF, FB, FF, BB = 0, 1, 2, 3

# exported symbols per header; the ABI test checks the library against these
PQ_SYMBOLS = [
    "pq_ctx_create", "pq_ctx_destroy", "pq_last_error", "pq_upload_reads", "pq_build_index",
    "pq_get_sa", "pq_get_bwt", "pq_get_reads_index", "pq_get_counts", "pq_get_occ_ref",
    "pq_get_shortcut", "pq_num_reads", "pq_text_len", "pq_find_overlaps", "pq_graph_sizes",
    "pq_graph_copy", "pq_learned_default", "pq_learned_train", "pq_learned_stats",
    "pq_learned_params", "pq_learned_table", "pq_mlp_gradients",
    "pq_get_timing", "pq_autotune", "pq_set_tuning", "pq_graph_copy_edges", "pq_sync",
    "pq_probe_copy", "pq_set_layout", "pq_get_layout", "pq_set_find_chains", "pq_synth_preset", "pq_synth_num_reads",
    "pq_synth_genome",
    "pq_synth_lengths", "pq_synth_reads", "pq_synth_fasta", "pq_import_fasta",
    "pq_dedup_reads", "pq_graph_nodes", "pq_probe_read", "pq_probe_gather", "pq_set_edge_capacity",
    "pq_get_slot_sizes", "pq_graph_export", "pq_host_alloc", "pq_host_free",
    "pq_mlp_eval", "pq_set_verify", "pq_set_seed_verify", "pq_dedup_reads_noalloc", "pq_dev_alloc", "pq_dev_free", "pq_dev_upload", "pq_dev_download",
    "pq_find_overlaps_export", "pq_set_search_batch", "pq_reserve_search", "pq_set_learned_keys",
]
NOALLOC_SCRATCH_BYTES = 32 << 20   # PQ_NOALLOC_SCRATCH_BYTES
REF_SYMBOLS = ["construct_index", "construct_string_graph", "destroy_graph", "free_index",
               "time_importreads"]


class PqError(RuntimeError):
    pass


class SynthCfg(C.Structure):
    _fields_ = [
        ("genome_len", C.c_uint64),
        ("n_units", C.c_uint32), ("unit_len", C.c_uint32), ("n_copies", C.c_uint32),
        ("copy_div", C.c_double),
        ("family_frac", C.c_double),
        ("fam_len_min", C.c_uint32), ("fam_len_max", C.c_uint32),
        ("fam_div_min", C.c_double), ("fam_div_max", C.c_double),
        ("fam_copies_min", C.c_uint32), ("fam_copies_max", C.c_uint32),
        ("len_min", C.c_uint32), ("len_max", C.c_uint32),
        ("coverage", C.c_double),
        ("n_reads", C.c_uint64),
        ("sub_rate", C.c_double),
        ("min_overlap", C.c_uint32),
        ("seed", C.c_uint64),
    ]


class ImportResult(C.Structure):
    _fields_ = [
        ("reads", C.POINTER(C.c_uint8)), ("n", C.c_uint64), ("removed", C.c_uint64),
        ("C", C.c_uint64 * 6), ("suffix_count", C.c_uint64 * 4), ("prefix_count", C.c_uint64 * 4),
        ("ms", C.c_double), ("bad_line", C.c_long), ("bloom_bits", C.c_uint64),
    ]


class LearnedCfg(C.Structure):
    _fields_ = [("k", C.c_uint32), ("hidden", C.c_uint32), ("leaves", C.c_uint32),
                ("segments", C.c_uint32), ("epochs", C.c_uint32), ("batch", C.c_uint32),
                ("samples", C.c_uint32), ("lr", C.c_float), ("closed_form", C.c_uint32),
                ("seed", C.c_uint64), ("sample_stride", C.c_uint64)]


class LearnedStats(C.Structure):
    _fields_ = [("mean_window", C.c_double * 2), ("max_window", C.c_uint64 * 2),
                ("train_loss", C.c_double * 2), ("fit_loss", C.c_double * 2),
                ("train_ms", C.c_double), ("lookups", C.c_uint64 * 2),
                ("probes", C.c_uint64 * 2), ("table_size", C.c_uint64 * 2),
                ("absent", C.c_uint64), ("fallback", C.c_uint64),
                ("depth", C.c_uint32), ("k", C.c_uint32), ("segments", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("group_lanes", C.c_uint32), ("items", C.c_uint32),
                ("nshapes", C.c_uint32), ("extract_ms", C.c_double * 4),
                ("find_chains", C.c_uint32), ("find_ms", C.c_double * 4),
                ("verify_rows", C.c_uint32), ("verify_ms", C.c_double * 3),
                ("seed_rows", C.c_uint32), ("seed_ms", C.c_double * 3)]


class Timing(C.Structure):
    _fields_ = [("upload_ms", C.c_double), ("sa_ms", C.c_double * 2), ("occ_ms", C.c_double * 2),
                ("find_ms", C.c_double), ("extract_ms", C.c_double), ("compact_ms", C.c_double),
                ("learned_ms", C.c_double), ("seed_kernel_ms", C.c_double),
                ("find_launches", C.c_uint64),
                ("find_kernel_ms", C.c_double), ("find_positions", C.c_uint64),
                ("find_lines", C.c_uint64), ("verify_rows", C.c_uint64),
                ("seed_rows", C.c_uint64), ("extract_items", C.c_uint64),
                ("extract_reads", C.c_uint64 * 4)]

    def as_dict(self):
        return {"upload_ms": self.upload_ms, "sa_ms": list(self.sa_ms),
                "occ_ms": list(self.occ_ms), "find_ms": self.find_ms,
                "extract_ms": self.extract_ms, "compact_ms": self.compact_ms,
                "learned_ms": self.learned_ms, "seed_kernel_ms": self.seed_kernel_ms,
                "find_launches": self.find_launches,
                "find_kernel_ms": self.find_kernel_ms, "find_positions": self.find_positions,
                "find_lines": self.find_lines, "verify_rows": self.verify_rows,
                "seed_rows": self.seed_rows, "extract_items": self.extract_items,
                "extract_reads": list(self.extract_reads)}


_LIBS = {}


def lib_path(width="u32"):
    return os.path.join(LIBDIR, f"libpasqual_amd_{width}.so")


def lib(width="u32"):
    """Load the in-tree library (raises if it was not built)."""
    if width in _LIBS:
        return _LIBS[width]
    path = lib_path(width)
    if not os.path.exists(path):
        raise PqError(f"{path} is missing: build it with `python -m pasqual_amd.build`")
    L = C.CDLL(path)
    P, U64, U32, I32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    sig = {
        "pq_ctx_create": (I32, [C.POINTER(P), I32]),
        "pq_ctx_destroy": (None, [P]),
        "pq_last_error": (C.c_char_p, [P]),
        "pq_upload_reads": (I32, [P, P, P, U64, U32]),
        "pq_build_index": (I32, [P]),
        "pq_get_sa": (I32, [P, I32, P]),
        "pq_get_bwt": (I32, [P, I32, P]),
        "pq_get_reads_index": (I32, [P, I32, P]),
        "pq_get_counts": (I32, [P, P, P, P]),
        "pq_get_occ_ref": (I32, [P, I32, P, P, I32]),
        "pq_get_shortcut": (I32, [P, I32, P, P, P]),
        "pq_num_reads": (U64, [P]),
        "pq_text_len": (U64, [P]),
        "pq_find_overlaps": (I32, [P, I32]),
        "pq_graph_sizes": (I32, [P, P, P]),
        "pq_graph_copy": (I32, [P, P, P, P, P, P, P, P, P]),
        "pq_learned_default": (None, [P]),
        "pq_learned_train": (I32, [P, P]),
        "pq_learned_stats": (I32, [P, P]),
        "pq_learned_params": (I32, [P, I32, P, P]),
        "pq_learned_table": (I32, [P, I32, P, P, P, P]),
        "pq_mlp_gradients": (I32, [P, U32, U32, P, P, P, U32, P, P]),
        "pq_get_timing": (I32, [P, P]),
        "pq_autotune": (I32, [P, I32, P]),
        "pq_graph_copy_edges": (I32, [P, P, U32]),
        "pq_set_tuning": (I32, [P, U32]),
        "pq_sync": (I32, [P]),
        "pq_probe_copy": (I32, [P, U64, P]),
        "pq_set_layout": (I32, [P, U32, U64]),
        "pq_set_find_chains": (I32, [P, U32]),
        "pq_set_edge_capacity": (I32, [P, U64]),
        "pq_get_slot_sizes": (I32, [P, I32, P, P, P]),
        "pq_get_layout": (I32, [P, P, P]),
        "pq_synth_preset": (I32, [I32, P]),
        "pq_synth_num_reads": (U64, [P]),
        "pq_synth_genome": (I32, [P, P]),
        "pq_synth_lengths": (I32, [P, P]),
        "pq_synth_reads": (I32, [P, P, P, P]),
        "pq_synth_fasta": (I32, [P, C.c_char_p]),
        "pq_import_fasta": (I32, [P, C.c_char_p, I32, I32, I32, I32, P]),
        "pq_dedup_reads": (I32, [P, P, P, U64, I32, U64, I32, I32, P, P]),
        "pq_graph_nodes": (I32, [P, P, P]),
        "pq_graph_export": (I32, [P, P, P, U32, P]),
        "pq_find_overlaps_export": (I32, [P, I32, P, P, U64, U32, P, P]),
        "pq_set_search_batch": (I32, [P, U64]),
        "pq_reserve_search": (I32, [P, U32]),
        "pq_set_learned_keys": (I32, [P, U32]),
        "pq_host_alloc": (I32, [P, U64, C.POINTER(P)]),
        "pq_host_free": (None, [P]),
        "pq_mlp_eval": (I32, [P, I32, P, U32, I32, P]),
        "pq_set_verify": (I32, [P, U32]),
        "pq_set_seed_verify": (I32, [P, U32]),
        "pq_dedup_reads_noalloc": (I32, [P, P, P, U64, I32, U64, I32, I32, P, P]),
        "pq_dev_alloc": (I32, [P, U64, C.POINTER(P)]),
        "pq_dev_free": (None, [P]),
        "pq_dev_upload": (I32, [P, P, P, U64]),
        "pq_dev_download": (I32, [P, P, P, U64]),
        "pq_probe_read": (I32, [P, U64, P]),
        "pq_probe_gather": (I32, [P, U64, I32, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[width] = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---- host utilities ------------------------------------------------------------------

def synth_preset(config):
    cfg = SynthCfg()
    if lib().pq_synth_preset(config, C.byref(cfg)) != 0:
        raise PqError(f"unknown workload config {config}")
    return cfg


def synth_reads(cfg):
    """(codes uint8[N] in 1..4, off uint64[n+1]) of a synthetic workload."""
    L = lib()
    n = L.pq_synth_num_reads(C.byref(cfg))
    genome = np.empty(max(1, cfg.genome_len), dtype=np.uint8)
    if L.pq_synth_genome(C.byref(cfg), _ptr(genome)) != 0:
        raise PqError("genome synthesis failed")
    off = np.empty(n + 1, dtype=np.uint64)
    if L.pq_synth_lengths(C.byref(cfg), _ptr(off)) != 0:
        raise PqError("read length synthesis failed")
    codes = np.empty(max(1, int(off[-1])), dtype=np.uint8)
    L.pq_synth_reads(C.byref(cfg), _ptr(genome), _ptr(off), _ptr(codes))
    return codes[:int(off[-1])], off


def synth_fasta(cfg, path):
    if lib().pq_synth_fasta(C.byref(cfg), str(path).encode()) != 0:
        raise PqError(f"writing {path} failed")


def import_fasta(path, read_len, bloom_m=36, n_hashfunc=2, index_bits=32, ctx=None):
    """Reads import with duplicate removal (index.c:145-405), the duplicate
    filter on ctx's device (a context of its own when ctx is None); returns a
    dict."""
    if ctx is None:
        with Context() as c:
            return import_fasta(path, read_len, bloom_m, n_hashfunc, index_bits, c)
    res = ImportResult()
    rc = ctx._L.pq_import_fasta(ctx._h, str(path).encode(), read_len, bloom_m, n_hashfunc,
                                index_bits, C.byref(res))
    if rc != 0:
        msg = ctx._L.pq_last_error(ctx._h) if rc not in (-2, -3) else b""
        raise PqError(f"import of {path} failed: {ERRORS.get(rc, rc)} line {res.bad_line}"
                      + (f": {msg.decode()}" if msg else ""))
    n = int(res.n)
    reads = np.ctypeslib.as_array(res.reads, shape=(n * read_len,)).copy() if n else \
        np.zeros(0, np.uint8)
    C.CDLL(None).free(C.cast(res.reads, C.c_void_p))
    return {"reads": reads, "n": n, "removed": int(res.removed), "C": list(res.C),
            "suffix_count": list(res.suffix_count), "prefix_count": list(res.prefix_count),
            "ms": res.ms, "bloom_bits": int(res.bloom_bits)}


# ---- device context ------------------------------------------------------------------

class PinnedBuffer:
    """Page-locked host memory from pq_host_alloc; `.array` is a uint8 view."""

    def __init__(self, ctx, nbytes):
        self._L = ctx._L
        self.ptr = C.c_void_p()
        ctx._check(self._L.pq_host_alloc(ctx._h, max(16, int(nbytes)), C.byref(self.ptr)),
                   "host_alloc")
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array(C.cast(self.ptr, C.POINTER(C.c_uint8)),
                                           shape=(max(16, self.nbytes),))[:self.nbytes]

    def free(self):
        if self.ptr:
            self.array = None
            self._L.pq_host_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device, one stream, one index (not thread-safe)."""

    def __init__(self, device=0, width="u32"):
        self._L = lib(width)
        self._h = C.c_void_p()
        rc = self._L.pq_ctx_create(C.byref(self._h), device)
        if rc != 0:
            raise PqError(f"pq_ctx_create failed: {ERRORS.get(rc, rc)}")

    def close(self):
        if self._h:
            self._L.pq_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.pq_last_error(self._h)
            raise PqError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def dedup_reads(self, codes, ok, read_len, bloom_bits, n_hashfunc=2, index_bits=32):
        """The import's duplicate filter over parsed records (pq_dedup_reads):
        codes = records * read_len codes 1..4, ok = 0 for records holding N.
        Returns (keep flags, passes)."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        ok = np.ascontiguousarray(ok, dtype=np.uint8)
        nrec = len(ok)
        assert codes.size == nrec * read_len
        keep = np.zeros(max(nrec, 1), np.uint8)
        passes = C.c_uint32()
        self._check(self._L.pq_dedup_reads(self._h, _ptr(codes), _ptr(ok), nrec, read_len,
                                           int(bloom_bits), n_hashfunc, index_bits, _ptr(keep),
                                           C.byref(passes)), "dedup_reads")
        return keep[:nrec], int(passes.value)

    @property
    def n(self):
        return int(self._L.pq_num_reads(self._h))

    @property
    def N(self):
        return int(self._L.pq_text_len(self._h))

    def upload(self, codes, off, min_overlap):
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        self._check(self._L.pq_upload_reads(self._h, _ptr(codes), _ptr(off), len(off) - 1,
                                             int(min_overlap)), "upload")

    def build_index(self):
        self._check(self._L.pq_build_index(self._h), "build_index")

    def set_layout(self, row_bytes=0, slice_suffixes=0):
        """Row width 4 or 8 (0: by text size; from the next upload) and the most
        suffixes sorted at once (0: 2^31; from the next build)."""
        self._check(self._L.pq_set_layout(self._h, int(row_bytes), int(slice_suffixes)),
                    "set_layout")

    def layout(self):
        """(row bytes in use, slices of the last build)."""
        rb, sl = C.c_uint32(0), C.c_uint32(0)
        self._check(self._L.pq_get_layout(self._h, C.byref(rb), C.byref(sl)), "get_layout")
        return rb.value, sl.value

    def sa(self, d):
        out = np.empty(max(1, self.N), dtype=np.uint64)
        self._check(self._L.pq_get_sa(self._h, d, _ptr(out)), "get_sa")
        return out[:self.N]

    def bwt(self, d):
        out = np.empty(max(1, self.N), dtype=np.uint8)
        self._check(self._L.pq_get_bwt(self._h, d, _ptr(out)), "get_bwt")
        return out[:self.N]

    def reads_index(self, d):
        out = np.empty(max(1, self.n), dtype=np.uint64)
        self._check(self._L.pq_get_reads_index(self._h, d, _ptr(out)), "get_reads_index")
        return out[:self.n]

    def counts(self):
        Cc = np.zeros(6, np.uint64)
        s = np.zeros(4, np.uint64)
        p = np.zeros(4, np.uint64)
        self._check(self._L.pq_get_counts(self._h, _ptr(Cc), _ptr(s), _ptr(p)), "get_counts")
        return Cc, s, p

    def occ_ref(self, d, sample_bytes=4):
        rows = (self.N + 64) // 64
        bm = np.empty(5 * rows, np.uint64)
        sm = np.empty(5 * rows, np.uint32 if sample_bytes == 4 else np.uint64)
        self._check(self._L.pq_get_occ_ref(self._h, d, _ptr(bm), _ptr(sm), sample_bytes),
                    "get_occ_ref")
        return bm.reshape(5, rows), sm

    def shortcut(self, d):
        n = self.n
        st = np.empty(max(1, n), np.uint64)
        sz = np.empty(max(1, n), np.uint64)
        dp = np.empty(max(1, n), np.uint16)
        self._check(self._L.pq_get_shortcut(self._h, d, _ptr(st), _ptr(sz), _ptr(dp)),
                    "get_shortcut")
        return st[:n], sz[:n], dp[:n]

    def slot_sizes(self, d):
        """index_t's first-triplet slot sizes of direction d: (size[5], offset[5], max)."""
        size = np.zeros(5, np.uint64)
        offs = np.zeros(5, np.uint64)
        mx = C.c_uint64()
        self._check(self._L.pq_get_slot_sizes(self._h, d, _ptr(size), _ptr(offs), C.byref(mx)),
                    "get_slot_sizes")
        return size.tolist(), offs.tolist(), int(mx.value)

    def find_overlaps(self, mode=SEARCH_PLAIN):
        self._check(self._L.pq_find_overlaps(self._h, mode), "find_overlaps")

    def graph(self):
        n, E = C.c_uint64(), C.c_uint64()
        self._check(self._L.pq_graph_sizes(self._h, C.byref(n), C.byref(E)), "graph_sizes")
        n, E = n.value, E.value
        ovB = np.empty(max(1, n), np.int32)
        ovF = np.empty(max(1, n), np.int32)
        nB = np.empty(max(1, n), np.uint32)
        nF = np.empty(max(1, n), np.uint32)
        eo = np.empty(n + 1, np.uint64)
        dst = np.empty(max(1, E), np.uint64)
        ln = np.empty(max(1, E), np.uint16)
        ty = np.empty(max(1, E), np.uint8)
        self._check(self._L.pq_graph_copy(self._h, _ptr(ovB), _ptr(ovF), _ptr(nB), _ptr(nF),
                                          _ptr(eo), _ptr(dst), _ptr(ln), _ptr(ty)), "graph_copy")
        return {"overlaps_B": ovB[:n], "overlaps_F": ovF[:n], "num_B": nB[:n], "num_F": nF[:n],
                "edge_off": eo, "dest": dst[:E], "len": ln[:E], "type": ty[:E]}

    NODE_DTYPE = np.dtype([("num_B", "<u2"), ("num_F", "<u2"), ("B_type", "i1"),
                           ("F_type", "i1"), ("status", "i1"), ("marked", "i1")])

    def graph_nodes(self):
        """Node records of the last search as construct_string_graph sets them
        (pq_graph_nodes) and the sum of the edges' overlap lengths."""
        n, E = C.c_uint64(), C.c_uint64()
        self._check(self._L.pq_graph_sizes(self._h, C.byref(n), C.byref(E)), "graph_sizes")
        rec = np.zeros(max(1, n.value), self.NODE_DTYPE)
        total = C.c_uint64()
        self._check(self._L.pq_graph_nodes(self._h, _ptr(rec), C.byref(total)), "graph_nodes")
        return rec[:n.value], int(total.value)

    def host_buffer(self, nbytes):
        """Page-locked host memory of nbytes (pq_host_alloc) as a uint8 array; freed
        with the array's owner (PinnedBuffer.free or garbage collection)."""
        return PinnedBuffer(self, nbytes)

    def graph_export(self, nodes, edges, index_bytes):
        """Node records and edge_t records of the last search into host buffers
        (pq_graph_export, one host wait): nodes holds n * 8 bytes, edges E *
        (8 if index_bytes == 4 else 16); either may be None.  Returns the sum of
        the overlap lengths."""
        total = C.c_uint64()
        self._check(self._L.pq_graph_export(self._h, None if nodes is None else nodes.ptr,
                                            None if edges is None else edges.ptr, index_bytes,
                                            C.byref(total)), "graph_export")
        return int(total.value)

    def find_overlaps_export(self, nodes, edges, index_bytes, mode=SEARCH_PLAIN):
        """Search that hands over the graph as it goes (pq_find_overlaps_export):
        node records into `nodes` (n * 8 bytes) and edge_t records into `edges`
        while later batches are searched; either may be None.  Returns (edges,
        overlap-length sum).  An edges buffer too small raises PqError with the
        edge count left in graph_size()."""
        E, total = C.c_uint64(), C.c_uint64()
        rec = 8 if index_bytes == 4 else 16
        cap = 0 if edges is None else edges.nbytes // rec
        self._check(self._L.pq_find_overlaps_export(
            self._h, mode, None if nodes is None else nodes.ptr,
            None if edges is None else edges.ptr, cap, index_bytes, C.byref(E),
            C.byref(total)), "find_overlaps_export")
        return int(E.value), int(total.value)

    def reserve_search(self, index_bytes):
        """Device buffers of a later plain export search (pq_reserve_search)."""
        self._check(self._L.pq_reserve_search(self._h, int(index_bytes)), "reserve_search")

    def set_search_batch(self, reads):
        """Reads per search batch (0: automatic)."""
        self._check(self._L.pq_set_search_batch(self._h, int(reads)), "set_search_batch")

    def graph_size(self):
        """(reads, edges) of the last search."""
        n, E = C.c_uint64(), C.c_uint64()
        self._check(self._L.pq_graph_sizes(self._h, C.byref(n), C.byref(E)), "graph_sizes")
        return n.value, E.value

    def autotune(self, mode=SEARCH_PLAIN):
        t = Tune()
        self._check(self._L.pq_autotune(self._h, mode, C.byref(t)), "autotune")
        return {"shape": t.shape, "group_lanes": t.group_lanes, "items": t.items,
                "extract_ms": list(t.extract_ms)[:t.nshapes], "find_chains": t.find_chains,
                "find_ms": list(t.find_ms), "verify_rows": t.verify_rows,
                "verify_ms": list(t.verify_ms), "seed_rows": t.seed_rows,
                "seed_ms": list(t.seed_ms)}

    def set_tuning(self, shape):
        self._check(self._L.pq_set_tuning(self._h, shape), "set_tuning")

    def set_find_chains(self, chains):
        self._check(self._L.pq_set_find_chains(self._h, chains), "set_find_chains")

    def set_verify(self, rows):
        """Chains of at most `rows` rows (<= 32) at depths <= t finish on the text
        (0: backward search to the end)."""
        self._check(self._L.pq_set_verify(self._h, int(rows)), "set_verify")

    def set_seed_verify(self, rows):
        """Seeded search: chains whose depth-D interval holds at most `rows` rows
        (<= 32) are finished by the seed kernel (0: all go to the search kernel)."""
        self._check(self._L.pq_set_seed_verify(self._h, int(rows)), "set_seed_verify")

    def set_edge_capacity(self, edges):
        self._check(self._L.pq_set_edge_capacity(self._h, int(edges)), "set_edge_capacity")

    def learned_defaults(self):
        """The learned index's default training configuration (pq_learned_default)."""
        cfg = LearnedCfg()
        self._L.pq_learned_default(C.byref(cfg))
        return {f[0]: getattr(cfg, f[0]) for f in cfg._fields_}

    def set_learned_keys(self, stride):
        """Later index builds also write the learned index's sampled keys of every
        stride-th row (pq_set_learned_keys; 0 = no)."""
        self._check(self._L.pq_set_learned_keys(self._h, int(stride)), "set_learned_keys")

    def learned_train(self, **kw):
        cfg = LearnedCfg()
        self._L.pq_learned_default(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        self._check(self._L.pq_learned_train(self._h, C.byref(cfg)), "learned_train")

    def learned_stats(self):
        st = LearnedStats()
        self._check(self._L.pq_learned_stats(self._h, C.byref(st)), "learned_stats")
        return {"mean_window": list(st.mean_window), "max_window": list(st.max_window),
                "train_loss": list(st.train_loss), "fit_loss": list(st.fit_loss),
                "train_ms": st.train_ms, "lookups": list(st.lookups), "probes": list(st.probes),
                "table_size": list(st.table_size), "absent": st.absent,
                "fallback": st.fallback, "depth": st.depth, "k": st.k,
                "segments": st.segments}

    def learned_params(self, d):
        cnt = C.c_uint32(0)
        self._check(self._L.pq_learned_params(self._h, d, None, C.byref(cnt)), "learned_params")
        out = np.zeros(cnt.value, np.float32)
        self._check(self._L.pq_learned_params(self._h, d, out.ctypes.data, C.byref(cnt)),
                    "learned_params")
        return out

    def learned_table(self, d):
        """(keys uint64[n, 2] high/low halves, first SA row, row count) of direction d."""
        cnt = C.c_uint64(0)
        self._check(self._L.pq_learned_table(self._h, d, None, None, None, C.byref(cnt)),
                    "learned_table")
        n = cnt.value
        keys = np.zeros((n, 2), np.uint64)
        first = np.zeros(n, np.uint64)
        rows = np.zeros(n, np.uint32)
        self._check(self._L.pq_learned_table(self._h, d, keys.ctypes.data, first.ctypes.data,
                                             rows.ctypes.data, C.byref(cnt)), "learned_table")
        return keys, first, rows

    def mlp_eval(self, d, xs, mfma=True):
        """Outputs of direction d's trained network on prefix values xs (the
        search's matrix-core form, or the scalar form)."""
        xs = np.ascontiguousarray(xs, np.uint64)
        out = np.zeros(max(1, xs.size), np.float32)
        self._check(self._L.pq_mlp_eval(self._h, d, xs.ctypes.data, xs.size, int(bool(mfma)),
                                        out.ctypes.data), "mlp_eval")
        return out[:xs.size]

    def mlp_gradients(self, k, hidden, params, xs, targets):
        params = np.ascontiguousarray(params, np.float32)
        xs = np.ascontiguousarray(xs, np.uint64)
        targets = np.ascontiguousarray(targets, np.float32)
        grad = np.zeros(params.size, np.float32)
        loss = C.c_float(0)
        self._check(self._L.pq_mlp_gradients(self._h, k, hidden, params.ctypes.data,
                                              xs.ctypes.data, targets.ctypes.data, xs.size,
                                              grad.ctypes.data, C.byref(loss)), "mlp_gradients")
        return grad, float(loss.value)

    def timing(self):
        t = Timing()
        self._check(self._L.pq_get_timing(self._h, C.byref(t)), "get_timing")
        return t.as_dict()

    def sync(self):
        self._check(self._L.pq_sync(self._h), "sync")

    def probe_copy(self, nbytes):
        ms = C.c_double()
        self._check(self._L.pq_probe_copy(self._h, nbytes, C.byref(ms)), "probe_copy")
        return ms.value

    def probe_gather(self, nbytes, mode):
        """Milliseconds for 2^25 scattered 64-byte line reads over nbytes (mode:
        0 whole lines per lane, 1 dependent chains, 2 four lanes per line,
        3 40 bytes per line)."""
        ms = C.c_double()
        self._check(self._L.pq_probe_gather(self._h, nbytes, int(mode), C.byref(ms)),
                    "probe_gather")
        return ms.value

    def probe_read(self, nbytes):
        ms = C.c_double()
        self._check(self._L.pq_probe_read(self._h, nbytes, C.byref(ms)), "probe_read")
        return ms.value
