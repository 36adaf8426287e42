/*
 * pasqual_amd.h -- C ABI of the MI355X (gfx950) suffix-array / FM-index
 * overlap library.
 *
 * Two layers are exported by libpasqual_amd_u32.so / libpasqual_amd_u64.so:
 *
 *  1. The reference's own entry points, with the reference's struct layouts
 *     for the INDEX_TYPE_T of the library variant (declared in
 *     pasqual_refabi.h): construct_index, construct_string_graph,
 *     destroy_graph, free_index and the global time_importreads.  These are
 *     drop-in replacements for index.o + sufsort.o + string_graph.o of the
 *     reference build.
 *
 *  2. The lower interface below: status-returning calls on an opaque context.
 *     Every call returns PQ_OK (0) or a negative PQ_ERR_* code and never exits;
 *     pq_last_error() describes the last failure.  Plain pointers and sizes
 *     only.  A context owns one HIP stream and all device memory; it is not
 *     re-entrant (one caller thread at a time).
 *
 * Codes: reads are arrays of nucleotide codes 1..4 (A,C,G,T), the encoding of
 * index.c:280-321.  Read r occupies [off[r], off[r+1]) of the concatenated
 * array; all reads may have the same length (the reference's contract) or
 * different lengths (the generalisation used for batches of uneven reads).
 */
#ifndef PASQUAL_AMD_H
#define PASQUAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQ_OK            0
#define PQ_ERR_ARG      -1   /* invalid argument */
#define PQ_ERR_IO       -2   /* file could not be opened / read */
#define PQ_ERR_FORMAT   -3   /* malformed reads file (index.c:218-321 rules) */
#define PQ_ERR_HIP      -4   /* HIP runtime error or no device */
#define PQ_ERR_NOMEM    -5   /* host or device allocation failed */
#define PQ_ERR_STATE    -6   /* call out of order (e.g. search before build) */
#define PQ_ERR_LIMIT    -7   /* a node needs more than 65535 edges
                                (string_graph.c:469-472), or a size limit */

#define PQ_DIR_FW 0          /* index over the reads */
#define PQ_DIR_BW 1          /* index over the reversed reads */

#define PQ_SEARCH_PLAIN   0  /* bidirectional backward search from depth 1 */
#define PQ_SEARCH_LEARNED 1  /* search seeded by the learned index */
#define PQ_SEARCH_GRAPH 0x100 /* or'ed into the mode: run the search as one HIP graph,
                                 captured on first use and replayed while the uploaded
                                 reads, index buffers, learned model and tuning stay
                                 the same (same result; no per-kernel timings) */

typedef struct pq_ctx pq_ctx_t;

/* ---- context ---------------------------------------------------------------- */
int pq_ctx_create(pq_ctx_t **out, int device);
void pq_ctx_destroy(pq_ctx_t *ctx);
const char *pq_last_error(const pq_ctx_t *ctx);

/* ---- input -------------------------------------------------------------------- */
/* Copy reads to the device.  codes: off[n] bytes of 1..4; off: n+1 offsets.
 * min_overlap: t of the overlap search (t >= 1, t < every read length that is
 * expected to overlap).  Replaces any previous reads and index. */
int pq_upload_reads(pq_ctx_t *ctx, const uint8_t *codes, const uint64_t *off,
                    uint64_t n, uint32_t min_overlap);

/* Device index layout, the runtime counterpart of the reference's INDEX_TYPE_T
 * build choice (U32 / U64, common.h): row_bytes 4 or 8 is the width of suffix
 * positions and BWT rows on the device, 0 (default) takes 4 below 2^32 text
 * codes and 8 from there on; applies from the next pq_upload_reads.
 * slice_suffixes bounds the suffixes sorted at once (0: 2^31, also the
 * maximum); a larger text is sorted in key-range slices, which bounds the
 * build's scratch memory; applies from the next pq_build_index.  Results do
 * not depend on either. */
int pq_set_layout(pq_ctx_t *ctx, uint32_t row_bytes, uint64_t slice_suffixes);
/* Row width in use and the number of slices of the last build. */
int pq_get_layout(pq_ctx_t *ctx, uint32_t *row_bytes, uint32_t *slices);

/* ---- index build (index.c:408-706 + sufsort.c) ------------------------------------ */
/* Builds both directions: suffix arrays, BWT occurrence lines, reads_index,
 * C and the end-of-read counts. */
int pq_build_index(pq_ctx_t *ctx);

int pq_get_sa(pq_ctx_t *ctx, int dir, uint64_t *sa);              /* N entries */
int pq_get_bwt(pq_ctx_t *ctx, int dir, uint8_t *bwt);             /* N codes 0..4, 0='$' */
int pq_get_reads_index(pq_ctx_t *ctx, int dir, uint64_t *ri);     /* n entries */
int pq_get_counts(pq_ctx_t *ctx, uint64_t C[6], uint64_t suffix_count[4],
                  uint64_t prefix_count[4]);
/* Occurrence table in the reference host layout (index.c:26-142, 587-615):
 * rows = (N+64)/64; bitmap gets 5*rows words (row c at c*rows), sample gets
 * 5*rows entries of sample_bytes (4 or 8) each. */
int pq_get_occ_ref(pq_ctx_t *ctx, int dir, uint64_t *bitmap, void *sample,
                   int sample_bytes);
/* Interval shortcut of every read for the reference's search entry
 * (index.h:31-37): start, size and depth of the SA interval the reference's
 * sort records for the read's last P codes (fw; first P codes reversed, bw):
 * the deepest P in 9, 12, ... <= min(t, L) whose bucket the sort visits
 * (sufsort.c:177-209, 322-334, 354-376, 594-611), else depth 1 (the interval of
 * the last / first code, index.c:495-502, 548-555).  The deeper records need
 * reads of one length (the reference's only input) and a text of fewer than
 * 2^32 codes; otherwise every read gets the depth-1 interval. */
int pq_get_shortcut(pq_ctx_t *ctx, int dir, uint64_t *start, uint64_t *size,
                    uint16_t *depth);
/* First-triplet slot sizes of index_t (size_fw / offset_fw / max_size_fw and the
 * bw ones, index.h:52-57), as initial_sort computes them (sufsort.c:386-450):
 * size[k] = text positions whose code of the (up to) three codes from there to
 * the end of the read, 1 + 21 c0 + [rem >= 2] (1 + 5 c1 + [rem >= 3] (1 + c2)),
 * lies in 17k .. 17k+16; offset = exclusive prefix; *max_size = largest size. */
int pq_get_slot_sizes(pq_ctx_t *ctx, int dir, uint64_t size[5], uint64_t offset[5],
                      uint64_t *max_size);
uint64_t pq_num_reads(const pq_ctx_t *ctx);
uint64_t pq_text_len(const pq_ctx_t *ctx);

/* ---- overlap search (string_graph.c:162-731) ------------------------------------------- */
/* Any number of overlap candidates per read side: sides beyond the largest
 * extraction tier (65536 candidates) are sorted by the whole device; a graph
 * replay that meets one runs the search again eagerly.  PQ_ERR_NOMEM if their
 * scratch (about 90 bytes per candidate) does not fit. */
int pq_find_overlaps(pq_ctx_t *ctx, int mode);
/* Sizes of the last search result. */
int pq_graph_sizes(pq_ctx_t *ctx, uint64_t *n_reads, uint64_t *n_edges);
/* Copy the last search result: per read the candidate counts overlaps_B/F
 * (string_graph.c:820), the number of irreducible B and F edges, and the edges
 * in CSR order (read r: [edge_off[r], edge_off[r+1]), B list then F list),
 * as (dest read, overlap length, type BF=0 FB=1 FF=2 BB=3).  Any output
 * pointer may be NULL. */
int pq_graph_copy(pq_ctx_t *ctx, int32_t *overlaps_B, int32_t *overlaps_F,
                  uint32_t *num_B, uint32_t *num_F, uint64_t *edge_off,
                  uint64_t *dest, uint16_t *len, uint8_t *type);

/* The edges of the last search in the reference's edge_t layout (string_graph.h:
 * INDEX_TYPE_T dest, unsigned short overlap_len, char type; 8 bytes per edge for
 * index_bytes 4, 16 for 8), in edge_off order: what construct_string_graph copies
 * into the node records. */
int pq_graph_copy_edges(pq_ctx_t *ctx, void *edges, uint32_t index_bytes);
/* Bytes of one edge_t record: 8 for index_bytes 4, 16 for 8. */
#define PQ_EDGE_BYTES(index_bytes) ((index_bytes) == 4 ? 8u : 16u)

/* Node record of one read as construct_string_graph sets it (string_graph.c:
 * 832-889), with the reference's values (string_graph.h; PQ_REF_* of
 * pasqual_refabi.h): B_type / F_type ONE 1, BRANCH 2, ENDING 4; status DEFAULT 0
 * or REMOVED 8 (no edge on either side); marked SUSPECT 4 when
 * |overlaps_B - overlaps_F| / max(overlaps_B, overlaps_F) > 0.6, else 0. */
typedef struct {
    uint16_t num_B, num_F;   /* irreducible edges per side (max_B/max_F start equal) */
    int8_t B_type, F_type;
    int8_t status, marked;
} pq_node_t;

/* Node records of the last search, computed on the device (8 bytes per read
 * instead of the 24 of pq_graph_copy's per-read arrays), and the sum of the
 * overlap lengths of all edges (average_overlap = sum / edges, string_graph.c:
 * 1051).  PQ_ERR_LIMIT if a node has more than 65535 edges on one side
 * (string_graph.c:469-472). */
int pq_graph_nodes(pq_ctx_t *ctx, pq_node_t *nodes, uint64_t *overlap_len_sum);

/* Node records (as pq_graph_nodes) and edges in the edge_t layout (as
 * pq_graph_copy_edges) of the last search, with one host wait: everything
 * construct_string_graph (string_graph.c:832-1021) needs in host memory.
 * nodes or edges may be NULL.  Into memory from pq_host_alloc the copies run
 * at the host link's rate. */
int pq_graph_export(pq_ctx_t *ctx, pq_node_t *nodes, void *edges, uint32_t index_bytes,
                    uint64_t *overlap_len_sum);
/* Search (as pq_find_overlaps; eager modes only, PQ_SEARCH_GRAPH is refused)
 * that exports the graph as it goes: once a batch's edges are extracted, its
 * node records (pq_graph_nodes) and its edges in the edge_t layout
 * (pq_graph_copy_edges) leave for nodes[n] and edges[edge_capacity] on a second
 * stream while the next batch is searched, so that only the last batch's
 * transfer follows the search.  Fills *n_edges and the overlap-length sum (either
 * may be NULL).  An edges buffer too small for the graph: the node records are
 * delivered, *n_edges says how many edges there are, the result stays on the
 * device for pq_graph_export, and PQ_ERR_LIMIT is returned.  Both the search and
 * the graph it hands over are the ones construct_string_graph produces
 * (string_graph.c:734-1021); nodes or edges may be NULL.  Into memory from
 * pq_host_alloc the copies run at the host link's rate. */
int pq_find_overlaps_export(pq_ctx_t *ctx, int mode, pq_node_t *nodes, void *edges,
                            uint64_t edge_capacity, uint32_t index_bytes, uint64_t *n_edges,
                            uint64_t *overlap_len_sum);
/* Allocates the device buffers a plain pq_find_overlaps_export of the uploaded,
 * indexed reads will use (staging for about four edges per read), so that the
 * search itself allocates nothing unless its graph is larger; drops the last
 * search's result.  No counterpart in the reference; for callers that time the
 * search apart from its set-up (construct_index prepares what
 * construct_string_graph uses). */
int pq_reserve_search(pq_ctx_t *ctx, uint32_t index_bytes);
/* Reads per search batch (0: automatic, batches whose candidate buffers take a
 * third of the free device memory).  Smaller batches cost launches; the result
 * is the same for any batch size.  No counterpart in the reference, which
 * searches read by read. */
int pq_set_search_batch(pq_ctx_t *ctx, uint64_t reads);
/* Page-locked host memory (hipHostMalloc) for the transfers above; released
 * by pq_host_free.  No counterpart in the reference, whose results never leave
 * host memory. */
int pq_host_alloc(pq_ctx_t *ctx, uint64_t bytes, void **p);
void pq_host_free(void *p);

/* ---- learned index --------------------------------------------------------------------- */
/* Counterpart of the reference's interval shortcut (sufsort.c:179-189, used at
 * string_graph.c:203-212, 289-298): every chain of the overlap search resumes at
 * depth D = min(63, t) from the suffix-array interval of its first D codes,
 * which a learned index finds.  An FP32 MLP (k-code prefix -> hidden ReLU layer
 * -> 1), fitted on the device to the freshly built suffix array, routes a
 * pattern to one of `leaves` piecewise-linear leaves (`segments` pieces with
 * exact end points) that predict its suffix-array row (to the sampling stride);
 * an exact correction finds the interval: galloping + binary search over the
 * keys of every sample_stride-th row (the exact part, 16 bytes per sampled row),
 * then the rows within the bracketing strides on the suffix array and the text:
 * an interval that may hold at most the seed kernel's limit of rows (pq_set_seed_verify)
 * is scanned outward from the samples, and its chain finished on the text on
 * the way; a wider one by binary search.
 * Results equal PQ_SEARCH_PLAIN's for any model.  The model is dropped by
 * pq_upload_reads and pq_build_index. */
/* Index builds from now on also write the sampled keys of the learned index
 * (every stride-th suffix-array row; stride a power of two, 0 = no) while they
 * sort, for a later pq_learned_train with that sample_stride, which then skips
 * its own pass over the suffix array and the text.  Results are identical. */
int pq_set_learned_keys(pq_ctx_t *ctx, uint32_t stride);
typedef struct {
    uint32_t k;              /* prefix length fed to the model (<= 21; capped at D), 21 */
    uint32_t hidden;         /* MLP hidden width (<= 64), 64 */
    uint32_t leaves;         /* piecewise-linear leaves, 4096 */
    uint32_t segments;       /* linear pieces per leaf (<= 65536); 0 (the default):
                                about 8 sampled rows per piece, 16..65536 */
    uint32_t epochs;         /* Adam passes over the sampled training set, 1 */
    uint32_t batch;          /* minibatch size (multiple of 64), 16384 */
    uint32_t samples;        /* training samples drawn from the sampled rows (multiple of 64) */
    float lr;                /* Adam step size (decays linearly to 1%) */
    uint32_t closed_form;    /* 1: solve the output layer by least squares before Adam */
    uint64_t seed;
    uint64_t sample_stride;  /* suffix-array rows per sampled key of the exact part (8) */
} pq_learned_cfg_t;

typedef struct {
    double mean_window[2];   /* mean |exact - predicted sampled row| over the last learned
                                search's lookups, per index (fw, bw), in sampled rows */
    uint64_t max_window[2];  /* largest such correction window */
    double train_loss[2];    /* mean squared error of the MLP on row / rows */
    double fit_loss[2];      /* the same before Adam (after the closed-form layer) */
    double train_ms;         /* wall time of pq_learned_train (both directions) */
    uint64_t lookups[2];     /* lookups in the last learned search, per index */
    uint64_t probes[2];      /* key reads of those lookups (sampled keys and
                                suffix-array rows) */
    uint64_t table_size[2];  /* sampled rows per index */
    uint64_t absent;         /* chains whose depth-D pattern does not occur */
    uint64_t fallback;       /* always 0: a seed is the searched index's own interval */
    uint32_t depth;          /* D: depth the backward search resumes from */
    uint32_t k;              /* model prefix length used */
    uint32_t segments;       /* linear pieces per leaf used */
} pq_learned_stats_t;

void pq_learned_default(pq_learned_cfg_t *cfg);
int pq_learned_train(pq_ctx_t *ctx, const pq_learned_cfg_t *cfg);
int pq_learned_stats(pq_ctx_t *ctx, pq_learned_stats_t *st);
/* Host copy of the trained parameters of direction dir (0 fw, 1 bw): W1[hidden][k],
 * b1[hidden], w2[hidden], b2.  *count in: capacity, out: number of parameters
 * (params may be NULL to query it). */
int pq_learned_params(pq_ctx_t *ctx, int dir, float *params, uint32_t *count);
/* The trained network of direction dir evaluated on n k-code prefix values xs
 * (first code most significant), as the learned search evaluates it on the
 * matrix cores (mfma = 1) or by the scalar form (mfma = 0). */
int pq_mlp_eval(pq_ctx_t *ctx, int dir, const uint64_t *xs, uint32_t n, int mfma, float *out);
/* Host copy of the exact part of the learned index of direction dir: the keys of
 * every sample_stride-th suffix-array row, two uint64 per sample (high half
 * first; the suffix's first D codes, first code in the top bits, codes past the
 * end of its read 0, bit 0 of the low half set when it has at least D codes);
 * first: the row; rows: the rows the sample stands for.  *count in: capacity,
 * out: number of samples (outputs may be NULL to query). */
int pq_learned_table(pq_ctx_t *ctx, int dir, uint64_t *keys, uint64_t *first, uint32_t *rows,
                     uint64_t *count);
/* Gradient of (1/n) sum (MLP(x_i) - y_i)^2 over n samples, by the device training
 * kernels; x_i are k-code prefix values (first code most significant), params as
 * above.  grad receives one float per parameter, *loss the mean squared error. */
int pq_mlp_gradients(pq_ctx_t *ctx, uint32_t k, uint32_t hidden, const float *params,
                     const uint64_t *xs, const float *targets, uint32_t n, float *grad,
                     float *loss);

/* ---- autotuner ------------------------------------------------------------------------- */
/* The first extraction tier processes one read per group of lanes holding up to
 * `items` overlap candidates per read side (4 per lane in registers; in LDS for
 * extension keys over 128 codes); larger sides fall through to wider tiers; the
 * best shape depends on the read set (default: shape 0).  pq_autotune runs the
 * search (mode as pq_find_overlaps) once per shape on the uploaded, indexed reads,
 * keeps the fastest for later searches of this context and leaves that shape's
 * (identical) result in place; then the search kernel and the verification
 * threshold the same way.  pq_set_tuning forces a shape. */
typedef struct {
    uint32_t shape;          /* chosen shape index */
    uint32_t group_lanes;    /* lanes per read of that shape (4, 8, 16; twice that
                                for keys over 128 codes) */
    uint32_t items;          /* candidates per side the first tier holds (16, 32, 64) */
    uint32_t nshapes;
    double extract_ms[4];    /* extraction time measured per shape */
    uint32_t find_chains;    /* candidate-search kernel chosen (1 to 4, below; 0 for
                                PQ_SEARCH_LEARNED, which has one) */
    double find_ms[4];       /* candidate-search time of kernels 1 to 4 (with
                                PQ_SEARCH_LEARNED: seed phase + search) */
    uint32_t verify_rows;    /* verification threshold chosen (0, 8 or 32, pq_set_verify) */
    double verify_ms[3];     /* candidate search + verification time with 0, 8 and 32
                                (with PQ_SEARCH_LEARNED: seed phase included) */
    uint32_t seed_rows;      /* PQ_SEARCH_LEARNED: seed kernel's threshold chosen (0, 8
                                or 32, pq_set_seed_verify); PQ_SEARCH_PLAIN: unchanged */
    double seed_ms[3];       /* PQ_SEARCH_LEARNED: seed phase + search with 0, 8 and 32 */
} pq_tune_t;
int pq_autotune(pq_ctx_t *ctx, int mode, pq_tune_t *out);
int pq_set_tuning(pq_ctx_t *ctx, uint32_t shape);
/* The candidate search runs the four chains of a read on four lanes (1), the
 * two chains of each index interleaved on one lane (2), or on four lanes that
 * load each other's occurrence lines together, one 64-byte access per line (3),
 * or one chain per lane with each lane taking the batch's next chain as its own
 * ends (4); same result.  pq_autotune times all four and keeps the fastest; until then
 * (or after chains = 0) the kernel follows the size of both indexes'
 * occurrence lines: 2 below 1 GiB, 1 below 4 GiB, 3 from there on.  The seeded
 * search (PQ_SEARCH_LEARNED) ignores the choice: it walks the chains the seed
 * kernel left unfinished from a work list, one chain per lane, each lane taking
 * the next listed chain as its own ends. */
int pq_set_find_chains(pq_ctx_t *ctx, uint32_t chains);
/* A chain of the candidate search whose interval holds at most `rows` suffix-
 * array rows (<= 32) at a depth below the minimum overlap is finished on the
 * text: each row's read and offset come from the suffix array, and how far the
 * chain's next codes match left of it from the packed texts, 32 codes per
 * compare; the rows' reaches give the depths the backward search would have
 * gone on to and the reads it would have recorded there (same result).  32 by
 * default; 0: every chain is searched backward to its end. */
int pq_set_verify(pq_ctx_t *ctx, uint32_t rows);
/* The seeded search's own threshold (PQ_SEARCH_LEARNED): a chain whose depth-D
 * interval may hold at most `rows` rows (<= 32) is scanned and finished on the
 * text by the seed kernel; wider ones, and all with 0, are handed to the search
 * kernel (which verifies by pq_set_verify's threshold).  32 by default; same
 * result for every value.  pq_autotune times 0, 8 and 32 for the seeded search. */
int pq_set_seed_verify(pq_ctx_t *ctx, uint32_t rows);
/* Initial capacity of the search's shared edge buffer: edges of read sides with
 * more than 16 irreducible edges (the first 16 of every side have slots of
 * their own).  0 (the default): max(2^20, reads / 4).  A search that needs more
 * counts what it needed and runs once more with the buffer at that size (kept
 * for later searches of the same reads). */
int pq_set_edge_capacity(pq_ctx_t *ctx, uint64_t edges);

/* ---- timing ---------------------------------------------------------------------------- */
typedef struct {
    double upload_ms;
    double sa_ms[2];         /* suffix sort per direction */
    double occ_ms[2];        /* BWT + occurrence lines + reads_index */
    double find_ms;          /* candidate intervals (bidirectional search) */
    double extract_ms;       /* irreducible-edge extraction */
    double compact_ms;       /* edge CSR compaction */
    double learned_ms;       /* learned seeding (learned mode only): prediction + seeds */
    double seed_kernel_ms;   /* the seed kernel alone (k_learned_seed) */
    uint64_t find_launches;
    double find_kernel_ms;   /* summed duration of the candidate-search kernel */
    uint64_t find_positions; /* rank positions evaluated by it (2 per step) */
    uint64_t find_lines;     /* occurrence lines it loaded (1 or 2 per step) */
    uint64_t verify_rows;    /* suffix-array rows finished on the text after backward
                                steps (pq_set_verify) */
    uint64_t seed_rows;      /* rows the learned seed kernel read on the text (members
                                of the intervals it scanned and the rows just past them) */
    uint64_t extract_items;  /* candidate items the extraction expanded (both sides) */
    uint64_t extract_reads[4]; /* reads finished per extraction tier: the three LDS
                                tiers in launch order, then global scratch */
} pq_timing_t;
int pq_get_timing(pq_ctx_t *ctx, pq_timing_t *t);
int pq_sync(pq_ctx_t *ctx);
/* Measurement helper: one plain device-to-device copy of `bytes` bytes by a
 * streaming kernel (k_copy_probe), the known transfer against which hardware
 * traffic counters are scaled; *ms gets its duration. */
int pq_probe_copy(pq_ctx_t *ctx, uint64_t bytes, double *ms);
/* Measurement helper: one streaming read of `bytes` bytes (k_read_probe), the
 * read-only ceiling of the device; *ms gets its duration. */
int pq_probe_read(pq_ctx_t *ctx, uint64_t bytes, double *ms);
/* Measurement helper: 2^25 scattered 64-byte line reads over a buffer of
 * `bytes` bytes (k_gather_probe).  mode 0: every lane loads whole lines as the
 * search loads occurrence lines, independent positions (throughput ceiling of
 * scattered lines); 1: each position depends on the previous line the lane
 * read; 2: four lanes load one line together, 16 bytes each; 3: every lane
 * loads 40 bytes of each line (three loads). */
int pq_probe_gather(pq_ctx_t *ctx, uint64_t bytes, int mode, double *ms);

/* ---- host utilities (no device use) --------------------------------------------------- */
/* Seeded synthetic workloads (see pq_synth.c for the generator contract). */
typedef struct {
    uint64_t genome_len;
    uint32_t n_units, unit_len, n_copies;   /* interspersed repeat units */
    double copy_div;
    double family_frac;                      /* repeat families */
    uint32_t fam_len_min, fam_len_max;
    double fam_div_min, fam_div_max;
    uint32_t fam_copies_min, fam_copies_max;
    uint32_t len_min, len_max;               /* read length range */
    double coverage;
    uint64_t n_reads;                        /* 0: ceil(coverage*G/mean length) */
    double sub_rate;                         /* i.i.d. substitutions per base */
    uint32_t min_overlap;                    /* t of the workload */
    uint64_t seed;
} pq_synth_cfg_t;

int pq_synth_preset(int config, pq_synth_cfg_t *cfg);
uint64_t pq_synth_num_reads(const pq_synth_cfg_t *cfg);
int pq_synth_genome(const pq_synth_cfg_t *cfg, uint8_t *genome);
int pq_synth_lengths(const pq_synth_cfg_t *cfg, uint64_t *off);
int pq_synth_reads(const pq_synth_cfg_t *cfg, const uint8_t *genome,
                   const uint64_t *off, uint8_t *codes);
int pq_synth_fasta(const pq_synth_cfg_t *cfg, const char *path);

/* Read import with duplicate removal, the rules of import_singlereads
 * (index.c:145-405) and bloom_filter (bloom.c:91-135) with hashes in
 * index_bits (32 or 64) arithmetic; replaces import_singlereads behind
 * construct_index (index.c:709-745).  The file is split into records on the
 * host in parallel (lines as the reference's 1000-byte fgets buffer cuts them,
 * index.c:218), and the duplicate filter runs on ctx's device
 * (pq_dedup_reads).  *reads is malloc'ed (n*read_len codes, input order); on
 * PQ_ERR_FORMAT bad_line is the line number the reference reports
 * (2 * reads kept before the bad record + 2, index.c:223). */
typedef struct {
    uint8_t *reads;
    uint64_t n;
    uint64_t removed;
    uint64_t C[6], suffix_count[4], prefix_count[4];
    double ms;
    long bad_line;           /* line number reported on PQ_ERR_FORMAT */
    uint64_t bloom_bits;     /* filter size: records * bloom_m in the index width */
} pq_import_t;
int pq_import_fasta(pq_ctx_t *ctx, const char *path, int read_len, int bloom_m,
                    int n_hashfunc, int index_bits, pq_import_t *out);

/* The duplicate filter of import_singlereads (index.c:327-333, bloom.c:91-135)
 * over nrec parsed records: codes = nrec*read_len codes 1..4 in file order,
 * ok[r] = 0 for a record holding N (dropped, sets no bits).  keep[r] = 1 for
 * the records the reference keeps, given a filter of bloom_bits bits (computed
 * in the index width by the caller, index.c:206) and n_hashfunc hashes.  The
 * order-dependent filter is resolved exactly by passes over all records until
 * no verdict changes; *passes (may be NULL) reports how many ran. */
int pq_dedup_reads(pq_ctx_t *ctx, const uint8_t *codes, const uint8_t *ok, uint64_t nrec,
                   int read_len, uint64_t bloom_bits, int n_hashfunc, int index_bits,
                   uint8_t *keep, uint32_t *passes);

/* The duplicate filter of pq_dedup_reads on records already in device memory,
 * allocating nothing: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  d_codes (nrec * read_len codes), d_ok and
 * d_keep (nrec bytes each) are device pointers; keep[r] as pq_dedup_reads sets
 * it.  Its scratch is a fixed device array of PQ_NOALLOC_SCRATCH_BYTES per
 * device, held until the process ends: the call needs
 * 2 * nrec * n_hashfunc * 8 + bloom_bits * 4 bytes of it (each part rounded up
 * to 256 bytes, plus 256) and fails with PQ_ERR_LIMIT beyond that.  Every call
 * on one device shares that array: calls that overlap on one device race.  It
 * waits on ctx's stream for each pass. */
#define PQ_NOALLOC_SCRATCH_BYTES (32u << 20)
int pq_dedup_reads_noalloc(pq_ctx_t *ctx, const uint8_t *d_codes, const uint8_t *d_ok,
                           uint64_t nrec, int read_len, uint64_t bloom_bits, int n_hashfunc,
                           int index_bits, uint8_t *d_keep, uint32_t *passes);

/* Device memory of ctx's device for callers of the entry points that take
 * device pointers (pq_dedup_reads_noalloc), and copies on ctx's stream that
 * return when the copy is done.  No counterpart in the reference. */
int pq_dev_alloc(pq_ctx_t *ctx, uint64_t bytes, void **p);
void pq_dev_free(void *p);
int pq_dev_upload(pq_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);
int pq_dev_download(pq_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
