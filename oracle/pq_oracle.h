/*
 * pq_oracle.h -- CPU oracle for the suffix-array build and the FM-index overlap
 * search of PASQUAL.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: tests, the package's
 * smoke check and the CPU-baseline leg of the benchmark may call it; the
 * product (the HIP library under pasqual_amd/) never links, loads or calls it.
 *
 * Reads are passed as one concatenated array of nucleotide codes 1..4
 * (A,C,G,T; index.c:280-321) with read offsets off[0..n] (off[0] = 0).  All
 * index values are returned as uint64_t whatever INDEX_TYPE_T the reference was
 * built with; unsigned wrap-around of the reference's index arithmetic is
 * reproduced in 64 bits, which gives the same outcomes for every input whose
 * text fits the narrower type.
 *
 * Pinning: every function here is checked in tests/ against golden vectors
 * written by the reference itself (tests/golden/), against brute force on
 * small inputs, and against the explicit-string restatement below.
 */
#ifndef PQ_ORACLE_H
#define PQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Direction text: dir 0 = reads as given; dir 1 = every read reversed (not
 * complemented) -- transform_to_triplet_reverse, sufsort.c:80-121. */
void orc_dir_text(const uint8_t *text, const uint64_t *off, uint64_t n, int dir,
                  uint8_t *out);

/* Canonical suffix array of a direction text: all positions g sorted by the
 * suffix string up to the end of its read (end of read smallest, codes 1<2<3<4),
 * ties by g.  This is the order the reference's MSD sort produces at one
 * thread (sufsort.c:154-237 stable insertion sort with STNI_strcmp's
 * shorter-first rule; sufsort.c:240-383 stable counting passes; sufsort.c:
 * 483-546 text-order scatter).  Returns 0 on success. */
int orc_sa(const uint8_t *dtext, const uint64_t *off, uint64_t n, uint64_t *sa);

/* BWT codes (0 = '$' for a suffix starting a read, else the preceding code)
 * and reads_index (read ids of '$' rows in SA order): index.c:53-79. */
void orc_bwt(const uint8_t *dtext, const uint64_t *off, uint64_t n,
             const uint64_t *sa, uint8_t *bwt, uint64_t *reads_index);

/* C[6], suffix_count[4] (reads ending in c), prefix_count[4] (reads starting
 * with c): index.c:335-346, 366-372. */
void orc_counts(const uint8_t *text, const uint64_t *off, uint64_t n,
                uint64_t C[6], uint64_t suffix_count[4],
                uint64_t prefix_count[4]);

/* Occurrence table in the reference host layout: rows = (N+64)/64;
 * bitmap[c*rows + w] has bit (j+1) set iff bwt[64w+j] == c for j <= 62;
 * sample[5w + c] = fix[c] + #{i < 64w : bwt[i] == c}, fix = {0, f0..f3}
 * (index.c:26-142, 587-615). */
void orc_occ_ref(const uint8_t *bwt, uint64_t N, const uint64_t fix4[4],
                 uint64_t *bitmap, uint64_t *sample);

/* Interval shortcut of every read (fixed read length L): the reference's
 * interval_start/end/depth arrays (index.c:495-502, 548-555 initialise depth 1;
 * sufsort.c:177-209, 322-334 record deeper buckets; sufsort.c:354-376 and
 * 594-611 decide which buckets are visited). */
int orc_shortcut(const uint8_t *dtext, uint64_t n, uint32_t L, uint32_t t,
                 const uint64_t *sa, const uint64_t C[6], uint64_t *istart,
                 uint64_t *iend, uint64_t *idepth);

/* ---- overlap search, restated on the reference host layout ---------------- */

typedef struct {
    uint64_t n, N;
    uint32_t L, t;
    const uint8_t *reads;          /* fixed-length reads, codes 1..4 */
    uint64_t C[6];
    const uint64_t *bitmap[2];     /* [dir] -> 5*rows words */
    const uint64_t *sample[2];     /* [dir] -> 5*rows */
    const uint64_t *reads_index[2];
    const uint64_t *istart[2], *iend[2], *idepth[2];
} orc_fmindex_t;

/* Per-read search result in CSR form; edges of read r are
 * [edge_off[r], edge_off[r+1]), the B list first (num_B of them), then F. */
typedef struct {
    uint64_t n;
    int64_t *overlaps_B, *overlaps_F;
    uint64_t *num_B, *num_F;
    uint64_t *edge_off;
    uint64_t *dest;
    uint32_t *len;
    uint8_t *type;                 /* BF=0 FB=1 FF=2 BB=3 (string_graph.h:13-16) */
    uint64_t rank_evals;           /* rank-position evaluations (see .c) */
    uint64_t rank_evals_find;      /* ... of which in find_intervals */
} orc_graph_t;

/* find_intervals + extract_{bw,fw}_edges for every read (string_graph.c:
 * 162-731, driven as string_graph.c:800-830).  Returns 0, or -1 if a node
 * exceeds 65535 edges (string_graph.c:469-472). */
int orc_search_fm(const orc_fmindex_t *ix, orc_graph_t *out);

/* Independent restatement on explicit strings (generalised to reads of
 * different lengths): candidates (A.8) and irreducible edges (A.9) of every
 * read by hashing read ends, no FM index involved. */
int orc_search_explicit(const uint8_t *text, const uint64_t *off, uint64_t n,
                        uint32_t t, orc_graph_t *out);

void orc_graph_free(orc_graph_t *g);

/* Work of the device's candidate search (no counterpart in the reference, whose
 * find_intervals counts differently: the chains of pq_search.hip's k_find
 * kernels).  Per read of L codes four backward searches, each extending its
 * pattern by one code per step: FB n[L-o..L) in the forward index, FF
 * comp(n[L-o..L)) in the backward index, BF rev(n[0..o)) in the backward index,
 * BB comp(rev(n[0..o))) in the forward index.  A chain starts at depth o = 1
 * with the interval of its first code (or, when seed_depth is in 1..L-1, at
 * o = seed_depth with the interval of its seed_depth-code pattern) and steps
 * while o < L and the interval holds at least 2 rows (FB, BF) or 1 row (FF,
 * BB), the last step at o = L - 1 -- except that at a depth o <= t where the
 * interval holds at most vmax rows (vmax > 0) it stops stepping and its rows
 * are verified on the text instead.  A step evaluates 2 rank positions, l and
 * u + 1 of the interval [l, u], on 1 occurrence line of 128 rows when both fall
 * in the same one, else on 2.  out[0] = rank positions, out[1] = lines, out[2] =
 * rows handed to verification, summed over all reads.  stride > 0 (with
 * seed_depth): the seeded chains are found as the learned seed kernel finds
 * them, every stride-th suffix-array row sampled, and a seeded interval that may
 * hold at most vmax rows (at depth <= t) is scanned outward from the samples on
 * the text instead: out[2] counts the rows the scan reads (the members and the
 * rows just past its ends; for an interval between two samples, the rows from
 * the upper sample down to one past the interval), and a scanned interval of
 * minsize..vmax rows is finished there.  Returns 0, or -5 without memory. */
int orc_chain_work(const uint8_t *text, const uint64_t *off, uint64_t n, uint32_t seed_depth,
                   uint32_t t, uint32_t vmax, uint32_t stride, uint64_t out[3]);

/* Property check of a suffix array at any size: a permutation of 0..N-1 in
 * strictly increasing (suffix, position) order.  Returns -1 when it holds,
 * else the first offending SA index (or N if not a permutation). */
int64_t orc_check_sa(const uint8_t *dtext, const uint64_t *off, uint64_t n,
                     const uint64_t *sa);

/* Learned index model (no counterpart in the reference): the FP32 MLP of the
 * device learned index evaluated in double.  x_i is a k-code prefix value (first
 * code most significant); code c enters as c * 2/3 - 1.  params = W1[h][k],
 * b1[h], w2[h], b2; out = b2 + sum_i w2[i] relu(b1[i] + sum_j W1[i][j] f_j).
 * Computes loss = (1/n) sum (out_i - y_i)^2 and its gradient with respect to
 * every parameter (grad may be NULL).  Pinned in tests by central differences
 * of the loss, which this function also gives. */
double orc_mlp_loss_grad(uint32_t k, uint32_t h, const double *params, const uint64_t *xs,
                         const double *targets, uint64_t n, double *grad);

/* Outputs of the same MLP in double: out[i] = b2 + sum_j w2[j] relu(b1[j] +
 * sum_m W1[j][m] f_m(x_i)), the function whose loss orc_mlp_loss_grad takes. */
void orc_mlp_forward(uint32_t k, uint32_t h, const double *params, const uint64_t *xs,
                     uint64_t n, double *out);

/* First-triplet slot sizes of a direction text (initial_sort, sufsort.c:
 * 386-450): the triplet code of position g is 1 + 21 c0 + [rem >= 2] (1 + 5 c1 +
 * [rem >= 3] (1 + c2)) for the codes c0, c1, c2 (0..3) at g, g+1, g+2 and rem the
 * codes from g to the end of its read (sufsort.c:21-40, 124-151); slot k holds
 * codes 17k .. 17k+16.  size[k] = positions in slot k, offset = exclusive
 * prefix of size, *max_size = the largest size. */
void orc_slot_sizes(const uint8_t *dtext, const uint64_t *off, uint64_t n, uint64_t size[5],
                    uint64_t offset[5], uint64_t *max_size);

/* Reads import with duplicate removal, sequential, as import_singlereads
 * does it (index.c:145-405): a first pass counts the fgets(line, 1000) pieces
 * starting with '>' (index.c:195-206), the filter has records * bloom_m bits
 * computed in index_bits arithmetic, then pieces are taken two at a time
 * (header, sequence; the sequence's last byte dropped, index.c:253), codes
 * mapped (index.c:280-321), reads with N dropped, and bloom_filter
 * (bloom.c:91-135) with the hashes of bloom.c:18-86 decides duplicates.
 * Returns 0, or -2 (file cannot be opened), -3 (malformed: *bad_line = the
 * line number the reference prints, index.c:223), -5 (no memory), -7 (filter
 * of 0 bits).  *reads is malloc'ed (n * read_len codes); free with free(). */
int orc_import_fasta(const char *path, int read_len, int bloom_m, int n_hashfunc,
                     int index_bits, uint8_t **reads, uint64_t *n, uint64_t *removed,
                     long *bad_line);

/* bloom_filter (bloom.c:91-135) applied in record order to nrec parsed
 * records (codes 1..4, read_len each); records with ok[r] = 0 (holding N) are
 * skipped.  keep[r] = 1 when the record is kept. */
void orc_dedup(const uint8_t *codes, const uint8_t *ok, uint64_t nrec, int read_len,
               uint64_t bloom_bits, int n_hashfunc, int index_bits, uint8_t *keep);

#ifdef __cplusplus
}
#endif
#endif
