/*
 * pq_dropin.c -- the reference's entry points on top of the device library.
 *
 * Compiled once per index width (-D__INDEX_U32__ / -D__INDEX_U64__) so that the
 * structs of pasqual_refabi.h have the layout of the matching reference build.
 * Linked in place of index.o + sufsort.o + string_graph.o, the reference's CLI
 * (test.c), bloom.c, common.c and assemble.c run unchanged on top of it.
 *
 *   construct_index         index.c:709-745: reads import (pq_import.c, its
 *                           duplicate filter on the device), then
 *                           both suffix arrays / BWTs / occurrence tables on the
 *                           device; the host-layout arrays of index_t are filled
 *                           from the device index unless PQ_HOST_INDEX=0.
 *   construct_string_graph  string_graph.c:734-1054: overlap search on the
 *                           device, node records built here exactly as
 *                           string_graph.c:832-1021 builds them.
 *
 * Messages and exit behaviour follow the reference: its errors print
 * "Error: ..." to stderr and exit with status 0 (common.h:87-94).  Device
 * failures, which the reference cannot have, exit with status 1.
 *
 * Environment: PQ_DEVICE (HIP device, default 0), PQ_SEARCH=learned (seed the
 * search with the learned index), PQ_HOST_INDEX=0 (leave the host occurrence
 * arrays of index_t NULL; free(NULL) in the caller stays valid).
 */
#include <malloc.h>
#include <omp.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../../include/pasqual_amd.h"
#include "../../../include/pasqual_refabi.h"

double time_importreads = 0.0;

static pq_ctx_t *g_ctx;
static const index_t *g_index;

/* Page-locked staging for the search result (node records, edge_t records),
 * set up by construct_index for about four edges per read and grown by
 * construct_string_graph if the search finds more: device-to-host copies into
 * pageable memory go through a bounce buffer at a fraction of the link's rate. */
static void *g_stage_nodes, *g_stage_edges;
static uint64_t g_stage_nodes_bytes, g_stage_edges_bytes;

static void *stage(void **buf, uint64_t *have, uint64_t bytes)
{
    if (*buf && *have >= bytes)
        return *buf;
    pq_host_free(*buf);
    *buf = NULL;
    *have = 0;
    if (pq_host_alloc(g_ctx, bytes, buf) != PQ_OK)
        return NULL;
    *have = bytes;
    return *buf;
}

#define REF_ERROR_EXIT(...)                                                    \
    do {                                                                       \
        fprintf(stderr, "Error: ");                                            \
        fprintf(stderr, __VA_ARGS__);                                          \
        fflush(stderr);                                                        \
        exit(0);                                                               \
    } while (0)

static void dev_fail(const char *what, int rc)
{
    fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc,
            g_ctx ? pq_last_error(g_ctx) : "no device context");
    fflush(stderr);
    exit(1);
}

static double now_ms(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

static void ensure_ctx(void)
{
    if (g_ctx)
        return;
    int rc = pq_ctx_create(&g_ctx, env_int("PQ_DEVICE", 0));
    if (rc)
        dev_fail("device context", rc);
}

static void device_build(const char *reads, uint64_t n, int L, int t)
{
    ensure_ctx();
    uint64_t *off = malloc(sizeof(uint64_t) * (n + 1));
    if (!off)
        dev_fail("offsets", PQ_ERR_NOMEM);
    for (uint64_t i = 0; i <= n; i++)
        off[i] = i * (uint64_t) L;
    int rc = pq_upload_reads(g_ctx, (const uint8_t *) reads, off, n, (uint32_t) t);
    free(off);
    if (rc)
        dev_fail("read upload", rc);
    if ((rc = pq_build_index(g_ctx)))
        dev_fail("index build", rc);
}

static void export_host_index(index_t *ix)
{
    uint64_t n = ix->num_reads, N = n * (uint64_t) ix->read_len, rows = (N + 64) / 64;
    uint64_t *bm = malloc(sizeof(uint64_t) * 5 * rows);
    uint64_t *ri = malloc(sizeof(uint64_t) * (n + 1));
    uint64_t *ss = malloc(sizeof(uint64_t) * (n + 1));
    uint64_t *sz = malloc(sizeof(uint64_t) * (n + 1));
    if (!bm || !ri || !ss || !sz)
        dev_fail("host index buffers", PQ_ERR_NOMEM);
    for (int d = 0; d < 2; d++) {
        INDEX_TYPE_T *sample = malloc(sizeof(INDEX_TYPE_T) * 5 * rows);
        if (!sample)
            dev_fail("host index buffers", PQ_ERR_NOMEM);
        int rc = pq_get_occ_ref(g_ctx, d, bm, sample, (int) sizeof(INDEX_TYPE_T));
        if (rc)
            dev_fail("occurrence export", rc);
        unsigned long long **rowp = d ? ix->OCC_bitmap_bw : ix->OCC_bitmap;
        for (int c = 0; c < 5; c++) {
            rowp[c] = malloc(sizeof(unsigned long long) * rows);
            if (!rowp[c])
                dev_fail("host index buffers", PQ_ERR_NOMEM);
            memcpy(rowp[c], bm + (uint64_t) c * rows, sizeof(uint64_t) * rows);
        }
        INDEX_TYPE_T *rix = malloc(sizeof(INDEX_TYPE_T) * (n ? n : 1));
        INDEX_TYPE_T *st = malloc(sizeof(INDEX_TYPE_T) * (n ? n : 1));
        INDEX_TYPE_T *en = malloc(sizeof(INDEX_TYPE_T) * (n ? n : 1));
        unsigned short *dp = malloc(sizeof(unsigned short) * (n ? n : 1));
        if (!rix || !st || !en || !dp)
            dev_fail("host index buffers", PQ_ERR_NOMEM);
        if ((rc = pq_get_reads_index(g_ctx, d, ri)))
            dev_fail("reads_index export", rc);
        if ((rc = pq_get_shortcut(g_ctx, d, ss, sz, dp)))
            dev_fail("shortcut export", rc);
        for (uint64_t i = 0; i < n; i++) {
            rix[i] = (INDEX_TYPE_T) ri[i];
            st[i] = (INDEX_TYPE_T) ss[i];
            en[i] = (INDEX_TYPE_T) sz[i];
        }
        if (d == 0) {
            ix->OCC_sample = sample;
            ix->reads_index = rix;
            ix->interval_start = st;
            ix->interval_end = en;
            ix->interval_depth = dp;
        } else {
            ix->OCC_sample_bw = sample;
            ix->reads_index_bw = rix;
            ix->interval_start_bw = st;
            ix->interval_end_bw = en;
            ix->interval_depth_bw = dp;
        }
    }
    free(bm);
    free(ri);
    free(ss);
    free(sz);
}

index_t *construct_index(char *single_file, char *pair_file1, char *pair_file2,
                         char *mate_file1, char *mate_file2, int min_overlap,
                         int read_len, int bloom_m, int n_hashfunc)
{
    index_t *ix = calloc(1, sizeof(index_t));
    if (!ix)
        dev_fail("index record", PQ_ERR_NOMEM);
    ix->min_overlap = (unsigned short) min_overlap;
    ix->read_len = (unsigned short) read_len;
    ix->n_hashfunc = n_hashfunc;
    if (single_file != NULL) {
        pq_import_t imp;
        printf("Importing reads from %s ... ", single_file);
        fflush(stdout);
        ensure_ctx();
        int rc = pq_import_fasta(g_ctx, single_file, read_len, bloom_m, n_hashfunc,
                                 (int) (8 * sizeof(INDEX_TYPE_T)), &imp);
        if (rc == PQ_ERR_IO)
            REF_ERROR_EXIT("open reads file %s failed.\n", single_file);
        if (rc == PQ_ERR_FORMAT)
            REF_ERROR_EXIT("incorrect reads file at line %ld.\n", imp.bad_line);
        if (rc)
            dev_fail("reads import", rc);
        ix->reads = (char *) imp.reads;
        ix->num_reads = (INDEX_TYPE_T) imp.n;
        ix->bloom_size = (INDEX_TYPE_T) imp.bloom_bits;   /* index.c:205 */
        for (int c = 0; c < 6; c++)
            ix->C[c] = (INDEX_TYPE_T) imp.C[c];
        for (int c = 0; c < 4; c++) {
            ix->suffix_count[c] = (INDEX_TYPE_T) imp.suffix_count[c];
            ix->prefix_count[c] = (INDEX_TYPE_T) imp.prefix_count[c];
        }
        printf("done\n");
        printf("  has imported %llu reads (removed %llu duplicate reads)\n",
               (unsigned long long) imp.n, (unsigned long long) imp.removed);
        printf("  takes %.3lf ms\n", imp.ms);
        fflush(stdout);
        time_importreads = imp.ms;
    } else if (pair_file1 != NULL && pair_file2 != NULL) {
        REF_ERROR_EXIT("We do not support paired end reads.\nPlease input them as single reads.\n");
    } else if (mate_file1 != NULL && mate_file2 != NULL) {
        REF_ERROR_EXIT("We do not support mate pair reads.\nPlease input them as single reads.\n");
    }

    device_build(ix->reads, ix->num_reads, read_len, min_overlap);
    g_index = ix;
    pq_timing_t tm;
    pq_get_timing(g_ctx, &tm);
    printf("Sorting forward suffixes ... done\n  takes %.3lf ms\n", tm.sa_ms[0] + tm.occ_ms[0]);
    printf("Sorting backward suffixes ... done\n  takes %.3lf ms\n", tm.sa_ms[1] + tm.occ_ms[1]);
    fflush(stdout);
    /* first-triplet slot sizes of initial_sort (sufsort.c:386-450) */
    for (int d = 0; d < 2; d++) {
        uint64_t sz[5], of[5], mx = 0;
        int rc = pq_get_slot_sizes(g_ctx, d, sz, of, &mx);
        if (rc)
            dev_fail("slot sizes", rc);
        INDEX_TYPE_T *S = d ? ix->size_bw : ix->size_fw, *O = d ? ix->offset_bw : ix->offset_fw;
        for (int k = 0; k < 5; k++) {
            S[k] = (INDEX_TYPE_T) sz[k];
            O[k] = (INDEX_TYPE_T) of[k];
        }
        if (d)
            ix->max_size_bw = (INDEX_TYPE_T) mx;
        else
            ix->max_size_fw = (INDEX_TYPE_T) mx;
    }
    if (env_int("PQ_HOST_INDEX", 1))
        export_host_index(ix);
    /* staging for construct_string_graph's copies (its size is what the search
     * will need for about four edges per read) */
    stage(&g_stage_nodes, &g_stage_nodes_bytes, sizeof(pq_node_t) * ((uint64_t) ix->num_reads + 1));
    stage(&g_stage_edges, &g_stage_edges_bytes, sizeof(edge_t) * (4 * (uint64_t) ix->num_reads + 1));
    /* and the device buffers of the search */
    int rc = pq_reserve_search(g_ctx, (uint32_t) sizeof(INDEX_TYPE_T));
    if (rc)
        dev_fail("search buffers", rc);
    return ix;
}

void free_index(index_t *ix)
{
    if (!ix)
        return;
    for (int c = 0; c < 5; c++) {
        free(ix->OCC_bitmap[c]);
        free(ix->OCC_bitmap_bw[c]);
    }
    free(ix->OCC_sample);
    free(ix->OCC_sample_bw);
    free(ix->reads_index);
    free(ix->reads_index_bw);
    free(ix->interval_start);
    free(ix->interval_end);
    free(ix->interval_depth);
    free(ix->interval_start_bw);
    free(ix->interval_end_bw);
    free(ix->interval_depth_bw);
    free(ix->reads);
    if (g_index == ix)
        g_index = NULL;
    free(ix);
}

string_graph_t *construct_string_graph(index_t *ix)
{
    double t0 = now_ms();
    printf("Building overlap graph ... ");
    fflush(stdout);
    if (g_index != ix) {
        device_build(ix->reads, ix->num_reads, ix->read_len, ix->min_overlap);
        g_index = ix;
    }
    int mode = PQ_SEARCH_PLAIN;
    const char *sm = getenv("PQ_SEARCH");
    if (sm && strcmp(sm, "learned") == 0) {
        pq_learned_cfg_t cfg;
        pq_learned_default(&cfg);
        int rc = pq_learned_train(g_ctx, &cfg);
        if (rc)
            dev_fail("learned index training", rc);
        mode = PQ_SEARCH_LEARNED;
    }
    const int timing = env_int("PQ_TIMING", 0);
    double tp = now_ms();
    /* the search hands node records and edge_t records over batch by batch into
     * page-locked staging (set up by construct_index for about four edges per
     * read); a graph with more edges than the staging holds is exported again
     * after the staging has grown */
    const uint64_t n0 = pq_num_reads(g_ctx);
    pq_node_t *rec = stage(&g_stage_nodes, &g_stage_nodes_bytes, sizeof(pq_node_t) * (n0 + 1));
    edge_t *all = stage(&g_stage_edges, &g_stage_edges_bytes,
                        g_stage_edges_bytes > sizeof(edge_t) ? g_stage_edges_bytes
                                                             : sizeof(edge_t));
    if (!rec || !all)
        dev_fail("graph buffers", PQ_ERR_NOMEM);
    uint64_t n = n0, E = 0, len_sum = 0;
    int rc = pq_find_overlaps_export(g_ctx, mode, rec, all, g_stage_edges_bytes / sizeof(edge_t),
                                     (uint32_t) sizeof(INDEX_TYPE_T), &E, &len_sum);
    if (rc == PQ_ERR_LIMIT && E > g_stage_edges_bytes / sizeof(edge_t)) {
        all = stage(&g_stage_edges, &g_stage_edges_bytes, sizeof(edge_t) * (E + 1));
        if (!all)
            dev_fail("graph buffers", PQ_ERR_NOMEM);
        rc = pq_graph_export(g_ctx, rec, all, (uint32_t) sizeof(INDEX_TYPE_T), &len_sum);
    }
    if (timing) {
        pq_timing_t st;
        pq_get_timing(g_ctx, &st);
        fprintf(stderr, "[pq] search %.3f ms\n", now_ms() - tp);
        fprintf(stderr, "[pq] search find %.3f ms\n[pq] search extract %.3f ms\n",
                st.find_ms, st.extract_ms);
        tp = now_ms();
    }
    if (rc == PQ_ERR_LIMIT)
        REF_ERROR_EXIT("assembly failed. Try a larger min overlap length.\n");
    if (rc)
        dev_fail("overlap search", rc);
    pq_graph_sizes(g_ctx, &n, &E);
    string_graph_t *g = malloc(sizeof(string_graph_t));
    if (!g)
        dev_fail("graph buffers", PQ_ERR_NOMEM);
    if (timing)
        fprintf(stderr, "[pq] graph copy %.3f ms\n", now_ms() - tp), tp = now_ms();
    g->nodes = calloc(n ? n : 1, sizeof(node_t));
    if (!g->nodes)
        dev_fail("graph nodes", PQ_ERR_NOMEM);
    g->num_reads = (INDEX_TYPE_T) n;
    g->read_len = ix->read_len;
    g->min_overlap = ix->min_overlap;

    const int RL = ix->read_len;
    /* node records, string_graph.c:832-1021; edges are malloc'ed per node because
     * assemble.c frees them one node at a time.  The heaps grow in steps of 64 MiB
     * meanwhile (glibc's default step is 128 KiB): millions of small blocks from
     * many threads otherwise spend most of the loop growing the heaps, which
     * stalls the page faults of every other thread.  Each thread takes a
     * contiguous range of nodes, so its edge offset is a running sum. */
    mallopt(M_TOP_PAD, 64 << 20);
    const int T = omp_get_max_threads();
    uint64_t *tbase = calloc((size_t) T + 1, sizeof(uint64_t));
    if (!tbase)
        dev_fail("graph buffers", PQ_ERR_NOMEM);
    const uint64_t seg = (n + T - 1) / T;
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = seg * t < n ? seg * t : n, e = seg * (t + 1) < n ? seg * (t + 1) : n;
        uint64_t k = 0;
        for (uint64_t i = a; i < e; i++)
            k += (uint64_t) rec[i].num_B + rec[i].num_F;
        tbase[t + 1] = k;
#pragma omp barrier
#pragma omp single
        for (int q = 0; q < T; q++)
            tbase[q + 1] += tbase[q];
        uint64_t off = tbase[t];
        for (uint64_t i = a; i < e; i++) {
            node_t *nd = &g->nodes[i];
            const pq_node_t r = rec[i];
            const unsigned nb = r.num_B, nf = r.num_F;
            nd->max_B = nd->num_B = (unsigned short) nb;
            nd->max_F = nd->num_F = (unsigned short) nf;
            nd->B_type = r.B_type;
            nd->F_type = r.F_type;
            nd->status = r.status;
            nd->marked = r.marked;
            nd->path_q = NULL;
            nd->edges = malloc((nb + nf) * sizeof(edge_t));
            if (nb + nf)
                memcpy(nd->edges, all + off, (nb + nf) * sizeof(edge_t));
            off += nb + nf;
            int with_paths = (nd->B_type == PQ_REF_BRANCH && nd->F_type != PQ_REF_ENDING) ||
                             (nd->F_type == PQ_REF_BRANCH && nd->B_type != PQ_REF_ENDING);
            if (!with_paths)
                continue;
            path_q_t *pq = nd->path_q = malloc(sizeof(path_q_t));
            pq->max_B_endlen = pq->max_F_endlen = 0;
            pq->max_B_len = pq->max_F_len = 0;
            pq->max_F_trans = pq->max_B_trans = 0;   /* never read by assemble.c */
            pq->path = malloc(sizeof(path_t) * (nb + nf));
            for (unsigned j = 0; j < nb + nf; j++) {
                const edge_t *ed = &nd->edges[j];
                path_t *p = &pq->path[j];
                memset(p, 0, sizeof(*p));
                p->end = p->hop = ed->dest;
                p->len = 1;
                p->end_len = (unsigned short) (RL - ed->overlap_len);
                p->bubble_len = 0;
                p->type = PQ_REF_PATH_NONE;
                p->direction = j < nb ? (ed->type == PQ_REF_BB) : (ed->type == PQ_REF_FB);
            }
        }
    }
    mallopt(M_TOP_PAD, 128 * 1024);
    /* string_graph.c:1051: the sum of edge lengths over the number of edges */
    g->average_overlap = (double) len_sum / (double) E;
    if (timing)
        fprintf(stderr, "[pq] node build %.3f ms\n", now_ms() - tp);
    free(tbase);
    printf("  takes %.3lf ms\n", now_ms() - t0);   /* string_graph.c:1043 */
    printf("  average overlap length: %lf\n", g->average_overlap);
    fflush(stdout);
    return g;
}

void destroy_graph(string_graph_t *graph)
{
    (void) graph;       /* the reference leaves node memory to process exit */
}
