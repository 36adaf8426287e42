/*
 * Test infrastructure only.  A small driver linked against the reference's own
 * index/sufsort/string_graph objects (compiled from the reference sources by
 * oracle/Makefile) that writes every index_t array and every string-graph node
 * to one binary file, so tests can compare the oracle and the HIP library with
 * the reference bit for bit.
 *
 *   ref_dump <reads.fa> <read_len> <min_overlap> <out.bin> [nograph] [bloom=M] [nf=K]
 *
 * Prints the slot sizes of index_t (size/offset/max_size per direction) on a
 * line "slots fw <5 sizes> <5 offsets> <max> bw <5 sizes> <5 offsets> <max>".
 * Calls construct_index (index.h:64-68) with the CLI defaults bloom_m=36,
 * n_hashfunc=2 (test.c:136-148) unless bloom=/nf= say otherwise, and
 * construct_string_graph (string_graph.h:139).
 *
 * File layout: every integer is written as little-endian uint64.
 *   header  : magic 0x31464552514150 ("PQAREF1"), W, n, L, t, N
 *   reads   : N bytes (codes 1..4, input order), zero-padded to 8 bytes
 *   counts  : C[6], suffix_count[4], prefix_count[4]
 *   per dir (fw, bw):
 *             bitmap[5][(N+64)/64], sample[5*((N+64)/64)], reads_index[n],
 *             interval_start[n], interval_end[n], interval_depth[n]
 *   graph   : flag (1 if present); average_overlap as raw double bits;
 *             per node: num_B, num_F, max_B, max_F, B_type, F_type, status,
 *             marked, has_path_q, then (num_B+num_F) x (dest, len, type)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "index.h"
#include "string_graph.h"

static FILE *fo;

static void put(unsigned long long v)
{
    fwrite(&v, 8, 1, fo);
}

static void put_idx(const INDEX_TYPE_T *a, size_t k)
{
    for (size_t i = 0; i < k; i++)
        put((unsigned long long) a[i]);
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s reads.fa L t out.bin [nograph]\n", argv[0]);
        return 2;
    }
    int L = atoi(argv[2]), t = atoi(argv[3]);
    int nograph = 0, bloom_m = 36, nf = 2;
    for (int a = 5; a < argc; a++) {
        if (strcmp(argv[a], "nograph") == 0)
            nograph = 1;
        else if (strncmp(argv[a], "bloom=", 6) == 0)
            bloom_m = atoi(argv[a] + 6);
        else if (strncmp(argv[a], "nf=", 3) == 0)
            nf = atoi(argv[a] + 3);
    }
    index_t *ix = construct_index(argv[1], NULL, NULL, NULL, NULL, t, L, bloom_m, nf);
    size_t n = ix->num_reads, N = n * (size_t) L, rows = (N + 64) / 64;
    /* the first-triplet slot sizes of initial_sort (sufsort.c:386-450), one line */
    printf("slots fw");
    for (int k = 0; k < 5; k++)
        printf(" %llu", (unsigned long long) ix->size_fw[k]);
    for (int k = 0; k < 5; k++)
        printf(" %llu", (unsigned long long) ix->offset_fw[k]);
    printf(" %llu bw", (unsigned long long) ix->max_size_fw);
    for (int k = 0; k < 5; k++)
        printf(" %llu", (unsigned long long) ix->size_bw[k]);
    for (int k = 0; k < 5; k++)
        printf(" %llu", (unsigned long long) ix->offset_bw[k]);
    printf(" %llu\n", (unsigned long long) ix->max_size_bw);

    fo = fopen(argv[4], "wb");
    if (!fo) {
        perror("open");
        return 1;
    }
    put(0x31464552514150ULL);
    put(sizeof(INDEX_TYPE_T));
    put(n);
    put(L);
    put(t);
    put(N);
    fwrite(ix->reads, 1, N, fo);
    for (size_t pad = N; pad % 8; pad++)
        fputc(0, fo);
    put_idx(ix->C, 6);
    put_idx(ix->suffix_count, 4);
    put_idx(ix->prefix_count, 4);
    for (int d = 0; d < 2; d++) {
        u64 **bm = d ? ix->OCC_bitmap_bw : ix->OCC_bitmap;
        INDEX_TYPE_T *sm = d ? ix->OCC_sample_bw : ix->OCC_sample;
        for (int c = 0; c < 5; c++)
            fwrite(bm[c], 8, rows, fo);
        put_idx(sm, rows * 5);
        put_idx(d ? ix->reads_index_bw : ix->reads_index, n);
        put_idx(d ? ix->interval_start_bw : ix->interval_start, n);
        put_idx(d ? ix->interval_end_bw : ix->interval_end, n);
        u16 *dep = d ? ix->interval_depth_bw : ix->interval_depth;
        for (size_t i = 0; i < n; i++)
            put(dep[i]);
    }
    if (nograph) {
        put(0);
    } else {
        string_graph_t *g = construct_string_graph(ix);
        put(1);
        unsigned long long bits;
        memcpy(&bits, &g->average_overlap, 8);
        put(bits);
        for (size_t i = 0; i < n; i++) {
            node_t *nd = &g->nodes[i];
            put(nd->num_B);
            put(nd->num_F);
            put(nd->max_B);
            put(nd->max_F);
            put((unsigned char) nd->B_type);
            put((unsigned char) nd->F_type);
            put((unsigned char) nd->status);
            put((unsigned char) nd->marked);
            put(nd->path_q != NULL);
            for (int j = 0; j < nd->num_B + nd->num_F; j++) {
                put(nd->edges[j].dest);
                put(nd->edges[j].overlap_len);
                put((unsigned char) nd->edges[j].type);
            }
        }
    }
    fclose(fo);
    return 0;
}
