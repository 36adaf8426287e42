/*
 * Test infrastructure only.  The reference's own construct_index and
 * construct_string_graph (compiled from the reference sources by
 * oracle/Makefile) on a reads file, writing the overlap graph in the byte
 * layout the HIP library exports it (pq_graph_export), so that digests of the
 * two can be compared at sizes where a full dump is impractical:
 *
 *   ref_graph <reads.fa> <read_len> <min_overlap> <nodes.bin> <edges.bin>
 *
 *   nodes.bin: per node 8 bytes: num_B u16, num_F u16, B_type, F_type, status,
 *              marked (one byte each; string_graph.h node_t fields)
 *   edges.bin: per edge (node order, B list then F list) the edge_t layout of
 *              the index width: dest INDEX_TYPE_T, overlap_len u16, type u8, the
 *              padding zeroed (8 bytes for U32, 16 for U64)
 *
 * Prints the reads imported, the edges, and the sum of the overlap lengths.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "index.h"
#include "string_graph.h"

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s reads.fa L t nodes.bin edges.bin\n", argv[0]);
        return 2;
    }
    int L = atoi(argv[2]), t = atoi(argv[3]);
    index_t *ix = construct_index(argv[1], NULL, NULL, NULL, NULL, t, L, 36, 2);
    string_graph_t *g = construct_string_graph(ix);
    FILE *fn = fopen(argv[4], "wb"), *fe = fopen(argv[5], "wb");
    if (!fn || !fe) {
        perror("open");
        return 1;
    }
    unsigned long long edges = 0, len_sum = 0;
    const size_t rec = sizeof(INDEX_TYPE_T) == 4 ? 8 : 16;
    for (size_t i = 0; i < (size_t) ix->num_reads; i++) {
        node_t *nd = &g->nodes[i];
        unsigned char r[8];
        unsigned short nb = nd->num_B, nf = nd->num_F;
        memcpy(r, &nb, 2);
        memcpy(r + 2, &nf, 2);
        r[4] = (unsigned char) nd->B_type;
        r[5] = (unsigned char) nd->F_type;
        r[6] = (unsigned char) nd->status;
        r[7] = (unsigned char) nd->marked;
        fwrite(r, 1, 8, fn);
        for (int j = 0; j < nb + nf; j++) {
            unsigned char e[16];
            memset(e, 0, sizeof(e));
            INDEX_TYPE_T dest = nd->edges[j].dest;
            unsigned short len = nd->edges[j].overlap_len;
            memcpy(e, &dest, sizeof(dest));
            memcpy(e + sizeof(dest), &len, 2);
            e[sizeof(dest) + 2] = (unsigned char) nd->edges[j].type;
            fwrite(e, 1, rec, fe);
            edges++;
            len_sum += len;
        }
    }
    fclose(fn);
    fclose(fe);
    printf("ref_graph reads %llu edges %llu len_sum %llu\n",
           (unsigned long long) ix->num_reads, edges, len_sum);
    return 0;
}
