/*
 * pasqual_refabi.h -- the reference's own C interface, re-declared with the
 * same memory layout so that the library can replace the reference objects
 * index.o, sufsort.o and string_graph.o at link time.
 *
 * Layouts follow index.h:15-57 (index_t) and string_graph.h:47-132 (interval,
 * path, path list, edge, node and graph records); field order and types must
 * stay exactly as there.  The index width INDEX_TYPE_T is selected like the
 * reference build does, with -D__INDEX_U32__ or -D__INDEX_U64__ (default U32;
 * common.h:31-44).  Entry points: index.h:64-71, string_graph.h:139-142; the
 * global time_importreads: index.h:12.
 */
#ifndef PASQUAL_REFABI_H
#define PASQUAL_REFABI_H

#include <stdint.h>

#ifndef INDEX_TYPE_T
#if defined(__INDEX_U64__)
#define INDEX_TYPE_T unsigned long long
#else
#define INDEX_TYPE_T unsigned int
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _index_t {
    unsigned char *bloom_array;
    INDEX_TYPE_T bloom_size;
    int n_hashfunc;
    INDEX_TYPE_T C[6];                 /* $ A C G T # */
    INDEX_TYPE_T suffix_count[4];
    INDEX_TYPE_T prefix_count[4];
    unsigned long long *OCC_bitmap[5];
    unsigned long long *OCC_bitmap_bw[5];
    INDEX_TYPE_T *OCC_sample;
    INDEX_TYPE_T *OCC_sample_bw;
    INDEX_TYPE_T *reads_index;
    INDEX_TYPE_T *reads_index_bw;
    INDEX_TYPE_T *interval_start;
    INDEX_TYPE_T *interval_end;
    unsigned short *interval_depth;
    INDEX_TYPE_T *interval_start_bw;
    INDEX_TYPE_T *interval_end_bw;
    unsigned short *interval_depth_bw;
    void *read_lock;                   /* omp_lock_t * in the reference */
    char *reads;
    unsigned short read_len;
    INDEX_TYPE_T num_reads;
    unsigned short min_overlap;
    INDEX_TYPE_T size_bw[5];
    INDEX_TYPE_T offset_bw[5];
    INDEX_TYPE_T size_fw[5];
    INDEX_TYPE_T offset_fw[5];
    INDEX_TYPE_T max_size_fw;
    INDEX_TYPE_T max_size_bw;
} index_t;

/* edge types (string_graph.h:13-17) */
#define PQ_REF_BF 0
#define PQ_REF_FB 1
#define PQ_REF_FF 2
#define PQ_REF_BB 3
/* node types (string_graph.h:20-25) */
#define PQ_REF_ONE 1
#define PQ_REF_BRANCH 2
#define PQ_REF_ENDING 4
/* node status / marks (string_graph.h:28-38) */
#define PQ_REF_DEFAULT 0
#define PQ_REF_REMOVED (1 << 3)
#define PQ_REF_SUSPECT (1 << 2)
#define PQ_REF_PATH_NONE 0

typedef struct _interval_t {
    INDEX_TYPE_T ll;
    INDEX_TYPE_T uu;
    INDEX_TYPE_T remain_size;
    unsigned short len;
    char type;
} interval_t;

typedef struct _path_t {
    INDEX_TYPE_T end;
    INDEX_TYPE_T hop;
    unsigned short len;
    unsigned short end_len;
    unsigned short bubble_len;
    int trans;
    char updated;
    char direction;                    /* 1: F, 0: B */
    char type;
} path_t;

typedef struct _path_q_t {
    path_t *path;
    unsigned short max_F_endlen;
    unsigned short max_B_endlen;
    unsigned short max_B_len;
    unsigned short max_F_len;
    int max_F_trans;
    int max_B_trans;
} path_q_t;

typedef struct _edge_t {
    INDEX_TYPE_T dest;
    unsigned short overlap_len;
    char type;
} edge_t;

typedef struct _node_t {
    edge_t *edges;
    path_q_t *path_q;
    unsigned short num_F;
    unsigned short num_B;
    unsigned short max_F;
    unsigned short max_B;
    char B_type;
    char F_type;
    char status;
    char marked;
} node_t;

typedef struct _string_graph_t {
    unsigned short read_len;
    unsigned short min_overlap;
    double average_overlap;
    INDEX_TYPE_T num_reads;
    node_t *nodes;
} string_graph_t;

extern double time_importreads;

index_t *construct_index(char *single_file, char *pair_file1, char *pair_file2,
                         char *mate_file1, char *mate_file2, int min_overlap,
                         int read_len, int bloom_m, int n_hashfunc);
void free_index(index_t *index);
string_graph_t *construct_string_graph(index_t *index);
void destroy_graph(string_graph_t *graph);

#ifdef __cplusplus
}
#endif
#endif
