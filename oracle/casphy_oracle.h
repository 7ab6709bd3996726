/*
 * casphy_oracle.h -- CPU restatement of CasPhy's bit-packed phylogeny hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (casphy_amd/libcasphy_hip.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only to compute expected values or a
 * CPU timing.  The product never links, loads or calls anything here.
 *
 * Every function restates the reference's own algorithm (file:line cited on each
 * definition in casphy_oracle.c), keeping its float accumulation order, its
 * strict-compare tie rules and its output layouts.  Pinning of each function
 * (golden vectors, closed forms, brute force) lives in tests/test_oracle_*.py.
 */
#ifndef CASPHY_ORACLE_H
#define CASPHY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- (a-1) binary shared-mutation score ---------------------------------- */
/* Score of query row q against row t, both cell-major W words, float weights
 * (32*W of them).  Float accumulation in ascending set-bit order of q.       */
float orc_score_bit(const uint32_t *q, const uint32_t *t, uint32_t W, const float *w);
/* Full n x n matrix S[i*n+j] (diagonal = self score, as in the radius path). */
void orc_score_bit_matrix(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float *S);

/* ---- (a-2) indel shared-mutation score ----------------------------------- */
/* a, b: ncol int32 states (-1 missing, 0 unmutated, k>0 allele);
 * w: nState x ncol, state-major ((k-1)*ncol + s).                           */
float orc_score_indel(const int32_t *a, const int32_t *b, uint32_t ncol, const float *w);
void orc_score_indel_matrix(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w, float *S);

/* ---- (a-3) DNA similarity over fasta2bin byte codes ------------------------ */
/* a, b: nwRaw u32 words, 4 byte-codes each.  Count of positions whose codes
 * share a nucleotide bit (low nibble).                                       */
uint32_t orc_score_dna(const uint32_t *a, const uint32_t *b, uint32_t nwRaw);

/* ---- (a-7, a-8) radius neighbour lists / counts ---------------------------- */
/* Rows [i0, i1) of the radius lists.  mat cell-major.  counts[i-i0] gets
 * |N_r(i)|; when idx != NULL the ids are appended in ascending order at
 * idx[*pos...] (pos advanced).  Returns 0, or -1 when idx_cap is exceeded.  */
int orc_radius_bit_rows(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float radius,
                        uint32_t i0, uint32_t i1, uint32_t *counts, uint32_t *idx,
                        uint64_t idx_cap, uint64_t *pos);
int orc_radius_dna_rows(const uint32_t *mat, uint32_t n, uint32_t nwRaw, uint32_t radius,
                        uint32_t i0, uint32_t i1, uint32_t *counts, uint32_t *idx,
                        uint64_t idx_cap, uint64_t *pos);
/* tryRadius: counts[k] = |N_r(q[k])| (self included). */
void orc_try_radius_bit(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float radius,
                        const uint32_t *q, uint32_t R, uint32_t *counts);
void orc_try_radius_dna(const uint32_t *mat, uint32_t n, uint32_t nwRaw, uint32_t radius,
                        const uint32_t *q, uint32_t R, uint32_t *counts);

/* ---- (a-4, a-5) greedy shared-mutation join -------------------------------- */
/* Binary join, CPU twin greedyJoinEdgeBitReverse (O(n^2) score cache).
 * mat: n x W cell-major.  w: 32*W float.  edge: (2n-2) x 2 (parent, child) in
 * join order, 0-based.  node_state / node_bistate: (2n-1) x W indexed by node
 * id (tips are the input rows with bistate 0); node_state is seq & ~bistate. */
int orc_join_bit(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, int ancestral_type,
                 uint32_t *edge, uint32_t *node_state, uint32_t *node_bistate);
/* The same join driven the way the GPU executable drives it: O(n) lazy
 * local-max cache with recompute over all live nodes.                       */
int orc_join_bit_lazy(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, int ancestral_type,
                      uint32_t *edge, uint32_t *node_state, uint32_t *node_bistate);
/* Indel join, CPU twin greedyJoinEdge2.  mat: n x ncol int32 (-1/0/k).
 * w: nState x ncol state-major.  tip_size: n.  edge: (2n-2) x 2.
 * node_state: (2n-1) x ncol by node id.  branch: 2n-2 per-edge counts
 * (row 2 of greedyJoinEdge2's output).                                       */
int orc_join_indel(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w,
                   const uint32_t *tip_size, uint32_t min_size, float min_ratio,
                   uint32_t *edge, int32_t *node_state, uint32_t *branch);
int orc_join_indel_lazy(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w,
                        const uint32_t *tip_size, uint32_t min_size, float min_ratio,
                        uint32_t *edge, int32_t *node_state, uint32_t *branch);

/* ---- (a-6) cladewise emission --------------------------------------------- */
/* edge_raw: (2n-2) x 2 in join order.  edge_out: (2n-2) x 2 relabelled, in DFS
 * order.  internal_order[k] = raw node id of relabelled node n+k.            */
void orc_cladewise(const uint32_t *edge_raw, uint32_t n, uint32_t *edge_out, uint32_t *internal_order);

/* ---- (a-9) transfer distance ---------------------------------------------- */
void orc_transfer(const uint32_t *Y, uint32_t ny, const uint32_t *X, uint32_t nx,
                  uint32_t ncol, uint32_t ntip,
                  uint32_t *ymin, uint32_t *ysize, uint32_t *xmin, uint32_t *xsize);

/* ---- (f-1) tipMerge ----------------------------------------------------- */
/* Size of the intersection of two ascending id lists. */
uint32_t orc_list_shared(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb);
/* Cluster labels of cu/tipMerge.cu for ascending neighbour lists (nnb[i] ids
 * of tip i at idx + offsets[i]) and sharing fraction p; 1 when a list is
 * longer than the reference's 1,048,576-id limit. */
int orc_tip_merge(const uint32_t *nnb, const uint64_t *offsets, const uint32_t *idx, uint32_t ntip, double p,
                  int32_t *member_out);

/* ---- (f-3) nodeKeep ------------------------------------------------------- */
/* Kept internal nodes of nodeKeepBit (indel = 0: state / outgroup rows of
 * `width` u32 words of bits) or nodeKeepIndel (indel = 1: `width` u32 states
 * per row); edge[nedge][2] 0-based (tips 0..nnode, internal nodes after).
 * Returns the number of kept ids written to kept (at most nnode). */
uint32_t orc_node_keep(int indel, const uint32_t *edge, uint32_t nedge, const uint32_t *state, uint32_t nnode,
                       uint32_t width, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept);

/* ---- (f-2) bipartition producer ------------------------------------------- */
/* Clade bitsets of nodeDesBig: anc/des the columns of R's 1-based edge matrix;
 * rows_out holds (nnode - 1) x ceil(ntip / 32) words, size_out nnode - 1
 * entries (the reference's return vector).  Returns the rows written. */
uint32_t orc_clade_bitsets(const int32_t *anc, const int32_t *des, uint32_t nedge, uint32_t nnode, uint32_t ntip,
                           uint32_t *rows_out, int32_t *size_out);

/* ---- (f-4) DNA format edge ------------------------------------------------ */
/* voteRoot2: state nnode x nbin words (4 byte codes each), edge nedge x 2
 * 1-based (parent, child), samples nsample x nbin.  root = 0-based node of the
 * most shared positions under the reference's reduction order, nshared the
 * count. */
void orc_vote_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                   const uint32_t *samples, uint32_t nsample, uint32_t *root, uint32_t *nshared);
/* findRoot / findRootAVX2Big: first edge child of the largest count in edge
 * order; root 1-based (0 and nshared -1 with no edge). */
void orc_find_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                   const uint32_t *samples, uint32_t nsample, int32_t *root, int32_t *nshared);
/* fasta2bin on nseq equal-length rows of len characters: is_variable[len],
 * valid[nseq] (bases other than '-', '.', 'N', 'n'), out nseq x width words
 * of the variable columns' byte codes; returns width = ceil(nvar / 4). */
uint64_t orc_fasta_encode(const unsigned char *seqs, uint32_t nseq, uint64_t len, unsigned char *is_variable,
                          uint32_t *valid, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
