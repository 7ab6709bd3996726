/*
 * casphy_hip.h -- C ABI of the MI355X-native CasPhy hot path (libcasphy_hip.so).
 *
 * The library replaces the GPU executables of CasPhy (cu/joinEdgeBitReverse.cu,
 * cu/joinEdgeIndel.cu, cu/radiusNeighbor.cu, cu/radiusNeighborDNA.cu,
 * cu/tryRadius.cu, cu/tryRadiusDNA.cu, cu/TransferDisStream.cu) and the Rcpp CPU
 * twins they mirror (src/functions.cpp, registered in src/RcppExports.cpp).
 * Every computation runs in hand-written HIP kernels for gfx950; the host side
 * only moves buffers and launches work.  Results are integer and bit-exact with
 * the reference's CPU path.
 *
 * Conventions
 *  - Status: 0 = ok, < 0 = CASPHY_E* error; casphy_last_error() gives a
 *    thread-local message.  No exception crosses the ABI.
 *  - Buffers are caller-owned host memory unless a parameter says "library-
 *    allocated" (free those with casphy_free).
 *  - Layouts are the reference's on-disk / R-matrix layouts:
 *      cell-major  [n][W]   (.tbmat; an R matrix with cells as columns)
 *      word-major  [W][n]   (.bmat;  an R matrix with cells as rows)
 *  - Weights are the .w doubles.  They must be non-negative integers whose sum
 *    is < 2^24, the range in which the reference's float (and fp16 <= 2048 for
 *    the indel join) accumulation is exact; other weights return CASPHY_EWEIGHT.
 *    NULL means unit weights (R's default, R/CasPhy_functions.R:575-576).
 *  - Thread safety: entry points are re-entrant; concurrent calls on one GPU are
 *    serialised per call on the library's per-thread stream.
 */
#ifndef CASPHY_HIP_H
#define CASPHY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CASPHY_OK 0
#define CASPHY_EARG (-1)      /* bad argument (sizes, NULL pointers, ranges) */
#define CASPHY_EHIP (-2)      /* HIP runtime error */
#define CASPHY_ENOMEM (-3)    /* device or host allocation failed / over budget */
#define CASPHY_EWEIGHT (-4)   /* weights outside the exact integer contract */
#define CASPHY_EINTERNAL (-5)
#define CASPHY_ENODEV (-6)    /* no usable GPU */

/* Thread-local message of the last failing call on this thread. */
const char *casphy_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int casphy_version(void);
/* Frees a library-allocated host buffer. */
void casphy_free(void *p);
/* Selects the device used by this thread's later calls (default 0). */
int casphy_set_device(int device);
/* Blocks until this thread's stream is idle. */
int casphy_synchronize(void);
/* JSON list of the tile variants the autotuners chose so far (one entry per
 * shape class, with the time of every candidate): the matrix-core tiles, then
 * the VALU tiles (entries with "form": "valu").  Writes at most cap
 * bytes including the terminating NUL and stores the full length in *len.
 * No counterpart in the reference, which has one kernel shape per path. */
int casphy_mfma_tuning(char *buf, size_t cap, size_t *len);

/* ------------------------------------------------------------------ scoring */
/* Full n x n shared-mutation score matrix of bit-packed cells, S[i*n+j] =
 * sum_s w_s b_s(i) b_s(j) (self score on the diagonal).
 * Replaces: calcShared src/functions.cpp:3160-3170 used over all pairs at
 * :3347-3357; calculateSharedKernel cu/joinEdgeBitReverse.cu:53-75. */
int casphy_score_bit_matrix(const uint32_t *mat_cellmajor, uint32_t n, uint32_t W,
                            const double *weight, int32_t *S);

/* Full n x n indel score matrix: S[i*n+j] = sum_s w[a-1][s] [a_is = a_js = a > 0].
 * mat: cell-major int32 (-1 missing, 0 unmutated, k > 0 allele); weight:
 * nstate x ncol state-major (NULL = unit).
 * Replaces: countApoFloat src/functions.cpp:261-270 over all pairs at
 * :1906-1915; calculateSharedKernel cu/joinEdgeIndel.cu:67-85. */
int casphy_score_indel_matrix(const int32_t *mat_cellmajor, uint32_t n, uint32_t ncol,
                              const double *weight, uint32_t nstate, int32_t *S);

/* ------------------------------------------------------------------ radius */
/* Radius neighbour lists of bit-packed cells: N_r(i) = ascending {j : S(i,j) >= r}
 * (i itself included when its self score reaches r).  nneighbor[n] receives
 * the list lengths; *idx (library-allocated) receives the concatenated lists,
 * *nidx their total length: exactly the .nNeighbor / .idx contents.
 * Replaces: main of cu/radiusNeighbor.cu:162-470 (argv: prefix radius) and the
 * radiusWeightedParallel loop of R/CasPhy_functions.R:160-208
 * (src/functions.cpp:1825-1857). */
int casphy_radius_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n,
                      const double *weight, int32_t radius, uint32_t *nneighbor,
                      uint32_t **idx, uint64_t *nidx);

/* Same for DNA in fasta2bin byte codes (4 per u32, cell-major [n][nwRaw]):
 * S(i,j) = #{p : code_i[p] & code_j[p] & 0xF != 0}.
 * Replaces: main of cu/radiusNeighborDNA.cu:111-380 and
 * radiusWeightedDNAParallel src/functions.cpp:4021-4036 (unit weights). */
int casphy_radius_dna(const uint32_t *mat_cellmajor, uint32_t n, uint32_t nwRaw,
                      int32_t radius, uint32_t *nneighbor, uint32_t **idx, uint64_t *nidx);

/* tryRadius: counts[k] = |N_r(q[k])| for the R query rows q (self included).
 * The reference draws q with time-seeded rand(); here the caller passes them.
 * Replaces: cu/tryRadius.cu:26-50,52-189 (argv: prefix radius nReplicate). */
int casphy_try_radius_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n,
                          const double *weight, int32_t radius, const uint32_t *q,
                          uint32_t R, uint32_t *counts);
/* Replaces: cu/tryRadiusDNA.cu:45-74,76-178. */
int casphy_try_radius_dna(const uint32_t *mat_cellmajor, uint32_t n, uint32_t nwRaw,
                          int32_t radius, const uint32_t *q, uint32_t R, uint32_t *counts);
/* Batched tryRadius: all radii in one launch, counts[k*R + j] for radii[k]
 * (the autoSplit tuning loop of R/CasPhy_functions.R:2010-2021 runs one
 * tryRadius process per radius step; this answers every step at once).
 * nradii <= 64, radii ascending. */
int casphy_try_radii_bit(const uint32_t *mat_wordmajor, uint32_t W, uint32_t n,
                         const double *weight, const int32_t *radii, uint32_t nradii,
                         const uint32_t *q, uint32_t R, uint32_t *counts);
int casphy_try_radii_dna(const uint32_t *mat_cellmajor, uint32_t n, uint32_t nwRaw,
                         const int32_t *radii, uint32_t nradii, const uint32_t *q,
                         uint32_t R, uint32_t *counts);

/* tryRadius counts with no device allocation: casphy_try_radii_bit's
 * matrix-core path on caller-owned DEVICE buffers -- mat_cellmajor (n x W
 * u32 bit rows), radii (nradii <= 64, ascending), q (R row ids; an id of n
 * or more scores as an all-zero row) and counts (nradii x R, counts[k*R + j])
 * -- enqueued on `stream` (a hipStream_t; NULL = the null stream) with no
 * host synchronisation, so the call can be captured into a graph.  Every
 * intermediate buffer (padded rows, stage-major copies and their fp4
 * expansions, the R x n score matrix) lives in one library array of 128 MiB
 * per device, held until the process ends; sizes that need more return
 * CASPHY_EARG with the amount.  Unit weights only.  Calls that overlap on one
 * device share the array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops with fixed memory). */
int casphy_try_radii_bit_noalloc(const uint32_t *mat_cellmajor, uint32_t n, uint32_t W, const int32_t *radii,
                                 uint32_t nradii, const uint32_t *q, uint32_t R, uint32_t *counts, void *stream);

/* ------------------------------------------------------------------ joins */
/* Greedy shared-mutation join of bit-packed cells (n >= 2).
 * edge[(2n-2)*2]: (parent, child) pairs in join order, 0-based, new nodes
 * numbered n, n+1, ... (the raw edge of greedyJoinEdgeBitReverse).
 * node_state[(2n-1)*W] (optional): per node id, seq & ~bistate;
 * node_bistate[(2n-1)*W] (optional): per node id.
 * ancestral_type: 0 or 1 (calcAncestralBitReverse / ...2).
 * Replaces: greedyJoinEdgeBitReverse src/functions.cpp:3316-3461
 * (R signature greedyJoinEdgeBitReverse(mat, weight, verbose, ancestralType),
 * R/RcppExports.R:264-266) and the join loop of cu/joinEdgeBitReverse.cu:285-505. */
int casphy_join_bit(const uint32_t *mat_cellmajor, uint32_t n, uint32_t W, const double *weight,
                    int ancestral_type, uint32_t *edge, uint32_t *node_state,
                    uint32_t *node_bistate);

/* Greedy join of indel states (n >= 2).  mat: cell-major int32 (-1/0/k).
 * weight: nstate x ncol (NULL = unit).  tip_size: n (NULL = all 1).
 * edge as above; node_state[(2n-1)*ncol] (optional) per node id;
 * branch[2n-2] (optional): per edge, sites with parent 0 and child > 0.
 * Replaces: greedyJoinEdge2 src/functions.cpp:1880-2017 (R signature
 * greedyJoinEdge2(mat, weight, tipSize, minSize, minRatio, verbose),
 * R/RcppExports.R:168-170) and cu/joinEdgeIndel.cu:244-468. */
int casphy_join_indel(const int32_t *mat_cellmajor, uint32_t n, uint32_t ncol,
                      const double *weight, uint32_t nstate, const uint32_t *tip_size,
                      uint32_t min_size, float min_ratio, uint32_t *edge, int32_t *node_state,
                      uint32_t *branch);

/* Cladewise emission: DFS from the root 2n-2 taking the lowest unused edge row
 * first; tips keep ids, the root becomes n, internal nodes take the next id on
 * first discovery.  edge_out[(2n-2)*2] relabelled in DFS order (.edge2 payload);
 * internal_order[n-1] = raw id of relabelled node n+k (row order of .nodeState2).
 * Replaces: cu/joinEdgeBitReverse.cu:526-597 / cu/joinEdgeIndel.cu:486-558 and
 * reOrderEdge2 src/functions.cpp:2234-2260 + R/CasPhy_functions.R:1632-1641. */
int casphy_cladewise(const uint32_t *edge_raw, uint32_t n, uint32_t *edge_out,
                     uint32_t *internal_order);

/* Batched join: `njobs` independent binary trees in one persistent work-queue
 * launch (bootstrap replicates, split clusters).  Job k's cells are rows
 * [offsets[k], offsets[k+1]) of mat_cellmajor (all of width W); its edges go to
 * edge + 2*(2*offsets[k] - 2*k), i.e. 2n_k-2 rows per job in job order.  offsets[0]
 * must be 0.  One workgroup runs each job from an atomic work queue.
 * No counterpart in the reference (which runs one process per tree,
 * R/CasPhy_functions.R:2389-2392); each job's result equals casphy_join_bit. */
int casphy_join_bit_batch(const uint32_t *mat_cellmajor, const uint32_t *offsets, uint32_t njobs,
                          uint32_t W, const double *weight, int ancestral_type, uint32_t *edge);
/* Batched indel joins (same job layout; tip_size holds offsets[njobs] sizes or
 * is NULL).  No counterpart in the reference (one joinEdgeIndel process per
 * cluster, R/CasPhy_functions.R:2199-2210); each job equals casphy_join_indel. */
int casphy_join_indel_batch(const int32_t *mat_cellmajor, const uint32_t *offsets, uint32_t njobs,
                            uint32_t ncol, const double *weight, uint32_t nstate,
                            const uint32_t *tip_size, uint32_t min_size, float min_ratio,
                            uint32_t *edge);
/* Bootstrap replicates of one indel matrix as one batched launch: replicate r
 * is the n rows of mat_cellmajor (n x ncol allele codes, -1 missing) at the
 * columns cols[r*ncol .. r*ncol + ncol - 1] (every index below ncol); unit
 * weights.  The replicates are built on the device from the matrix and the
 * indices, so only those travel.  edge: nrep x (2n-2) x 2, replicate after
 * replicate, each equal to casphy_join_indel on that replicate's matrix.
 * No counterpart in the reference (a caller resamples the columns and runs
 * one join per replicate); the batched form of casphy_join_indel_batch. */
int casphy_join_indel_bootstrap(const int32_t *mat_cellmajor, uint32_t n, uint32_t ncol, const uint32_t *cols,
                                uint32_t nrep, uint32_t *edge);

/* ------------------------------------------------------------------ transfer */
/* Transfer distance between bipartition bitsets (row-major [nrow][ncol] u32,
 * ntip tips, zero padding bits).  For each y: ymin[y] = min over x of
 * min(H, ntip - H) with H = popcount(x ^ y) (u32 arithmetic), starting from
 * 32*ncol; ysize[y] = min(popc(y), ntip - popc(y)); likewise xmin / xsize.
 * Replaces: cu/TransferDisStream.cu:22-344 (argv: Yfile Xfile outprefix ntip). */
int casphy_transfer(const uint32_t *Y, uint32_t ny, const uint32_t *X, uint32_t nx,
                    uint32_t ncol, uint32_t ntip, uint32_t *ymin, uint32_t *ysize,
                    uint32_t *xmin, uint32_t *xsize);
/* Batched transfer distance: one reference set Y against nrep replicate sets
 * whose rows are X[xoffsets[b] .. xoffsets[b+1]), all in one launch.
 * ymin[b*ny + y] = min over x of replicate b; xmin/xsize per X row; ysize[ny].
 * No counterpart in the reference (one TransferDisStream process per pair of
 * files); replicate b's results equal casphy_transfer(Y, X_b). */
int casphy_transfer_batch(const uint32_t *Y, uint32_t ny, const uint32_t *X,
                          const uint32_t *xoffsets, uint32_t nrep, uint32_t ncol, uint32_t ntip,
                          uint32_t *ymin, uint32_t *ysize, uint32_t *xmin, uint32_t *xsize);

/* ------------------------------------------------------------------ tipMerge */
/* Cluster labels of tips whose neighbour lists overlap: nneighbor[ntip] list
 * lengths and idx[sum] the concatenated ascending lists (ids < ntip, no
 * repeats), as radiusNeighbor writes them to .nNeighbor / .idx.  A pair
 * (j, t > j) of linked tips joins when the number s of tips in both lists
 * passes s > nneighbor[j] * p or s > nneighbor[t] * p (double precision);
 * pairs are folded into categories in the reference's order (windows of
 * 131,072 tips, then j, then list order) and member[ntip] receives each
 * tip's category, tips in no category numbered after the categories in tip
 * order -- the lines of the .member file.  A list longer than 1,048,576 ids
 * is refused (CASPHY_EARG), as the reference refuses it.
 * Replaces: cu/tipMerge.cu:143-461 (argv: prefix p; called by
 * R/CasPhy_functions.R:316,2025). */
int casphy_tip_merge(const uint32_t *nneighbor, const uint32_t *idx, uint32_t ntip, double p, int32_t *member);

/* ------------------------------------------------------------ clade bitsets */
/* Clade bitsets of a rooted tree given as the two columns of R's 1-based
 * edge matrix (ape phylo$edge: tips 1..ntip, internal nodes ntip+1..
 * ntip+nnode, one root, each node's edge listed before its children's edges
 * as in cladewise order).  rows receives one row of ceil(ntip / 32) u32 words
 * per internal node except the root (bit t-1 set for each tip t below it), in
 * the order nodeDesBig writes them (internal children by a backward scan of
 * the edges), *nrows their number (nnode - 1 for a tree), and node_size the
 * vector nodeDesBig returns (node_size[k] = tips below internal node
 * nnode - 1 - k, 0-based among the internal nodes, k < nnode - 1).  Edge lists
 * in other orders are refused (CASPHY_EARG); there is no row cache limit.
 * Replaces: src/functions.cpp:1277-1366 (nodeDesBig, called by
 * R/bootstrap_functions.R:3). */
int casphy_clade_bitsets(const int32_t *anc, const int32_t *des, uint32_t nedge, uint32_t nnode, uint32_t ntip,
                         uint32_t *rows, uint32_t *nrows, int32_t *node_size);

/* ------------------------------------------------------------------ nodeKeep */
/* Internal nodes of a joined tree to keep against outgroup rows: edge[nedge][2]
 * (0-based, tips 0..nnode, internal nodes nnode+1.. as in .edge2), node_state
 * nnode rows and outmat nout rows of `nbin` bit words (_bit) or `ncol` u32
 * states shifted by +1 (_indel), as in .nodeState2 / .tboutmat.  Edges are
 * walked in order; an edge to an internal node d is skipped while the last
 * evaluated node was kept and the edge's ancestor id is not below it,
 * otherwise d is kept when every outgroup row lacks at least `cutoff` of its
 * mutations (bit: set bits of d clear in the row; indel: sites where d's
 * state is above 1 and the row's is at least 1 and differs).  kept (room for
 * nnode ids) receives the kept ids in edge order, *nkept their number -- the
 * lines of the .retainNode file.
 * Replaces: cu/nodeKeepBit.cu:33-51,186-211 and cu/nodeKeepIndel.cu:32-48,
 * 133-158 (argv: prefix cutoff; called by R/CasPhy_functions.R:2277,2281). */
int casphy_node_keep_bit(const uint32_t *edge, uint32_t nedge, const uint32_t *node_state, uint32_t nnode,
                         uint32_t nbin, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept,
                         uint32_t *nkept);
int casphy_node_keep_indel(const uint32_t *edge, uint32_t nedge, const uint32_t *node_state, uint32_t nnode,
                           uint32_t ncol, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept,
                           uint32_t *nkept);

/* ------------------------------------------------------------- DNA format edge */
/* Root votes of outgroup samples: state nnode rows of nbin words of fasta2bin
 * byte codes (4 per word; tips and internal nodes, as .nodeState),
 * edge[nedge][2] 1-based (parent, child) as .edge, samples the rows idx[s]
 * (< nmat) of mat (nmat rows of nbin words, the .bin matrix).  A node's edge
 * state is the OR of its and its parent's rows (zero for a node that is no
 * edge's child); a sample's count against it is the number of positions whose
 * AND has a base bit (low nibble).  root[s] / nshared[s] receive the node of
 * the largest count and the count: vote_root with voteRoot2's reduction order
 * (0-based node; among the largest counts the first by (bit-reversed block of
 * 8 mod 32, block / 32, bit-reversed node mod 8), node 0 when every count is
 * 0), find_root with
 * findRoot's (the first edge child in edge order; 1-based node, 0 and -1 with
 * no edge).
 * Replaces: cu/voteRoot2.cu:34-169,171-323 (argv: matfile prefix; called by
 * R/DNA_functions.R:119-133) and findRoot / findRootAVX2Big
 * src/functions.cpp:4170-4212,4317-4410 (called by R/DNA_functions.R:117). */
int casphy_vote_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                     const uint32_t *mat, uint64_t nmat, const uint32_t *idx, uint32_t nsample, uint32_t *root,
                     uint32_t *nshared);
int casphy_find_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                     const uint32_t *mat, uint64_t nmat, const uint32_t *idx, uint32_t nsample, int32_t *root,
                     int32_t *nshared);

/* fasta2bin's encoding of an alignment: seqs nseq rows of len characters
 * (row-major, the sequences of a FASTA file as read).  is_variable[len]
 * receives 1 for the kept columns (some base -- A, T/U, C, G, counted with
 * the IUPAC ambiguity codes' bases -- in some but not all sequences),
 * valid[nseq] each sequence's characters other than '-', '.', 'N', 'n', out
 * (room for nseq * ceil(len / 4) words) the kept columns as byte codes, 4 per
 * word, *width the words per row (ceil(kept / 4)): the payload of the .bin
 * file, whose header is [nseq, width].
 * Replaces: cu/fasta2bin.cu:12-118,189-210 (main :256-363 reads the FASTA
 * file and writes the .bin, _names.txt and _positions.txt files). */
int casphy_fasta_encode(const uint8_t *seqs, uint32_t nseq, uint64_t len, uint8_t *is_variable, uint32_t *valid,
                        uint32_t *out, uint64_t *width);

/* One Newick tree of `text` (leading blanks and [comments] skipped, up to and
 * including its ';'; *consumed = the bytes read, so a caller loops over a
 * file of trees; *ntip = 0 when only blanks remain).  Numbered as ape's
 * read.tree numbers it: tips 1..ntip in the order their labels appear, or,
 * with tip_order (NUL-separated labels, tip_order_bytes in all), by the
 * position of their label there; internal nodes ntip+1.. in the order of
 * their '(' (the root first); edges in that preorder (cladewise).  *edge
 * (library-allocated) holds R's edge matrix column-major: the nedge parent
 * ids, then the nedge child ids; *labels (library-allocated) the tip labels
 * NUL-separated in tip order.  Branch lengths and internal node labels are
 * dropped.  No counterpart in the reference, where R reads the trees with
 * ape::read.tree before nodeDesBig (R/bootstrap_functions.R:1-3). */
int casphy_newick_read(const char *text, uint64_t len, const char *tip_order, uint64_t tip_order_bytes,
                       uint64_t *consumed, uint32_t *ntip, uint32_t *nnode, int32_t **edge, uint32_t *nedge,
                       char **labels, uint64_t *labels_bytes);

/* Transfer bootstrap expectation (Lemoine et al., Nature 556:452, 2018) of
 * the reference tree's clade rows from casphy_transfer_trees' outputs: with p
 * = min(ysize[y], ntip - ysize[y]) and the transfer index of each replicate
 * min(ymin[b * ny + y], p - 1), support[y] = 1 - (sum over the nrep
 * replicates of the indices) / ((p - 1) * nrep), NaN when p < 2.  No
 * counterpart in the reference, which stops at the transfer minima
 * (cu/TransferDisStream.cu:286-344). */
int casphy_tbe_support(const uint32_t *ymin, uint32_t nrep, uint32_t ny, const uint32_t *ysize, uint32_t ntip,
                       double *support);

/* Transfer distances from trees: the reference tree and nrep replicate trees
 * as R edge-matrix columns (cladewise, as casphy_clade_bitsets; replicate b's
 * edges at x_edge_offsets[b] .. x_edge_offsets[b+1], x_nnode[b] internal
 * nodes, all trees on the same ntip tips).  The clade bitsets of every tree
 * are built on the device (rows in nodeDesBig's order) and scored as
 * casphy_transfer_batch with Y = the reference's rows and X = the replicates'
 * rows: ymin[b * ny + y], ysize[ny], xmin / xsize per replicate row, where
 * ny = y_nnode - 1.  No counterpart in the reference, where R writes the
 * bitsets of each tree with nodeDesBig (R/bootstrap_functions.R:3) before
 * TransferDisStream reads them back (cu/TransferDisStream.cu:119-160). */
int casphy_transfer_trees(const int32_t *y_anc, const int32_t *y_des, uint32_t y_nedge, uint32_t y_nnode,
                          const int32_t *x_anc, const int32_t *x_des, const uint32_t *x_edge_offsets,
                          const uint32_t *x_nnode, uint32_t nrep, uint32_t ntip, uint32_t *ymin, uint32_t *ysize,
                          uint32_t *xmin, uint32_t *xsize);

/* ------------------------------------------------------------------ resident */
/* A packed cell matrix kept in HBM across calls (the serving / tuning form: one
 * upload, many radius and tryRadius passes).  kind: CASPHY_KIND_BIT (bit-packed
 * sites, `mat` word-major [W][n] like .bmat when wordmajor != 0, else cell-major)
 * or CASPHY_KIND_DNA (fasta2bin codes, cell-major [n][W=nwRaw]).
 * No counterpart in the reference, which re-reads its input file in every
 * tryRadius / radiusNeighbor process (R/CasPhy_functions.R:2010-2024). */
#define CASPHY_KIND_BIT 0
#define CASPHY_KIND_DNA 1
typedef struct casphy_cells casphy_cells;
int casphy_cells_create(int kind, const uint32_t *mat, uint32_t n, uint32_t W, int wordmajor,
                        const double *weight, casphy_cells **out);
void casphy_cells_destroy(casphy_cells *h);
/* Radius lists computed and kept on the device; *nidx = total list length. */
int casphy_cells_radius(casphy_cells *h, int32_t radius, uint64_t *nidx);
/* Copies the lists of the last casphy_cells_radius call (nneighbor[n], idx[nidx]). */
int casphy_cells_fetch(casphy_cells *h, uint32_t *nneighbor, uint32_t *idx);
/* tryRadius counts for up to 64 ascending radii: counts[k*R + j]. */
int casphy_cells_try_radii(casphy_cells *h, const int32_t *radii, uint32_t nradii, const uint32_t *q,
                           uint32_t R, uint32_t *counts);
/* One autoSplit round with a single host synchronisation: tryRadius counts of
 * the R query rows at each of up to 64 ascending radii, then the first radius
 * whose mean count c (= sum / R in double) satisfies
 * log10(max(c, 1e-30)) <= bound -- the stopping rule of the autoSplit loop,
 * R/CasPhy_functions.R:2007-2021 (bound = nNeighborBound - log10(ntip)) --
 * chosen on the device, and the radius lists at that radius kept on the
 * device as after casphy_cells_radius (fetch / last_ms apply).  *radius = -1
 * and *nidx = 0 when no radius qualifies.  counts (R x nradii, as
 * casphy_cells_try_radii) may be NULL.
 * Replaces: the tryRadius / radiusNeighbor executable round trips of that
 * loop (R/CasPhy_functions.R:2012-2025). */
int casphy_cells_autosplit(casphy_cells *h, const int32_t *radii, uint32_t nradii, const uint32_t *q,
                           uint32_t R, double bound, int32_t *radius, uint64_t *nidx, uint32_t *counts);
/* casphy_tip_merge over the lists of the last casphy_cells_radius or
 * casphy_cells_autosplit call, which stay on the device (the radius mask is
 * the dense form's bit rows).  No counterpart in the reference, where the
 * lists go through .idx / .nNeighbor files between radiusNeighbor and
 * tipMerge (R/CasPhy_functions.R:2024-2026). */
int casphy_cells_tip_merge(casphy_cells *h, double p, int32_t *member);
/* autoSplit's radius search for many groups of the resident cells at once:
 * run.split (R/CasPhy_functions.R:2896-2937) calls autoSplit (:1964-2082) on
 * mat[inIdx, ] for each group of a split layer, one tryRadius process per
 * radius step and one radiusNeighbor process per group.  Group g is the
 * n_g = goff[g+1] - goff[g] rows rows[goff[g] .. goff[g+1]) of h (goff[0] = 0,
 * every group non-empty); row rows[goff[g] + i] is the group's row i.  Its
 * queries are the group rows q[qoff[g] .. qoff[g+1]) (ids < n_g, qoff[0] = 0,
 * at least one per group).  Per group: tryRadius counts of its queries at
 * radii[0..nradii) (ascending, nradii <= 64), radius[g] = the first radius
 * whose mean count passes log10(mean) <= nneighbor_bound - log10(n_g) (the
 * rule of casphy_cells_autosplit with the group's own bound), or -1 when none
 * does, and the group's radius lists at that radius, ids local to the group
 * (nidx[g] ids; none for radius -1).  On the matrix cores every phase is one
 * launch over all groups; cells on the VALU tiles run the groups in turn.
 * The lists stay on the device for casphy_cells_group_fetch /
 * casphy_cells_group_tip_merge until the next group search. */
int casphy_cells_split_groups(casphy_cells *h, const uint32_t *goff, const uint32_t *rows, uint32_t ngroups,
                              const int32_t *radii, uint32_t nradii, const uint32_t *qoff, const uint32_t *q,
                              double nneighbor_bound, int32_t *radius, uint64_t *nidx);
/* Group g's lists of the last group search: nneighbor[n_g], idx[nidx[g]]. */
int casphy_cells_group_fetch(casphy_cells *h, uint32_t g, uint32_t *nneighbor, uint32_t *idx);
/* casphy_tip_merge over group g's lists of the last group search (the rest
 * of autoSplit's split step for that group, cu/tipMerge.cu:143-461);
 * member[n_g] labels the group's rows. */
int casphy_cells_group_tip_merge(casphy_cells *h, uint32_t g, double p, int32_t *member);
/* Device time (ms, hipEvents on the handle's stream): tile_ms = the scoring
 * kernel of the last radius, try_radii or autosplit call (autosplit: the
 * radius-mask tiles); radius_ms = the whole last radius or autosplit call;
 * 0 for a time no call has measured yet. */
int casphy_cells_last_ms(casphy_cells *h, float *tile_ms, float *radius_ms);

/* ------------------------------------------------------------------ serving */
/* A request loop over one GPU: jobs of mixed sizes are submitted from any
 * thread and run by `workers` worker threads (one stream each, so their
 * launches overlap on the device).  A worker takes the first queued job whose
 * estimated device footprint fits `hbm_budget_bytes` next to the running jobs
 * and the resident matrices; once the oldest job has waited
 * CASPHY_SERVE_HOL_MS (default 500) without fitting, later jobs start only if
 * they leave room for it.  Joins of at most 2,048 cells that are waiting when
 * a worker frees up are batched into one persistent launch; with two or more
 * workers one of them takes only such joins.  Every job's output equals the
 * plain entry point's (casphy_join_bit / _indel with unit weights,
 * casphy_transfer, casphy_transfer_trees).  No counterpart in the reference
 * (R forks one executable per job, R/CasPhy_functions.R:2389-2392). */
#define CASPHY_KIND_INDEL 2
#define CASPHY_JOB_TRANSFER 3
typedef struct casphy_server casphy_server;
int casphy_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, casphy_server **out);
/* Stops the workers after the running launches; queued jobs are dropped. */
void casphy_server_destroy(casphy_server *s);
/* kind CASPHY_KIND_BIT: mat = n x width u32 cell-major bits;
 * kind CASPHY_KIND_INDEL: mat = n x width int32 states (-1/0/k).  The input is
 * copied before the call returns. */
int casphy_server_submit_join(casphy_server *s, int kind, const void *mat, uint32_t n, uint32_t width,
                              int ancestral_type, uint64_t *job);
int casphy_server_submit_transfer(casphy_server *s, const uint32_t *Y, uint32_t ny, const uint32_t *X,
                                  uint32_t nx, uint32_t ncol, uint32_t ntip, uint64_t *job);
/* A packed matrix made resident on the server's device (kind CASPHY_KIND_BIT:
 * n x width u32 cell-major bits; CASPHY_KIND_INDEL: n x width int32 states),
 * uploaded before the call returns and counted against the HBM budget until
 * dropped; *matrix names it for casphy_server_submit_join_rows. */
int casphy_server_put_matrix(casphy_server *s, int kind, const void *mat, uint32_t n, uint32_t width,
                             uint64_t *matrix);
/* Drops a resident matrix; jobs already submitted on it still run. */
int casphy_server_drop_matrix(casphy_server *s, uint64_t matrix);
/* A join over rows rows[0..nrows) of a resident matrix (row i of the job is
 * row rows[i]; the output equals casphy_join_bit / _indel on those rows, the
 * subsets R's h.partial.join joins, R/CasPhy_functions.R:2124-2131).  The
 * rows are gathered on the device. */
int casphy_server_submit_join_rows(casphy_server *s, uint64_t matrix, const uint32_t *rows, uint32_t nrows,
                                   int ancestral_type, uint64_t *job);
/* Transfer distances from trees, the arguments of casphy_transfer_trees; the
 * output is ymin[nrep * ny], ysize[ny], then xmin and xsize of the
 * replicates' clade rows back to back (ny = y_nnode - 1). */
int casphy_server_submit_transfer_trees(casphy_server *s, const int32_t *y_anc, const int32_t *y_des,
                                        uint32_t y_nedge, uint32_t y_nnode, const int32_t *x_anc,
                                        const int32_t *x_des, const uint32_t *x_edge_offsets,
                                        const uint32_t *x_nnode, uint32_t nrep, uint32_t ntip, uint64_t *job);
/* Blocks until `job` is done and copies its output: joins (2n-2)*2 raw edge
 * values, transfers ymin[ny] ysize[ny] xmin[nx] xsize[nx].  A job is waited
 * once. */
int casphy_server_wait(casphy_server *s, uint64_t job, uint32_t *out, uint64_t out_len,
                       double *latency_ms);
/* Completed jobs, p50 / p99 latency (submit to completion) and throughput
 * since the first submission; `launches` counts worker launches. */
int casphy_server_stats(casphy_server *s, uint64_t *completed, double *p50_ms, double *p99_ms,
                        double *jobs_per_s, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* CASPHY_HIP_H */
