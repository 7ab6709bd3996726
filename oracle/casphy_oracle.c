/*
 * casphy_oracle.c -- CPU restatement of CasPhy's bit-packed phylogeny hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library.  Built by oracle/Makefile with the reference's own host flags
 * (src/Makevars:1: -O3 -mavx2 -mbmi2 -march=native).
 *
 * Citations are path:line in the CasPhy reference (LiLin-biosoft/CasPhy).
 * Each function keeps the reference's float accumulation order, strict-compare
 * tie rules and output layout.  Pins (tests/test_oracle_*.py):
 *   orc_score_bit / _matrix   closed form popcount(AND) for unit weights, brute
 *                             force over sites for integer weights, symmetry.
 *   orc_score_indel           brute force over (site, allele) one-hot bits.
 *   orc_score_dna             IUPAC set-intersection on decoded FASTA text.
 *   orc_radius_* / try_*      brute force {j : S(i,j) >= r} from the closed form;
 *                             monotone in r; symmetric; self iff S(i,i) >= r.
 *   orc_join_*                O(n^3) lexicographic brute force on n <= 48 (pure
 *                             Python), CPU twin vs lazy GPU-style variant,
 *                             3-valued truth tables of the ancestral rules,
 *                             tree invariants.
 *   orc_cladewise             preorder / relabel invariants, reOrderEdge2 +
 *                             factor relabel restated in Python.
 *   orc_transfer              brute force min(H, n-H) in numpy, identical trees
 *                             give 0, complement invariance.
 *   orc_list_shared           numpy set intersection.
 *   orc_tip_merge             p >= 1: no merges (labels 0..n-1); disjoint
 *                             cliques: one label per clique numbered by (window
 *                             of its second tip, its first tip), singletons
 *                             after -- also past 131,072 tips; the OR of the
 *                             two sharing fractions and the strict > on
 *                             hand-built lists; labels never join tips that
 *                             the lists do not link.
 *   orc_clade_bitsets         on cladewise trees every row is the clade of
 *                             its node (numpy walk from each tip to the root),
 *                             sizes are the row popcounts, rows come in the
 *                             backward-scan order of the internal children.
 *   orc_node_keep             numpy counts of missing mutations per (node,
 *                             outgroup); on DFS-ordered join trees the kept set
 *                             is the retained nodes without a kept proper
 *                             ancestor (a tree-walk form of the skip rule).
 *   orc_vote_root             numpy counts over unpacked bytes; the winner in
 *                             closed form (first by bit-reversed block mod 32,
 *                             block / 32, bit-reversed node mod 8 among the
 *                             largest counts) against the literal simulation
 *                             of the reductions here; hand-worked ties.
 *   orc_find_root             numpy counts, first maximum in edge order.
 *   orc_fasta_encode          numpy final-count form of the column rule (OR
 *                             of the IUPAC base sets has a base the AND
 *                             lacks), codes from the IUPAC base sets, on
 *                             random alphabets and the reference's example
 *                             alignment; a hand-worked case.
 */
#include "casphy_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * (a-1) binary score.  src/functions.cpp:3160-3170 (calcShared) with the set
 * bits listed as in int2idx (src/functions.cpp:112-126); the GPU executables do
 * the same per thread (cu/joinEdgeBitReverse.cu:53-75, cu/radiusNeighbor.cu:26-50)
 * with float weights converted from the .w doubles (cu/radiusNeighbor.cu:240-243).
 * ------------------------------------------------------------------------- */
float orc_score_bit(const uint32_t *q, const uint32_t *t, uint32_t W, const float *w)
{
    float s = 0.0f;
    for (uint32_t x = 0; x < W; ++x) {
        uint32_t v = q[x];
        while (v) {
            uint32_t b = (uint32_t)__builtin_ctz(v);
            if ((t[x] >> b) & 1u) s += w[x * 32u + b];
            v &= v - 1u;
        }
    }
    return s;
}

void orc_score_bit_matrix(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float *S)
{
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < n; ++j)
            S[(size_t)i * n + j] = orc_score_bit(mat + (size_t)i * W, mat + (size_t)j * W, W, w);
}

/* ---------------------------------------------------------------------------
 * (a-2) indel score.  src/functions.cpp:261-270 (countApoFloat): a site counts
 * when both states are equal and > 0, weight indexed (state-1, site).  The GPU
 * twin (cu/joinEdgeIndel.cu:67-85) works on states shifted by +1 and accumulates
 * in fp16, which agrees with this float sum for integer weights summing <= 2048.
 * ------------------------------------------------------------------------- */
float orc_score_indel(const int32_t *a, const int32_t *b, uint32_t ncol, const float *w)
{
    float s = 0.0f;
    for (uint32_t k = 0; k < ncol; ++k)
        if (a[k] == b[k] && a[k] > 0) s += w[(size_t)(a[k] - 1) * ncol + k];
    return s;
}

/* All pairs of countApoFloat as greedyJoinEdge2 fills cacheShared
 * (src/functions.cpp:1906-1915), the diagonal included: S[i*n+j]. */
void orc_score_indel_matrix(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w, float *S)
{
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i; j < n; ++j) {
            const float v = orc_score_indel(mat + (size_t)i * ncol, mat + (size_t)j * ncol, ncol, w);
            S[(size_t)i * n + j] = v;
            S[(size_t)j * n + i] = v;
        }
}

/* ---------------------------------------------------------------------------
 * (a-3) DNA similarity.  cu/radiusNeighborDNA.cu:27-53 packs the low nibble of
 * every byte code (cu/fasta2bin.cu:12-32) and :55-87 counts the non-zero nibbles
 * of the AND; src/functions.cpp:3741-3754 (seqSimilarity) does the same per byte
 * with unit weights.  Padding bytes are 0 (cu/fasta2bin.cu:189-210).
 * ------------------------------------------------------------------------- */
uint32_t orc_score_dna(const uint32_t *a, const uint32_t *b, uint32_t nwRaw)
{
    uint32_t c = 0;
    for (uint32_t k = 0; k < nwRaw; ++k) {
        uint32_t v = a[k] & b[k];
        for (uint32_t j = 0; j < 4; ++j)
            if ((v >> (j * 8)) & 0x0Fu) ++c;
    }
    return c;
}

/* ---------------------------------------------------------------------------
 * (a-7) radius lists.  cu/radiusNeighbor.cu:322-437: query i scores the targets
 * j >= 32*floor(i/32) with its own set bits (:357-374); earlier targets are back-
 * filled from the lists of earlier queries (:388-396), so for those the score
 * is S(j, i).  The list is ascending and includes i itself when S(i,i) >= r.
 * ------------------------------------------------------------------------- */
int orc_radius_bit_rows(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float radius,
                        uint32_t i0, uint32_t i1, uint32_t *counts, uint32_t *idx,
                        uint64_t idx_cap, uint64_t *pos)
{
    int rc = 0;
    for (uint32_t i = i0; i < i1; ++i) {
        uint32_t c = 0;
        for (uint32_t j = 0; j < n; ++j) {
            float s;
            if ((j >> 5) < (i >> 5))
                s = orc_score_bit(mat + (size_t)j * W, mat + (size_t)i * W, W, w);
            else
                s = orc_score_bit(mat + (size_t)i * W, mat + (size_t)j * W, W, w);
            if (s >= radius) {
                if (idx) {
                    if (*pos >= idx_cap) rc = -1;
                    else idx[*pos] = j;
                    ++*pos;
                }
                ++c;
            }
        }
        counts[i - i0] = c;
    }
    return rc;
}

/* cu/radiusNeighborDNA.cu:259-363: same list contract, integer count vs the
 * integer radius (:120). */
int orc_radius_dna_rows(const uint32_t *mat, uint32_t n, uint32_t nwRaw, uint32_t radius,
                        uint32_t i0, uint32_t i1, uint32_t *counts, uint32_t *idx,
                        uint64_t idx_cap, uint64_t *pos)
{
    int rc = 0;
    for (uint32_t i = i0; i < i1; ++i) {
        uint32_t c = 0;
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t s = orc_score_dna(mat + (size_t)i * nwRaw, mat + (size_t)j * nwRaw, nwRaw);
            if (s >= radius) {
                if (idx) {
                    if (*pos >= idx_cap) rc = -1;
                    else idx[*pos] = j;
                    ++*pos;
                }
                ++c;
            }
        }
        counts[i - i0] = c;
    }
    return rc;
}

/* (a-8) cu/tryRadius.cu:26-50,129-157: count over all targets, self included. */
void orc_try_radius_bit(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, float radius,
                        const uint32_t *q, uint32_t R, uint32_t *counts)
{
    for (uint32_t k = 0; k < R; ++k) {
        uint32_t c = 0;
        const uint32_t *qr = mat + (size_t)q[k] * W;
        for (uint32_t j = 0; j < n; ++j)
            if (orc_score_bit(qr, mat + (size_t)j * W, W, w) >= radius) ++c;
        counts[k] = c;
    }
}

/* cu/tryRadiusDNA.cu:45-74,140-148. */
void orc_try_radius_dna(const uint32_t *mat, uint32_t n, uint32_t nwRaw, uint32_t radius,
                        const uint32_t *q, uint32_t R, uint32_t *counts)
{
    for (uint32_t k = 0; k < R; ++k) {
        uint32_t c = 0;
        const uint32_t *qr = mat + (size_t)q[k] * nwRaw;
        for (uint32_t j = 0; j < n; ++j)
            if (orc_score_dna(qr, mat + (size_t)j * nwRaw, nwRaw) >= radius) ++c;
        counts[k] = c;
    }
}

/* ---------------------------------------------------------------------------
 * (a-5) ancestral rules.
 * Binary type 0: src/functions.cpp:3199-3235 (calcAncestralBitReverse);
 * type 1 (the default passed by R/CasPhy_functions.R:607): :3262-3312
 * (calcAncestralBitReverse2); GPU twins cu/joinEdgeBitReverse.cu:77-113.
 * ------------------------------------------------------------------------- */
static void anc_bit(int type, uint32_t s1, uint32_t b1, uint32_t s2, uint32_t b2,
                    uint32_t *s, uint32_t *b)
{
    if (type == 0) {
        uint32_t diff = s1 ^ s2;
        uint32_t conflict = diff & (b1 | b2);
        *s = (s1 | s2) & ~conflict;
        *b = (diff & ~(b1 ^ b2)) | (b1 & b2);
    } else {
        *s = (s1 & s2) | (b1 & s2) | (b2 & s1);
        *b = (s1 ^ s2) & ~(b1 | b2);
    }
}

/* Indel rule: src/functions.cpp:1956-1974 (missing -> other; both observed and
 * different -> 0; else keep), with the size override of :1959-1962. */
static void anc_indel(const int32_t *a, const int32_t *b, uint32_t ncol, uint32_t sa, uint32_t sb,
                      uint32_t min_size, float min_ratio, int32_t *out)
{
    if (sa < min_size && (float)sa * min_ratio < (float)sb) {
        memcpy(out, b, sizeof(int32_t) * ncol);
    } else if (sb < min_size && (float)sb * min_ratio < (float)sa) {
        memcpy(out, a, sizeof(int32_t) * ncol);
    } else {
        for (uint32_t k = 0; k < ncol; ++k) {
            int32_t t1 = a[k], t2 = b[k];
            if (t1 == -1) t1 = t2;
            else if (t2 >= 0 && t1 != t2) t1 = 0;
            out[k] = t1;
        }
    }
}

/* ---------------------------------------------------------------------------
 * (a-4) join driver, CPU-twin form.  src/functions.cpp:3316-3461
 * (greedyJoinEdgeBitReverse) and :1880-2017 (greedyJoinEdge2): an n x n score
 * cache indexed by cache slot, colMax over live j > i with strict '<' updates,
 * recompute of colMax when its partner is gone (:3365-3380), and a strict '<'
 * scan for the global maximum (:3382-3386), so ties go to the smallest i and
 * then the smallest j.  The new node takes the slot of the smaller child.
 * A "model" abstracts the two data kinds.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint32_t n, width;       /* rows, words (or sites) per row */
    int kind;                /* 0 = binary, 1 = indel */
    int anc_type;            /* binary ancestral type */
    const float *w;
    uint32_t *seq, *bis;     /* binary slots: n x W each */
    int32_t *st;             /* indel slots: n x ncol */
    uint32_t *size;          /* indel node sizes by node id (2n-1) */
    uint32_t min_size;
    float min_ratio;
} model_t;

static float model_score(const model_t *m, uint32_t query_slot, uint32_t target_slot)
{
    if (m->kind == 0)
        return orc_score_bit(m->seq + (size_t)query_slot * m->width,
                             m->seq + (size_t)target_slot * m->width, m->width, m->w);
    return orc_score_indel(m->st + (size_t)query_slot * m->width,
                           m->st + (size_t)target_slot * m->width, m->width, m->w);
}

/* Writes the state of the node in `slot` into the per-node-id outputs. */
static void model_record(const model_t *m, uint32_t slot, uint32_t node,
                         uint32_t *node_state, uint32_t *node_bistate, int32_t *node_state_i)
{
    size_t W = m->width;
    if (m->kind == 0) {
        for (size_t k = 0; k < W; ++k) {
            uint32_t s = m->seq[slot * W + k], b = m->bis[slot * W + k];
            if (node_state) node_state[node * W + k] = s & ~b;
            if (node_bistate) node_bistate[node * W + k] = b;
        }
    } else if (node_state_i) {
        memcpy(node_state_i + node * W, m->st + slot * W, sizeof(int32_t) * W);
    }
}

/* Joins slot c1 (node j1) with slot c2 (node j2) into slot c1 (node newid). */
static void model_merge(model_t *m, uint32_t c1, uint32_t c2, uint32_t j1, uint32_t j2, uint32_t newid)
{
    size_t W = m->width;
    if (m->kind == 0) {
        for (size_t k = 0; k < W; ++k) {
            uint32_t s, b;
            anc_bit(m->anc_type, m->seq[c1 * W + k], m->bis[c1 * W + k],
                    m->seq[c2 * W + k], m->bis[c2 * W + k], &s, &b);
            m->seq[c1 * W + k] = s;
            m->bis[c1 * W + k] = b;
        }
    } else {
        int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * W);
        anc_indel(m->st + c1 * W, m->st + c2 * W, (uint32_t)W, m->size[j1], m->size[j2],
                  m->min_size, m->min_ratio, tmp);
        memcpy(m->st + c1 * W, tmp, sizeof(int32_t) * W);
        free(tmp);
        m->size[newid] = m->size[j1] + m->size[j2];
    }
}

static int join_cpu_twin(model_t *m, uint32_t *edge, uint32_t *node_state,
                         uint32_t *node_bistate, int32_t *node_state_i)
{
    const uint32_t n = m->n, nnode = 2 * n - 1;
    float *cache = (float *)malloc(sizeof(float) * (size_t)n * n);
    float *colMax = (float *)malloc(sizeof(float) * nnode);
    int32_t *colMaxIdx = (int32_t *)calloc(nnode, sizeof(int32_t));
    int32_t *node2cache = (int32_t *)malloc(sizeof(int32_t) * nnode);
    if (!cache || !colMax || !colMaxIdx || !node2cache) {
        free(cache); free(colMax); free(colMaxIdx); free(node2cache);
        return -1;
    }
    for (uint32_t i = 0; i < nnode; ++i) { colMax[i] = -1.0f; node2cache[i] = -1; }
    for (uint32_t i = 0; i < n; ++i) {
        node2cache[i] = (int32_t)i;
        model_record(m, i, i, node_state, node_bistate, node_state_i);
    }
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j) {
            float s = model_score(m, i, j);
            cache[(size_t)i * n + j] = s;
            if (colMax[i] < s) { colMax[i] = s; colMaxIdx[i] = (int32_t)j; }
        }
    for (uint32_t nodeCount = n; nodeCount < nnode; ++nodeCount) {
        float maxShared = -1.0f;
        uint32_t j1 = 0, j2 = 0;
        for (uint32_t i = 0; i + 1 < nodeCount; ++i) {
            int32_t ci = node2cache[i];
            if (ci == -1) continue;
            if (node2cache[colMaxIdx[i]] == -1) {
                colMax[i] = -1.0f;
                for (uint32_t j = i + 1; j < nodeCount; ++j) {
                    int32_t cj = node2cache[j];
                    if (cj != -1) {
                        float s = cache[(size_t)ci * n + cj];
                        if (colMax[i] < s) { colMax[i] = s; colMaxIdx[i] = (int32_t)j; }
                    }
                }
            }
            if (maxShared < colMax[i]) {
                maxShared = colMax[i];
                j1 = i;
                j2 = (uint32_t)colMaxIdx[i];
            }
        }
        uint32_t k = nodeCount - n;
        edge[4 * k + 0] = nodeCount; edge[4 * k + 1] = j1;
        edge[4 * k + 2] = nodeCount; edge[4 * k + 3] = j2;
        uint32_t c1 = (uint32_t)node2cache[j1], c2 = (uint32_t)node2cache[j2];
        model_merge(m, c1, c2, j1, j2, nodeCount);
        node2cache[nodeCount] = (int32_t)c1;
        node2cache[j1] = -1;
        node2cache[j2] = -1;
        model_record(m, c1, nodeCount, node_state, node_bistate, node_state_i);
        for (uint32_t i = 0; i < nodeCount; ++i) {
            int32_t t = node2cache[i];
            if (t == -1) continue;
            /* src/functions.cpp:3429-3437 score the new state's set bits against
             * the live node; :1995-2003 the same with countApoFloat. */
            float s = model_score(m, c1, (uint32_t)t);
            cache[(size_t)t * n + c1] = s;
            if (colMax[i] < s) { colMax[i] = s; colMaxIdx[i] = (int32_t)nodeCount; }
        }
    }
    free(cache); free(colMax); free(colMaxIdx); free(node2cache);
    return 0;
}

/* ---------------------------------------------------------------------------
 * (a-4) join driver, GPU-executable form.  cu/joinEdgeBitReverse.cu:347-505 and
 * cu/joinEdgeIndel.cu:306-468: per-slot localMax initialised over the later
 * tips (:380-397), a global (value, smallest node id) argmax starting from
 * (-1, 0) (:123-166), a recompute over every live slot when the partner is
 * gone (:407-432), and after each join a strict '>' merge of the new node's
 * scores into localMax (:490-501).  O(n) score memory.
 * ------------------------------------------------------------------------- */
static void argmax_pairs(const float *val, const uint32_t *ids, uint32_t len,
                         float *best_v, uint32_t *best_i)
{
    float bv = -1.0f;
    uint32_t bi = 0;
    for (uint32_t s = 0; s < len; ++s) {
        float v = val[s];
        uint32_t id = ids[s];
        if (v > bv || (v == bv && id < bi)) { bv = v; bi = id; }
    }
    *best_v = bv;
    *best_i = bi;
}

static int join_lazy(model_t *m, uint32_t *edge, uint32_t *node_state,
                     uint32_t *node_bistate, int32_t *node_state_i)
{
    const uint32_t n = m->n, nnode = 2 * n - 1;
    float *shared = (float *)malloc(sizeof(float) * n);
    float *localMax = (float *)malloc(sizeof(float) * n);
    uint32_t *localMaxIdx = (uint32_t *)calloc(n, sizeof(uint32_t));
    uint32_t *node2cache = (uint32_t *)malloc(sizeof(uint32_t) * nnode);
    uint32_t *cache2node = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *valid = (uint32_t *)malloc(sizeof(uint32_t) * n);
    if (!shared || !localMax || !localMaxIdx || !node2cache || !cache2node || !valid) {
        free(shared); free(localMax); free(localMaxIdx); free(node2cache); free(cache2node); free(valid);
        return -1;
    }
    uint32_t nValid = n;
    for (uint32_t i = 0; i < nnode; ++i) node2cache[i] = n;
    for (uint32_t i = 0; i < n; ++i) {
        node2cache[i] = i; cache2node[i] = i; valid[i] = i;
        model_record(m, i, i, node_state, node_bistate, node_state_i);
    }
    for (uint32_t i = 0; i + 1 < n; ++i) {
        for (uint32_t j = i + 1; j < n; ++j) shared[j] = model_score(m, i, j);
        argmax_pairs(shared + i + 1, cache2node + i + 1, n - i - 1, &localMax[i], &localMaxIdx[i]);
    }
    localMax[n - 1] = -1.0f;
    for (uint32_t nodeCount = n; nodeCount < nnode; ++nodeCount) {
        uint32_t j1, j2, c1, c2;
        for (;;) {
            float v;
            argmax_pairs(localMax, cache2node, n, &v, &j1);
            c1 = node2cache[j1];
            j2 = localMaxIdx[c1];
            c2 = node2cache[j2];
            if (c2 < n) break;
            for (uint32_t s = 0; s < n; ++s) shared[s] = -1.0f;
            for (uint32_t t = 0; t < nValid; ++t)
                if (valid[t] != c1) shared[valid[t]] = model_score(m, c1, valid[t]);
            argmax_pairs(shared, cache2node, n, &localMax[c1], &localMaxIdx[c1]);
        }
        uint32_t k = nodeCount - n;
        edge[4 * k + 0] = nodeCount; edge[4 * k + 1] = j1;
        edge[4 * k + 2] = nodeCount; edge[4 * k + 3] = j2;
        model_merge(m, c1, c2, j1, j2, nodeCount);
        model_record(m, c1, nodeCount, node_state, node_bistate, node_state_i);
        node2cache[j1] = n; node2cache[j2] = n; node2cache[nodeCount] = c1;
        cache2node[c1] = nodeCount; cache2node[c2] = nnode;
        /* remove c2 from the ascending live-slot list (:472-480) */
        uint32_t p = 0;
        while (valid[p] < c2) ++p;
        --nValid;
        for (; p < nValid; ++p) valid[p] = valid[p + 1];
        for (uint32_t t = 0; t < nValid; ++t) {
            uint32_t s = valid[t];
            if (s == c1) continue;
            float sc = model_score(m, c1, s);
            if (sc > localMax[s]) { localMax[s] = sc; localMaxIdx[s] = nodeCount; }
        }
        localMax[c1] = -1.0f;
        localMax[c2] = -1.0f;
    }
    free(shared); free(localMax); free(localMaxIdx); free(node2cache); free(cache2node); free(valid);
    return 0;
}

static int bit_model(model_t *m, const uint32_t *mat, uint32_t n, uint32_t W, const float *w, int type)
{
    memset(m, 0, sizeof(*m));
    m->n = n; m->width = W; m->kind = 0; m->anc_type = type; m->w = w;
    m->seq = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * W);
    m->bis = (uint32_t *)calloc((size_t)n * W, sizeof(uint32_t));
    if (!m->seq || !m->bis) return -1;
    memcpy(m->seq, mat, sizeof(uint32_t) * (size_t)n * W);
    return 0;
}

static int indel_model(model_t *m, const int32_t *mat, uint32_t n, uint32_t ncol, const float *w,
                       const uint32_t *tip_size, uint32_t min_size, float min_ratio)
{
    memset(m, 0, sizeof(*m));
    m->n = n; m->width = ncol; m->kind = 1; m->w = w;
    m->min_size = min_size; m->min_ratio = min_ratio;
    m->st = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * ncol);
    m->size = (uint32_t *)calloc(2 * (size_t)n, sizeof(uint32_t));
    if (!m->st || !m->size) return -1;
    memcpy(m->st, mat, sizeof(int32_t) * (size_t)n * ncol);
    for (uint32_t i = 0; i < n; ++i) m->size[i] = tip_size ? tip_size[i] : 1u;
    return 0;
}

static void model_free(model_t *m)
{
    free(m->seq); free(m->bis); free(m->st); free(m->size);
}

/* greedyJoinEdge2 row 2 (src/functions.cpp:1978-1985): per edge, the sites
 * where the parent state is 0 and the child state is > 0. */
static void indel_branch(const uint32_t *edge, uint32_t n, uint32_t ncol, const int32_t *ns, uint32_t *branch)
{
    for (uint32_t e = 0; e < 2 * n - 2; ++e) {
        const int32_t *p = ns + (size_t)edge[2 * e] * ncol, *c = ns + (size_t)edge[2 * e + 1] * ncol;
        uint32_t cnt = 0;
        for (uint32_t k = 0; k < ncol; ++k)
            if (p[k] == 0 && c[k] > 0) ++cnt;
        branch[e] = cnt;
    }
}

int orc_join_bit(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, int ancestral_type,
                 uint32_t *edge, uint32_t *node_state, uint32_t *node_bistate)
{
    if (n < 2) return -2;
    model_t m;
    int rc = bit_model(&m, mat, n, W, w, ancestral_type);
    if (rc == 0) rc = join_cpu_twin(&m, edge, node_state, node_bistate, NULL);
    model_free(&m);
    return rc;
}

int orc_join_bit_lazy(const uint32_t *mat, uint32_t n, uint32_t W, const float *w, int ancestral_type,
                      uint32_t *edge, uint32_t *node_state, uint32_t *node_bistate)
{
    if (n < 2) return -2;
    model_t m;
    int rc = bit_model(&m, mat, n, W, w, ancestral_type);
    if (rc == 0) rc = join_lazy(&m, edge, node_state, node_bistate, NULL);
    model_free(&m);
    return rc;
}

int orc_join_indel(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w,
                   const uint32_t *tip_size, uint32_t min_size, float min_ratio,
                   uint32_t *edge, int32_t *node_state, uint32_t *branch)
{
    if (n < 2) return -2;
    model_t m;
    int rc = indel_model(&m, mat, n, ncol, w, tip_size, min_size, min_ratio);
    if (rc == 0) rc = join_cpu_twin(&m, edge, NULL, NULL, node_state);
    if (rc == 0 && branch) indel_branch(edge, n, ncol, node_state, branch);
    model_free(&m);
    return rc;
}

int orc_join_indel_lazy(const int32_t *mat, uint32_t n, uint32_t ncol, const float *w,
                        const uint32_t *tip_size, uint32_t min_size, float min_ratio,
                        uint32_t *edge, int32_t *node_state, uint32_t *branch)
{
    if (n < 2) return -2;
    model_t m;
    int rc = indel_model(&m, mat, n, ncol, w, tip_size, min_size, min_ratio);
    if (rc == 0) rc = join_lazy(&m, edge, NULL, NULL, node_state);
    if (rc == 0 && branch) indel_branch(edge, n, ncol, node_state, branch);
    model_free(&m);
    return rc;
}

/* ---------------------------------------------------------------------------
 * (a-6) cladewise emission.  cu/joinEdgeBitReverse.cu:526-597 (and the same at
 * cu/joinEdgeIndel.cu:486-558): depth-first from the root 2n-2, always taking
 * the lowest unused edge row whose parent is on top of the stack
 * (findValueKernel :168-206); tips keep their ids, the root becomes n and each
 * internal node takes the next id on first discovery.  This equals
 * reOrderEdge2 (src/functions.cpp:2234-2260) plus the factor relabel of
 * R/CasPhy_functions.R:1632-1641.
 * ------------------------------------------------------------------------- */
void orc_cladewise(const uint32_t *edge_raw, uint32_t n, uint32_t *edge_out, uint32_t *internal_order)
{
    const uint32_t nedge = 2 * n - 2, nnode = 2 * n - 1;
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * (nedge ? nedge : 1));
    uint32_t *relabel = (uint32_t *)malloc(sizeof(uint32_t) * nnode);
    uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (nedge + 1));
    for (uint32_t e = 0; e < nedge; ++e) par[e] = edge_raw[2 * e];
    for (uint32_t i = 0; i < nnode; ++i) relabel[i] = i < n ? i : nnode;
    uint32_t next = n, out = 0;
    int sp = 0;
    stack[0] = nedge;
    while (sp > -1) {
        uint32_t found = nedge + 1;
        for (uint32_t e = 0; e < nedge; ++e)
            if (par[e] == stack[sp]) { found = e; break; }
        if (found >= nedge) { --sp; continue; }
        par[found] = nedge + 1;
        uint32_t p = edge_raw[2 * found], c = edge_raw[2 * found + 1];
        stack[++sp] = c;
        if (relabel[p] == nnode) {
            internal_order[next - n] = p;
            relabel[p] = next++;
        }
        if (relabel[c] == nnode) {
            internal_order[next - n] = c;
            relabel[c] = next++;
        }
        edge_out[2 * out] = relabel[p];
        edge_out[2 * out + 1] = relabel[c];
        ++out;
    }
    free(par); free(relabel); free(stack);
}

/* ---------------------------------------------------------------------------
 * (a-9) transfer distance.  cu/TransferDisStream.cu:22-121: H = sum popc(x^y),
 * Hc = sum popc(x^~y) - (32*ncol - ntip) in u32 arithmetic (so Hc = ntip - H
 * modulo 2^32), d = min(H, Hc); xmin/ymin start at 32*ncol (:182-183) and take
 * the minimum over the other set; sizes are min(p, ntip - p) (:107-121).
 * Bitsets must carry zero padding bits beyond ntip: with non-zero padding the
 * reference's result depends on its 8x8 tiling (:39-42,66-71).
 * ------------------------------------------------------------------------- */
void orc_transfer(const uint32_t *Y, uint32_t ny, const uint32_t *X, uint32_t nx,
                  uint32_t ncol, uint32_t ntip,
                  uint32_t *ymin, uint32_t *ysize, uint32_t *xmin, uint32_t *xsize)
{
    const uint32_t init = ncol * 32u;
    for (uint32_t y = 0; y < ny; ++y) ymin[y] = init;
    for (uint32_t x = 0; x < nx; ++x) xmin[x] = init;
    for (uint32_t y = 0; y < ny; ++y) {
        const uint32_t *yr = Y + (size_t)y * ncol;
        for (uint32_t x = 0; x < nx; ++x) {
            const uint32_t *xr = X + (size_t)x * ncol;
            uint32_t h = 0;
            for (uint32_t k = 0; k < ncol; ++k) h += (uint32_t)__builtin_popcount(xr[k] ^ yr[k]);
            uint32_t hc = ntip - h;
            uint32_t d = h < hc ? h : hc;
            if (d < xmin[x]) xmin[x] = d;
            if (d < ymin[y]) ymin[y] = d;
        }
    }
    for (uint32_t y = 0; y < ny; ++y) {
        uint32_t p = 0;
        for (uint32_t k = 0; k < ncol; ++k) p += (uint32_t)__builtin_popcount(Y[(size_t)y * ncol + k]);
        ysize[y] = p > ntip - p ? ntip - p : p;
    }
    for (uint32_t x = 0; x < nx; ++x) {
        uint32_t p = 0;
        for (uint32_t k = 0; k < ncol; ++k) p += (uint32_t)__builtin_popcount(X[(size_t)x * ncol + k]);
        xsize[x] = p > ntip - p ? ntip - p : p;
    }
}

/* ---- (f-1) tipMerge: clusters of tips sharing most of their neighbours ----- */

/* |A & B| of two ascending id lists: the count the reference's warp kernel
 * sharedIdxKernel (cu/tipMerge.cu:25-129) forms by intersecting 32-wide
 * windows of both lists; for ascending lists without repeats that is the
 * size of the set intersection, taken here by one merge pass. */
uint32_t orc_list_shared(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb)
{
    uint32_t i = 0, j = 0, c = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) {
            ++i;
        } else if (a[i] > b[j]) {
            ++j;
        } else {
            ++c;
            ++i;
            ++j;
        }
    }
    return c;
}

/* transitionRoot (cu/tipMerge.cu:130-141): follow category links to a root;
 * -1 stays -1.  (The reference's host loop also calls it with a tip index in
 * place of a category, cu/tipMerge.cu:331; that call is kept as it is.) */
static int32_t tm_root(const int32_t *transition, int32_t x)
{
    if (x == -1) return -1;
    while (transition[x] != x && transition[x] != -1) x = transition[x];
    return x;
}

/* The host loop of cu/tipMerge.cu main (:143-461): tips are visited in
 * windows of 131,072 (CACHEMAX, :8); for window [c0, c1) every tip j < c1
 * takes, in list order, its neighbours t with max(j + 1, c0) <= t < c1 whose
 * id, read as a category, does not lead to j's current root (:323-338, with
 * the state before j's merges); each such pair is merged when the shared
 * count exceeds p times either list length (:379-405), with the reference's
 * category bookkeeping; unassigned tips are numbered after the categories
 * (:443-451).  Lists longer than 1,048,576 ids (MAXNEIGHBOR, :7, :183-186)
 * are refused (return 1).  offsets[ntip + 1] delimit the lists in idx. */
int orc_tip_merge(const uint32_t *nnb, const uint64_t *offsets, const uint32_t *idx, uint32_t ntip, double p,
                  int32_t *member_out)
{
    const uint32_t CHUNK = 1024u * 128u, MAXNB = 1024u * 1024u;
    uint32_t maxnb = 0;
    for (uint32_t i = 0; i < ntip; ++i) {
        if (nnb[i] > MAXNB) return 1;
        if (nnb[i] > maxnb) maxnb = nnb[i];
    }
    int32_t *member = (int32_t *)malloc(sizeof(int32_t) * (ntip ? ntip : 1));
    int32_t *transition = (int32_t *)malloc(sizeof(int32_t) * (ntip ? ntip : 1));
    uint32_t *cursor = (uint32_t *)malloc(sizeof(uint32_t) * (ntip ? ntip : 1));
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (maxnb ? maxnb : 1));
    for (uint32_t i = 0; i < ntip; ++i) {
        member[i] = -1;
        transition[i] = -1;
        const uint32_t *l = idx + offsets[i];
        uint32_t k = 0;
        while (k < nnb[i] && l[k] <= i) ++k;  /* searchOffset, :204-210 */
        cursor[i] = k;
    }
    int32_t ncat = 0;
    for (uint32_t c0 = 0; c0 < ntip; c0 += CHUNK) {
        const uint32_t c1 = ntip - c0 < CHUNK ? ntip : c0 + CHUNK;
        for (uint32_t j = 0; j < c1; ++j) {
            const uint32_t *lj = idx + offsets[j];
            int32_t r1 = member[j] == -1 ? -1 : tm_root(transition, member[j]);
            uint32_t nc = 0;
            for (uint32_t k = cursor[j]; k < nnb[j]; ++k) {
                const uint32_t t = lj[k];
                if (t >= c1) {
                    cursor[j] = k;
                    break;
                }
                if (t >= c0 && (r1 == -1 || r1 != tm_root(transition, (int32_t)t))) cand[nc++] = t;
            }
            for (uint32_t c = 0; c < nc; ++c) {
                const uint32_t t = cand[c];
                const uint32_t s = orc_list_shared(lj, nnb[j], idx + offsets[t], nnb[t]);
                if ((double)s > nnb[j] * p || (double)s > nnb[t] * p) {
                    const int32_t r2 = tm_root(transition, member[t]);
                    if (r1 == -1) {
                        if (r2 == -1) {
                            transition[ncat] = ncat;
                            member[j] = ncat;
                            member[t] = ncat;
                            r1 = ncat;
                            ++ncat;
                        } else {
                            r1 = r2;
                            member[j] = r2;
                        }
                    } else if (r2 == -1) {
                        member[t] = r1;
                    } else {
                        transition[r1] = r2;
                        member[j] = r2;
                        r1 = r2;
                    }
                }
            }
        }
    }
    for (uint32_t i = 0; i < ntip; ++i)
        member_out[i] = member[i] == -1 ? ncat++ : tm_root(transition, member[i]);
    free(member);
    free(transition);
    free(cursor);
    free(cand);
    return 0;
}

/* ---- (f-3) nodeKeep: internal nodes kept against the outgroup rows ------------ */

/* Mutations of a node's state row that an outgroup row lacks:
 * bit form (cu/nodeKeepBit.cu:31-47): set bits of the node whose bit is clear
 * in the outgroup row; indel form (cu/nodeKeepIndel.cu:32-48): sites where
 * the node's state is above 1, the outgroup's is at least 1 and differs. */
static uint32_t nk_found(int indel, const uint32_t *node, const uint32_t *out, uint32_t width)
{
    uint32_t c = 0;
    for (uint32_t w = 0; w < width; ++w) {
        if (indel)
            c += node[w] > 1u && out[w] != node[w] && out[w] >= 1u;
        else
            c += (uint32_t)__builtin_popcount(node[w] & ~out[w]);
    }
    return c;
}

/* The host loop of nodeKeepBit / nodeKeepIndel (cu/nodeKeepBit.cu:185-214,
 * cu/nodeKeepIndel.cu:133-158): edges in file order; an edge to an internal
 * node d (d >= ntip = nnode + 1) is skipped while the last evaluation kept a
 * node and the edge's ancestor id is not below that node's id; otherwise d is
 * kept when every outgroup row lacks at least `cutoff` of its mutations (no
 * outgroup rows: kept), and a kept d becomes the current node.  kept[] gets
 * the kept ids in order; returns their number. */
uint32_t orc_node_keep(int indel, const uint32_t *edge, uint32_t nedge, const uint32_t *state, uint32_t nnode,
                       uint32_t width, const uint32_t *outmat, uint32_t nout, uint32_t cutoff, uint32_t *kept)
{
    const uint32_t ntip = nnode + 1;
    uint32_t current = ntip - 1, flag = 0, nk = 0;
    for (uint32_t i = 0; i < nedge; ++i) {
        const uint32_t a = edge[2 * i], d = edge[2 * i + 1];
        if (d + 1 > ntip) {
            if (flag == 1 && current <= a) continue;
            const uint32_t *row = state + (size_t)(d - ntip) * width;
            flag = 1;
            for (uint32_t j = 0; j < nout; ++j)
                if (nk_found(indel, row, outmat + (size_t)j * width, width) < cutoff) {
                    flag = 0;
                    break;
                }
            if (flag == 1) {
                current = d;
                kept[nk++] = d;
            }
        }
    }
    return nk;
}

/* ---- (f-2) bipartition producer: clade bitsets of a tree ------------------- */

/* nodeDesBig (src/functions.cpp:1277-1366) with an unbounded row cache: edges
 * (1-based ids, tips 1..ntip, internal nodes ntip+1..ntip+nnode) are taken
 * from the last to the first; a tip child sets its bit in the ancestor's row
 * and adds 1 to its size, an internal child ORs its row into the ancestor's,
 * adds its size and is written out as the next row (so rows come in the order
 * the backward scan completes them; the root is never written).  size_out
 * receives the reference's return vector: size_out[k] = size of internal node
 * nnode - 1 - k (0-based internal index), k < nnode - 1.  Returns the number of
 * rows written. */
uint32_t orc_clade_bitsets(const int32_t *anc, const int32_t *des, uint32_t nedge, uint32_t nnode, uint32_t ntip,
                           uint32_t *rows_out, int32_t *size_out)
{
    const uint32_t nbin = ((ntip - 1) >> 5) + 1;
    uint32_t *rows = (uint32_t *)calloc((size_t)(nnode ? nnode : 1) * nbin, sizeof(uint32_t));
    int32_t *size = (int32_t *)calloc(nnode ? nnode : 1, sizeof(int32_t));
    uint32_t nrow = 0;
    for (uint32_t e = nedge; e-- > 0;) {
        const int32_t a = anc[e] - 1, d = des[e] - 1;
        uint32_t *ra = rows + (size_t)(a - (int32_t)ntip) * nbin;
        if (d < (int32_t)ntip) {
            size[a - (int32_t)ntip] += 1;
            ra[d >> 5] |= 1u << (d & 31);
        } else {
            uint32_t *rd = rows + (size_t)(d - (int32_t)ntip) * nbin;
            size[a - (int32_t)ntip] += size[d - (int32_t)ntip];
            for (uint32_t w = 0; w < nbin; ++w) ra[w] |= rd[w];
            memcpy(rows_out + (size_t)nrow * nbin, rd, (size_t)nbin * 4);
            ++nrow;
        }
    }
    for (uint32_t i = 1; i < nnode; ++i) size_out[nnode - 1 - i] = size[i];
    free(rows);
    free(size);
    return nrow;
}

/* ---- (f-4) voteRoot2: root votes of outgroup samples ------------------------- */

/* Positions (bytes) of a word pair whose AND has a nucleotide bit (low nibble). */
static uint32_t vr_shared(uint32_t a, uint32_t b)
{
    const uint32_t v = a & b;
    uint32_t c = 0;
    for (int k = 0; k < 4; ++k) c += ((v >> (8 * k)) & 0x0Fu) != 0;
    return c;
}

/* cu/voteRoot2.cu:34-50 (inferEdgeStateKernel, launched :229-232 on a zeroed
 * buffer): edge state of node d = state(parent) | state(d) per word for every
 * edge (1-based ids; nodes that are no edge's child keep a zero state);
 * :76-138 (findRootKernel): per sample row, the count of shared positions with
 * every node's edge state, nodes in blocks of 8 reduced by the tree (4, 2, 1)
 * in which the lower slot keeps ties (:125-133; node ids past nnode count 0);
 * :140-169 (rowMaxKernel, 32 threads): lane L keeps the first strictly larger
 * block value over blocks L, L + 32, ... starting from (0, node 0), and the 32
 * lanes reduce by the tree (16, 8, 4, 2, 1) in which the lower lane keeps
 * ties.  root[s] = 0-based node, nshared[s] its count (:309-322). */
void orc_vote_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                   const uint32_t *samples, uint32_t nsample, uint32_t *root, uint32_t *nshared)
{
    uint32_t *es = (uint32_t *)calloc((size_t)(nnode ? nnode : 1) * (nbin ? nbin : 1), sizeof(uint32_t));
    for (uint32_t e = 0; e < nedge; ++e) {
        const uint32_t a = edge[2 * e] - 1, d = edge[2 * e + 1] - 1;
        for (uint32_t w = 0; w < nbin; ++w) es[(size_t)d * nbin + w] = state[(size_t)a * nbin + w] | state[(size_t)d * nbin + w];
    }
    const uint32_t nblk = (nnode + 7) / 8;
    uint32_t *bval = (uint32_t *)malloc(sizeof(uint32_t) * (nblk ? nblk : 1));
    uint32_t *bidx = (uint32_t *)malloc(sizeof(uint32_t) * (nblk ? nblk : 1));
    for (uint32_t s = 0; s < nsample; ++s) {
        const uint32_t *row = samples + (size_t)s * nbin;
        for (uint32_t b = 0; b < nblk; ++b) {
            uint32_t v[8], id[8];
            for (uint32_t x = 0; x < 8; ++x) {
                const uint32_t node = 8 * b + x;
                uint32_t c = 0;
                if (node < nnode)
                    for (uint32_t w = 0; w < nbin; ++w) c += vr_shared(es[(size_t)node * nbin + w], row[w]);
                v[x] = c;
                id[x] = node;
            }
            for (uint32_t h = 4; h > 0; h >>= 1)
                for (uint32_t x = 0; x < h; ++x)
                    if (v[x] < v[x + h]) {
                        v[x] = v[x + h];
                        id[x] = id[x + h];
                    }
            bval[b] = v[0];
            bidx[b] = id[0];
        }
        uint32_t lv[32], li[32];
        for (uint32_t L = 0; L < 32; ++L) {
            lv[L] = 0;
            li[L] = 0;
            for (uint32_t i = L; i < nblk; i += 32)
                if (lv[L] < bval[i]) {
                    lv[L] = bval[i];
                    li[L] = bidx[i];
                }
        }
        for (uint32_t off = 16; off > 0; off >>= 1)
            for (uint32_t L = 0; L + off < 32; ++L)
                if (lv[L] < lv[L + off]) {
                    lv[L] = lv[L + off];
                    li[L] = li[L + off];
                }
        root[s] = li[0];
        nshared[s] = lv[0];
    }
    free(es);
    free(bval);
    free(bidx);
}

/* src/functions.cpp:4317-4410 (findRootAVX2Big; findRoot :4170-4212 is the
 * same): the same edge states and counts, candidates the edge children in
 * edge order, the first strictly larger count kept starting from -1; root[s]
 * = child id + 1 (1-based; 0 with no edge), nshared[s] its count (-1 with no
 * edge). */
void orc_find_root(const uint32_t *state, uint32_t nnode, uint32_t nbin, const uint32_t *edge, uint32_t nedge,
                   const uint32_t *samples, uint32_t nsample, int32_t *root, int32_t *nshared)
{
    uint32_t *es = (uint32_t *)calloc((size_t)(nnode ? nnode : 1) * (nbin ? nbin : 1), sizeof(uint32_t));
    for (uint32_t e = 0; e < nedge; ++e) {
        const uint32_t a = edge[2 * e] - 1, d = edge[2 * e + 1] - 1;
        for (uint32_t w = 0; w < nbin; ++w) es[(size_t)d * nbin + w] = state[(size_t)a * nbin + w] | state[(size_t)d * nbin + w];
    }
    for (uint32_t s = 0; s < nsample; ++s) {
        const uint32_t *row = samples + (size_t)s * nbin;
        int64_t best = -1, node = -1;
        for (uint32_t e = 0; e < nedge; ++e) {
            const uint32_t d = edge[2 * e + 1] - 1;
            int64_t c = 0;
            for (uint32_t w = 0; w < nbin; ++w) c += vr_shared(es[(size_t)d * nbin + w], row[w]);
            if (c > best) {
                best = c;
                node = d;
            }
        }
        root[s] = (int32_t)(node + 1);
        nshared[s] = (int32_t)best;
    }
    free(es);
}

/* ---- (f-4) fasta2bin: variable columns, valid bases, byte codes -------------- */

/* cu/fasta2bin.cu:12-32 (encode_char). */
static uint32_t fa_code(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 0x01;
    case 'T': case 't': case 'U': case 'u': return 0x02;
    case 'C': case 'c': return 0x04;
    case 'G': case 'g': return 0x08;
    case 'W': case 'w': return 0x03;
    case 'M': case 'm': return 0x05;
    case 'R': case 'r': return 0x09;
    case 'Y': case 'y': return 0x06;
    case 'K': case 'k': return 0x0A;
    case 'S': case 's': return 0x0C;
    case 'H': case 'h': return 0x07;
    case 'D': case 'd': return 0x0B;
    case 'V': case 'v': return 0x0D;
    case 'B': case 'b': return 0x0E;
    case 'N': case 'n': return 0x0F;
    default: return 0x10;
    }
}

/* seqs: nseq rows of len characters (row-major).  cu/fasta2bin.cu:35-45
 * (count_valid_bases: characters other than '-', '.', 'N', 'n'), :47-93
 * (find_variable_positions, restated with its early exit: the per-base counts
 * grow one sequence at a time and the column is variable as soon as some base
 * count is above 0 and below the number of sequences seen), :96-118 and
 * :189-210 (kept columns in order, 4 byte codes per word, byte j at bits 8j,
 * zero padding).  out: nseq x ceil(nvar / 4) words; returns that width. */
uint64_t orc_fasta_encode(const unsigned char *seqs, uint32_t nseq, uint64_t len, unsigned char *is_variable,
                          uint32_t *valid, uint32_t *out)
{
    for (uint32_t s = 0; s < nseq; ++s) {
        uint32_t v = 0;
        for (uint64_t p = 0; p < len; ++p) {
            const unsigned char c = seqs[(size_t)s * len + p];
            v += !(c == '-' || c == '.' || c == 'N' || c == 'n');
        }
        valid[s] = v;
    }
    uint64_t nvar = 0;
    for (uint64_t p = 0; p < len; ++p) {
        uint32_t cnt[4] = {0, 0, 0, 0}, seen = 0;
        int var = 0;
        for (uint32_t s = 0; s < nseq && !var; ++s) {
            const uint32_t code = fa_code(seqs[(size_t)s * len + p]);
            for (int b = 0; b < 4; ++b) cnt[b] += (code >> b) & 1u;
            ++seen;
            for (int b = 0; b < 4; ++b)
                if (cnt[b] > 0 && cnt[b] < seen) var = 1;
        }
        is_variable[p] = (unsigned char)var;
        nvar += (uint64_t)var;
    }
    const uint64_t width = (nvar + 3) / 4;
    for (uint32_t s = 0; s < nseq; ++s) {
        uint64_t k = 0;
        for (uint64_t w = 0; w < width; ++w) out[(size_t)s * width + w] = 0;
        for (uint64_t p = 0; p < len; ++p) {
            if (!is_variable[p]) continue;
            out[(size_t)s * width + k / 4] |= fa_code(seqs[(size_t)s * len + p]) << (8 * (k % 4));
            ++k;
        }
    }
    return width;
}
