// synth_native.cpp -- native generator of SPR-perturbed bootstrap trees for
// the synthetic inputs (casphy_amd/synth.py): the same moves as synth.spr on a
// rooted binary tree given as a parent array, then R's edge matrix as
// synth.ape_edges numbers it.  Seeded splitmix64 per replicate; plain C++,
// no GPU.
#include <cstdint>
#include <utility>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int64_t below(int64_t n) { return (int64_t)(next() % (uint64_t)n); }  // n << 2^64: bias negligible
};

}  // namespace

extern "C" {

// parent[nn] (root 0, parent[0] = -1, binary), leaves[ntip] (tip t at node
// leaves[t]); `moves` SPR moves with seed; edge_out receives (nn - 1) rows of
// (parent, child) 1-based: tips 1..ntip, internal nodes ntip + 1.. in
// preorder, edges depth first with children in increasing node id.
// Returns the number of internal nodes, or -1 on bad input.
int64_t casphy_synth_spr_edges(const int64_t *parent, int64_t nn, const int64_t *leaves, int64_t ntip, int64_t moves,
                               uint64_t seed, int32_t *edge_out)
{
    if (nn < 3 || ntip < 2 || parent[0] != -1) return -1;
    std::vector<int64_t> par(parent, parent + nn);
    std::vector<int64_t> k0(nn, -1), k1(nn, -1);  // the two children (binary tree)
    for (int64_t v = 1; v < nn; ++v) {
        const int64_t p = par[v];
        if (p < 0 || p >= nn) return -1;
        if (k0[p] < 0) k0[p] = v;
        else if (k1[p] < 0) k1[p] = v;
        else return -1;
    }
    Rng rng{seed};
    const int64_t root = 0;
    int64_t done = 0, guard = 0;
    while (done < moves) {
        if (++guard > 1000 * (moves + 1000)) return -1;
        const int64_t s = 1 + rng.below(nn - 1);
        const int64_t p = par[s];
        if (p == root) continue;
        const int64_t t = rng.below(nn);
        if (t == p || t == root || par[t] == p) continue;
        bool bad = false;
        for (int64_t a = t; a != -1; a = par[a])
            if (a == s) {
                bad = true;
                break;
            }
        if (bad) continue;
        const int64_t g = par[p];
        const int64_t sib = k0[p] == s ? k1[p] : k0[p];
        // sibling hangs from the grandparent
        if (k0[g] == p) k0[g] = sib; else k1[g] = sib;
        par[sib] = g;
        // p reused on edge (par[t], t)
        const int64_t q = par[t];
        if (k0[q] == t) k0[q] = p; else k1[q] = p;
        par[p] = q;
        k0[p] = t;
        k1[p] = s;
        par[t] = p;
        ++done;
    }
    // numbering: tips by leaves, internal nodes in preorder; children in
    // increasing node id
    std::vector<int64_t> label(nn, 0);
    for (int64_t t = 0; t < ntip; ++t) label[leaves[t]] = t + 1;
    int64_t nxt = ntip + 1;
    label[root] = nxt++;
    std::vector<std::pair<int64_t, int> > st;
    st.push_back({root, 0});
    int64_t ne = 0;
    while (!st.empty()) {
        auto &f = st.back();
        const int64_t v = f.first;
        int64_t a = k0[v], b = k1[v];
        if (a > b && b >= 0) std::swap(a, b);
        const int64_t c = f.second == 0 ? a : f.second == 1 ? b : -1;
        if (c < 0) {
            st.pop_back();
            continue;
        }
        ++f.second;
        const bool internal = k0[c] >= 0;
        if (internal) label[c] = nxt++;
        edge_out[2 * ne] = (int32_t)label[v];
        edge_out[2 * ne + 1] = (int32_t)label[c];
        ++ne;
        if (internal) st.push_back({c, 0});
    }
    return nxt - ntip - 1;
}

}  // extern "C"
