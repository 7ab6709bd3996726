// newick.hip -- Newick trees and transfer bootstrap support (host side).
//
// The trees of CasPhy's bootstrap step are Newick text (examples/test.bstrees)
// that R reads with ape's read.tree (ape >= 5.0, CasPhy's DESCRIPTION:10-12)
// before nodeDesBig turns them into clade bitsets.  read.tree numbers the tips
// 1..n in the order their labels appear, the internal nodes n+1.. in the
// order their '(' appears (the root is n+1), and lists the edges in that same
// preorder (cladewise order: an edge when its child is met).  The reader here
// numbers trees the same way, one tree per call, without recursion (bootstrap
// trees can be 100,000 levels deep).  Branch lengths, internal node labels and
// [comments] are read and dropped; labels are taken verbatim (quoted labels
// with '' for a quote).
#include <cmath>
#include <limits>
#include <unordered_map>

#include "newick.hpp"

namespace casphy {

namespace {

bool blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
bool stops(char c) { return c == '(' || c == ')' || c == ',' || c == ':' || c == ';' || c == '[' || blank(c); }

struct Reader {
    const char *t;
    u64 n, i;
    void skip()
    {
        while (i < n) {
            if (blank(t[i])) {
                ++i;
            } else if (t[i] == '[') {
                while (i < n && t[i] != ']') ++i;
                CASPHY_REQUIRE(i < n, CASPHY_EARG, "newick: unterminated [comment]");
                ++i;
            } else {
                break;
            }
        }
    }
    std::string label()
    {
        skip();
        std::string s;
        if (i < n && t[i] == '\'') {
            ++i;
            for (;;) {
                CASPHY_REQUIRE(i < n, CASPHY_EARG, "newick: unterminated quoted label");
                if (t[i] == '\'') {
                    if (i + 1 < n && t[i + 1] == '\'') {
                        s.push_back('\'');
                        i += 2;
                        continue;
                    }
                    ++i;
                    break;
                }
                s.push_back(t[i++]);
            }
        } else {
            while (i < n && !stops(t[i])) s.push_back(t[i++]);
        }
        return s;
    }
    void length()  // ":<number>", dropped
    {
        skip();
        if (i < n && t[i] == ':') {
            ++i;
            skip();
            while (i < n && !stops(t[i])) ++i;
        }
    }
};

}  // namespace

NewickTree read_newick(const char *text, u64 len, u64 &pos, const std::vector<std::string> *order)
{
    Reader r{text, len, pos};
    NewickTree tree;
    r.skip();
    if (r.i >= len) {
        pos = len;
        return tree;
    }
    constexpr u64 TIP = 1ull << 63;           // child is tip (low bits: order of appearance)
    std::vector<u32> open;                     // internal nodes whose ')' is still ahead
    std::vector<std::pair<u32, u64>> edges;    // (internal parent, child)
    std::vector<std::string> tips;
    u32 nint = 0;
    bool expect = true;  // a subtree comes next
    for (;;) {
        r.skip();
        CASPHY_REQUIRE(r.i < len, CASPHY_EARG, "newick: missing ';'");
        const char c = text[r.i];
        if (expect) {
            if (c == '(') {
                if (!open.empty()) edges.push_back({open.back(), nint});
                open.push_back(nint++);
                ++r.i;
                continue;
            }
            tips.push_back(r.label());
            if (!open.empty()) edges.push_back({open.back(), TIP | (tips.size() - 1)});
            r.length();
            expect = false;
            if (open.empty()) {  // a tree of one tip
                r.skip();
                CASPHY_REQUIRE(r.i < len && text[r.i] == ';', CASPHY_EARG, "newick: missing ';'");
                ++r.i;
                break;
            }
        } else if (c == ',') {
            CASPHY_REQUIRE(!open.empty(), CASPHY_EARG, "newick: ',' outside parentheses");
            ++r.i;
            expect = true;
        } else if (c == ')') {
            CASPHY_REQUIRE(!open.empty(), CASPHY_EARG, "newick: unbalanced ')'");
            open.pop_back();
            ++r.i;
            (void)r.label();  // internal node label
            r.length();
        } else if (c == ';') {
            CASPHY_REQUIRE(open.empty(), CASPHY_EARG, "newick: unbalanced '('");
            ++r.i;
            break;
        } else {
            CASPHY_REQUIRE(false, CASPHY_EARG, std::string("newick: unexpected '") + c + "'");
        }
    }
    pos = r.i;
    CASPHY_REQUIRE(tips.size() < (1ull << 30) && nint < (1u << 30), CASPHY_EARG, "newick: tree too large");
    const u32 ntip = (u32)tips.size();
    std::vector<u32> number(ntip);
    if (order) {
        CASPHY_REQUIRE(order->size() == ntip, CASPHY_EARG,
                       "newick: the tree has " + std::to_string(ntip) + " tips, the tip order " +
                           std::to_string(order->size()));
        std::unordered_map<std::string, u32> at;
        at.reserve(ntip * 2);
        for (u32 k = 0; k < ntip; ++k)
            CASPHY_REQUIRE(at.emplace((*order)[k], k).second, CASPHY_EARG,
                           "newick: tip order repeats label '" + (*order)[k] + "'");
        std::vector<unsigned char> seen(ntip, 0);
        for (u32 t = 0; t < ntip; ++t) {
            auto it = at.find(tips[t]);
            CASPHY_REQUIRE(it != at.end(), CASPHY_EARG, "newick: tip '" + tips[t] + "' is not in the tip order");
            CASPHY_REQUIRE(!seen[it->second], CASPHY_EARG, "newick: tip '" + tips[t] + "' appears twice");
            seen[it->second] = 1;
            number[t] = it->second + 1;
        }
        tree.labels = *order;
    } else {
        for (u32 t = 0; t < ntip; ++t) number[t] = t + 1;
        tree.labels = std::move(tips);
    }
    tree.ntip = ntip;
    tree.nnode = nint;
    tree.anc.resize(edges.size());
    tree.des.resize(edges.size());
    for (size_t e = 0; e < edges.size(); ++e) {
        tree.anc[e] = (int32_t)(ntip + 1 + edges[e].first);
        const u64 ch = edges[e].second;
        tree.des[e] = (ch & TIP) ? (int32_t)number[ch & ~TIP] : (int32_t)(ntip + 1 + ch);
    }
    return tree;
}

// Lemoine et al., Nature 556:452 (2018): a branch splitting p tips from the
// others (p the smaller side) has transfer index d = its smallest transfer
// distance to a branch of a replicate, at most p - 1 (a replicate's tip
// branch is always p - 1 away), and TBE = 1 - mean over replicates of
// d / (p - 1); a split with p < 2 has no TBE (NaN).
void tbe_support(const u32 *ymin, u32 nrep, u32 ny, const u32 *ysize, u32 ntip, double *support)
{
    for (u32 y = 0; y < ny; ++y) {
        const u32 sz = ysize[y], p = std::min(sz, ntip - std::min(sz, ntip));
        if (p < 2 || nrep == 0) {
            support[y] = std::numeric_limits<double>::quiet_NaN();
            continue;
        }
        u64 sum = 0;
        for (u32 b = 0; b < nrep; ++b) sum += std::min(ymin[(size_t)b * ny + y], p - 1);
        support[y] = 1.0 - (double)sum / ((double)(p - 1) * (double)nrep);
    }
}

}  // namespace casphy
