// fasta2bin <input FASTA file> <output prefix>
// Drop-in for cu/fasta2bin.cu: reads an alignment, keeps the variable
// columns and writes <prefix>.bin ([nseq, width] + nseq rows of width words,
// 4 byte codes a word), <prefix>_names.txt ("SequenceName,ValidBaseCount")
// and <prefix>_positions.txt ("Position,IsVariable", Yes / No per column).
// The FASTA text is read as the reference reads it (cu/fasta2bin.cu:121-187);
// the column filter, valid-base counts and codes come from
// casphy_fasta_encode.
#include <stdexcept>

#include "tool_common.hpp"

namespace {

bool blank(char c) { return c == ' ' || c == '\t'; }
bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

// Lines end at '\n'; empty lines are skipped; a '>' line closes the running
// sequence (kept only when non-empty, under the name in force) and sets the
// next name: the text after '>' from its first non-blank character, cut at
// the first blank if that comes later; other lines join the running sequence
// without their whitespace.
void read_fasta(const std::string &text, std::vector<std::string> &names, std::vector<std::string> &seqs)
{
    std::string name, cur;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t end = text.find('\n', pos);
        if (end == std::string::npos) end = text.size();
        const char *l = text.data() + pos;
        const size_t n = end - pos;
        pos = end + 1;
        if (!n) continue;
        if (l[0] == '>') {
            if (!cur.empty()) {
                seqs.push_back(cur);
                names.push_back(name);
                cur.clear();
            }
            name.assign(l + 1, n - 1);
            size_t first = std::string::npos, bl = std::string::npos;
            for (size_t i = 0; i < name.size(); ++i) {
                if (first == std::string::npos && !blank(name[i])) first = i;
                if (bl == std::string::npos && blank(name[i])) bl = i;
            }
            if (first != std::string::npos)
                name = (bl == std::string::npos || bl < first) ? name.substr(first) : name.substr(first, bl - first);
        } else {
            for (size_t i = 0; i < n; ++i)
                if (!space(l[i])) cur.push_back(l[i]);
        }
    }
    if (!cur.empty()) {
        seqs.push_back(cur);
        names.push_back(name);
    }
    if (seqs.empty()) throw std::runtime_error("No valid sequences found in file");
    for (size_t i = 1; i < seqs.size(); ++i)
        if (seqs[i].size() != seqs[0].size())
            throw std::runtime_error("All sequences must have the same length for variant filtering");
}

}  // namespace

int main(int argc, char **argv)
{
    using namespace tool;
    if (argc != 3) {
        std::printf("Usage: %s <input FASTA file> <output prefix>\n", argv[0]);
        return 1;
    }
    Timer t;
    const std::string input = argv[1], prefix = argv[2];
    try {
        std::string text;
        {
            FILE *f = std::fopen(input.c_str(), "rb");
            if (!f) throw std::runtime_error("Cannot open file: " + input);
            char buf[1 << 16];
            size_t got;
            while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
            std::fclose(f);
        }
        std::vector<std::string> names, seqs;
        read_fasta(text, names, seqs);
        text.clear();
        text.shrink_to_fit();
        const uint32_t nseq = (uint32_t)seqs.size();
        const uint64_t len = seqs[0].size();
        std::vector<uint8_t> mat((size_t)nseq * len), isv(len ? len : 1);
        for (uint32_t i = 0; i < nseq; ++i) {
            std::copy(seqs[i].begin(), seqs[i].end(), mat.begin() + (size_t)i * len);
            std::string().swap(seqs[i]);
        }
        std::vector<uint32_t> valid(nseq), out((size_t)nseq * ((len + 3) / 4) + 1);
        uint64_t width = 0;
        check(casphy_set_device(device_index()), "casphy_set_device");
        check(casphy_fasta_encode(mat.data(), nseq, len, isv.data(), valid.data(), out.data(), &width),
              "casphy_fasta_encode");
        uint64_t kept = 0;
        for (uint64_t p = 0; p < len; ++p) kept += isv[p];
        std::printf("Found %u sequences\nOriginal sequence length: %llu bases\n", nseq, (unsigned long long)len);
        std::printf("Removed %llu non-variable positions\nRetained %llu variable positions\n",
                    (unsigned long long)(len - kept), (unsigned long long)kept);
        {
            const std::string path = prefix + "_positions.txt";
            FILE *o = open_out(path);
            std::string body = "Position,IsVariable\n";
            for (uint64_t p = 0; p < len; ++p) body += std::to_string(p) + (isv[p] ? ",Yes\n" : ",No\n");
            write_all(o, body.data(), body.size(), path);
            if (std::fclose(o) != 0) die("close failed: " + path);
        }
        {
            const std::string path = prefix + "_names.txt";
            FILE *o = open_out(path);
            std::string body = "SequenceName,ValidBaseCount\n";
            for (uint32_t i = 0; i < nseq; ++i) body += names[i] + "," + std::to_string(valid[i]) + "\n";
            write_all(o, body.data(), body.size(), path);
            if (std::fclose(o) != 0) die("close failed: " + path);
        }
        if (width > 0xFFFFFFFFull) throw std::runtime_error("encoded rows beyond 2^32 words");
        write_bin(prefix + ".bin", nseq, (uint32_t)width, out.data(), (size_t)nseq * width);
        std::printf("Conversion completed! (%.2fs)\n", t.s());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
