// fasta.hpp -- fasta2bin's column filter and byte encoding (declarations).
#pragma once

#include "common.hpp"

namespace casphy {

// seqs: host nseq x len characters (row-major, equal lengths).  Fills
// is_variable[len] (0/1), valid[nseq] and out (nseq x width words, width =
// ceil(kept columns / 4)); returns width.
u64 fasta_encode(const unsigned char *seqs, u32 nseq, u64 len, unsigned char *is_variable, u32 *valid, u32 *out,
                 hipStream_t s);

}  // namespace casphy
