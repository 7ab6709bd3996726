// fasta.hip -- fasta2bin's column filter and byte encoding on the device.
//
// cu/fasta2bin.cu keeps the columns of an alignment where some base is in
// some but not all sequences (a sequence's base set is its IUPAC code:
// A=1, T/U=2, C=4, G=8, ambiguity codes the unions, N all four, anything
// else none), counts each sequence's valid bases (characters other than '-',
// '.', 'N', 'n') and writes the kept columns as byte codes, 4 per word.
// "Some but not all" is a per-column OR / AND of the code nibbles: the column
// is kept iff (OR & ~AND) has a base bit.  One pass over the resident
// characters folds both and the valid counts; a second writes the codes of the
// kept columns.
#include <algorithm>

#include "context.hpp"
#include "fasta.hpp"
#include "kernels.hpp"

namespace casphy {

namespace {

// Code per character (bits 0..4) and a valid-base flag (bit 7).
__constant__ unsigned char c_code[256];

unsigned char host_code(unsigned c)
{
    switch (c | 0x20u) {  // letters only reach the cases below
    case 'a': return 0x01;
    case 't': case 'u': return 0x02;
    case 'c': return 0x04;
    case 'g': return 0x08;
    case 'w': return 0x03;
    case 'm': return 0x05;
    case 'r': return 0x09;
    case 'y': return 0x06;
    case 'k': return 0x0A;
    case 's': return 0x0C;
    case 'h': return 0x07;
    case 'd': return 0x0B;
    case 'v': return 0x0D;
    case 'b': return 0x0E;
    case 'n': return 0x0F;
    default: return 0x10;
    }
}

std::vector<unsigned char> code_table()
{
    std::vector<unsigned char> t(256);
    for (unsigned c = 0; c < 256; ++c) {
        const bool letter = (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z');
        const unsigned char code = letter ? host_code(c) : 0x10;
        const bool valid = !(c == '-' || c == '.' || c == 'N' || c == 'n');
        t[c] = (unsigned char)(code | (valid ? 0x80 : 0));
    }
    return t;
}

constexpr u32 FOLD_BLOCKS = 512;  // row slices of the fold (partials per column)

// Rows of `pitch` bytes (pitch a multiple of 16, columns past len zero).  A
// workgroup is 4 waves over one slice of 64 16-byte chunks; wave k takes rows
// blockIdx.y * 4 + k, stepping 4 * gridDim.y, two loads in flight.  Lanes fold
// the code nibbles of their chunk by OR and AND (4 columns a word) and count
// the row's valid bases in columns < len (wave sum, one atomic per row and
// slice); the 4 waves' folds meet in LDS and leave as one partial per slice.
__global__ __launch_bounds__(256) void k_fasta_fold(const uint4 *seq, u32 nseq, u64 pitch, u64 len, uint4 *por,
                                                    uint4 *pand, u32 *valid)
{
    __shared__ unsigned char code[256];
    __shared__ uint4 so[4][64], sa[4][64];
    code[threadIdx.x] = c_code[threadIdx.x];
    __syncthreads();
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 cpr = pitch / 16, c = (u64)blockIdx.x * 64 + lane, c0 = c * 16;
    const bool live = c < cpr;
    const u32 inr = c0 >= len ? 0u : (len - c0 >= 16 ? 0xFFFFu : (1u << (u32)(len - c0)) - 1u);
    u32 o[4] = {0, 0, 0, 0}, a[4] = {0x0F0F0F0Fu, 0x0F0F0F0Fu, 0x0F0F0F0Fu, 0x0F0F0F0Fu};
    const u32 step = 4 * gridDim.y;
    for (u32 r = blockIdx.y * 4 + wave; r < nseq; r += 2 * step) {
        const u32 r2 = r + step;
        const bool has2 = r2 < nseq;
        uint4 q[2] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
        if (live) {
            q[0] = seq[(u64)r * cpr + c];
            if (has2) q[1] = seq[(u64)r2 * cpr + c];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !has2) break;  // wave-uniform
            const u32 wv[4] = {q[h].x, q[h].y, q[h].z, q[h].w};
            u32 vbits = 0;
            if (live) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    u32 nib = 0;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const u32 e = code[(wv[k] >> (8 * t)) & 0xFFu];
                        nib |= (e & 0x0Fu) << (8 * t);
                        vbits |= (e >> 7) << (4 * k + t);
                    }
                    o[k] |= nib;
                    a[k] &= nib;
                }
            }
            u32 cnt = (u32)__builtin_popcount(vbits & inr);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
            if (lane == 0 && cnt) atomicAdd(&valid[h ? r2 : r], cnt);
        }
    }
    so[wave][lane] = make_uint4(o[0], o[1], o[2], o[3]);
    sa[wave][lane] = make_uint4(a[0], a[1], a[2], a[3]);
    __syncthreads();
    if (wave == 0 && live) {
        uint4 x = so[0][lane], y = sa[0][lane];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            const uint4 u = so[k][lane], v = sa[k][lane];
            x.x |= u.x; x.y |= u.y; x.z |= u.z; x.w |= u.w;
            y.x &= v.x; y.y &= v.y; y.z &= v.z; y.w &= v.w;
        }
        por[(u64)blockIdx.y * cpr + c] = x;
        pand[(u64)blockIdx.y * cpr + c] = y;
    }
}

// flags (4 columns a word, one byte each): 1 where the OR of the slices'
// partials has a base bit their AND lacks.  One wave per word: lanes stride
// over the slices, a butterfly joins them.
__global__ __launch_bounds__(256) void k_fasta_cols(const u32 *por, const u32 *pand, u32 nslice, u64 pw, u32 *flags)
{
    const u32 lane = threadIdx.x & 63;
    for (u64 w = ((u64)blockIdx.x * 256 + threadIdx.x) / 64; w < pw; w += (u64)gridDim.x * 4) {
        u32 x = 0, y = 0x0F0F0F0Fu;
        for (u32 b = lane; b < nslice; b += 64) {
            x |= por[b * pw + w];
            y &= pand[b * pw + w];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            x |= __shfl_xor(x, off);
            y &= __shfl_xor(y, off);
        }
        const u32 v = x & ~y & 0x0F0F0F0Fu;
        u32 f = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) f |= (u32)(((v >> (8 * k)) & 0xFFu) != 0) << (8 * k);
        if (lane == 0) flags[w] = f;
    }
}

// Rows of len bytes back to back -> rows of `pitch` bytes, zero padded (one
// 16-byte chunk per thread; word loads when len is a multiple of 4).
__global__ __launch_bounds__(256) void k_repitch(const unsigned char *raw, u32 nseq, u64 len, u64 pitch, uint4 *dst)
{
    const u64 per = pitch / 16, total = (u64)nseq * per;
    const bool words = (len & 3) == 0;
    for (u64 t = (u64)blockIdx.x * 256 + threadIdx.x; t < total; t += (u64)gridDim.x * 256) {
        const u64 r = t / per, c0 = (t % per) * 16;
        const unsigned char *src = raw + r * len;
        u32 w[4] = {0, 0, 0, 0};
        if (words) {
            const u32 *s4 = reinterpret_cast<const u32 *>(src);
#pragma unroll
            for (u32 k = 0; k < 4; ++k)
                if (c0 + 4 * k < len) w[k] = s4[c0 / 4 + k];
        } else {
#pragma unroll
            for (u32 k = 0; k < 16; ++k)
                if (c0 + k < len) w[k / 4] |= (u32)src[c0 + k] << (8 * (k % 4));
        }
        dst[t] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Row r (workgroups step over rows), word w (threads step over the width):
// the codes of kept columns 4w .. 4w+3 (col4[w]; ~0u past the last one).
__global__ __launch_bounds__(256) void k_fasta_encode(const unsigned char *seq, u32 nseq, u64 pitch, const uint4 *col4,
                                                      u64 width, u32 *out)
{
    __shared__ unsigned char code[256];
    code[threadIdx.x] = c_code[threadIdx.x];
    __syncthreads();
    for (u32 r = blockIdx.x; r < nseq; r += gridDim.x) {
        const unsigned char *row = seq + (u64)r * pitch;
        for (u64 w = threadIdx.x; w < width; w += 256) {
            const uint4 cc = col4[w];
            const u32 cs[4] = {cc.x, cc.y, cc.z, cc.w};
            u32 v = 0;
#pragma unroll
            for (u32 j = 0; j < 4; ++j)
                if (cs[j] != 0xFFFFFFFFu) v |= (u32)(code[row[cs[j]]] & 0x1Fu) << (8 * j);
            out[(u64)r * width + w] = v;
        }
    }
}

}  // namespace

u64 fasta_encode(const unsigned char *seqs, u32 nseq, u64 len, unsigned char *is_variable, u32 *valid, u32 *out,
                 hipStream_t s)
{
    static const std::vector<unsigned char> table = code_table();
    CASPHY_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_code), table.data(), 256, 0, hipMemcpyHostToDevice, s));
    if (!nseq) return 0;
    if (!len) {
        std::fill(valid, valid + nseq, 0u);
        return 0;
    }
    const u64 pitch = (len + 15) / 16 * 16;
    DevBuf dseq((size_t)nseq * pitch), dvalid((size_t)nseq * 4);
    if (pitch == len) {
        h2d(dseq.p, seqs, (size_t)nseq * len, s);
    } else {  // one contiguous upload, rows padded on the device
        DevBuf raw((size_t)nseq * len);
        h2d(raw.p, seqs, (size_t)nseq * len, s);
        const u64 chunks = (u64)nseq * (pitch / 16);
        hipLaunchKernelGGL(k_repitch, dim3((u32)std::min<u64>((chunks + 255) / 256, 65536)), dim3(256), 0, s,
                           raw.as<unsigned char>(), nseq, len, pitch, reinterpret_cast<uint4 *>(dseq.p));
        CASPHY_HIP(hipGetLastError());
    }
    launch_fill_u32(dvalid.as<u32>(), nseq, 0u, s);
    const u64 cpr = pitch / 16, groups = (cpr + 63) / 64;
    CASPHY_REQUIRE(groups <= 0x7FFFFFFFull, CASPHY_EARG, "fasta encode: sequences too long");
    const u32 nslice = std::min<u32>(FOLD_BLOCKS, ceil_div(nseq, 8u));
    DevBuf por((size_t)nslice * pitch), pand((size_t)nslice * pitch), dflags(pitch);
    hipLaunchKernelGGL(k_fasta_fold, dim3((u32)groups, nslice), dim3(256), 0, s,
                       reinterpret_cast<const uint4 *>(dseq.p), nseq, pitch, len, por.as<uint4>(), pand.as<uint4>(),
                       dvalid.as<u32>());
    hipLaunchKernelGGL(k_fasta_cols, dim3((u32)std::min<u64>((pitch / 4 + 3) / 4, 65536)), dim3(256), 0, s,
                       por.as<u32>(), pand.as<u32>(), nslice, pitch / 4, dflags.as<u32>());
    CASPHY_HIP(hipGetLastError());
    std::vector<unsigned char> flags(pitch);
    d2h(flags.data(), dflags.p, pitch, s);
    d2h(valid, dvalid.p, (size_t)nseq * 4, s);
    std::vector<u32> col;
    for (u64 p = 0; p < len; ++p) {
        is_variable[p] = flags[p];
        if (flags[p]) {
            CASPHY_REQUIRE(p < 0xFFFFFFFFull, CASPHY_EARG, "fasta encode: kept column beyond 2^32 - 1");
            col.push_back((u32)p);
        }
    }
    const u64 nvar = col.size(), width = (nvar + 3) / 4;
    if (!width) return 0;
    col.resize(width * 4, 0xFFFFFFFFu);
    DevBuf dcol(width * 16), dout((size_t)nseq * width * 4);
    h2d(dcol.p, col.data(), width * 16, s);
    hipLaunchKernelGGL(k_fasta_encode, dim3(std::min<u32>(nseq, 65536)), dim3(256), 0, s, dseq.as<unsigned char>(),
                       nseq, pitch, dcol.as<uint4>(), width, dout.as<u32>());
    CASPHY_HIP(hipGetLastError());
    d2h(out, dout.p, (size_t)nseq * width * 4, s);
    return width;
}

}  // namespace casphy
