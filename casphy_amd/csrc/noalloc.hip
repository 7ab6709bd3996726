// noalloc.hip -- tryRadius counts that allocate nothing, for callers that
// cannot call hipMalloc (code under stream capture, tight serving loops).
//
// The same kernels as casphy_try_radii_bit's matrix-core path (padded rows,
// stage-major copies, fp4 expansions, query x cell score tiles, counting
// passes), with every intermediate buffer carved out of one file-scope device
// array of NOALLOC_SCRATCH_BYTES per device.  Inputs and outputs are the
// caller's device buffers; the work is enqueued on the caller's stream and
// nothing waits for it, so the call can be captured into a graph.  Every
// scratch region is fully written before it is read, so nothing needs
// clearing between calls; calls that overlap on one device share the array
// and race.
#include <mutex>

#include "context.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace casphy {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = 128ull << 20;  // per device, held until the process ends
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of the current device's scratch array, looked up once per
// device with that device current.  Host code uses only this address.
unsigned char *scratch_ptr()
{
    int dev = -1;
    CASPHY_HIP(hipGetDevice(&dev));
    CASPHY_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, CASPHY_ENODEV, "no-allocation entry point: device id");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        CASPHY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CASPHY_REQUIRE(p != nullptr, CASPHY_EHIP, "no-allocation entry point: scratch array has no device address");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace
}  // namespace casphy

using namespace casphy;

extern "C" int casphy_try_radii_bit_noalloc(const uint32_t *mat_cellmajor, uint32_t n, uint32_t W,
                                             const int32_t *radii, uint32_t nradii, const uint32_t *q, uint32_t R,
                                             uint32_t *counts, void *stream)
{
    return guarded([&] {
        CASPHY_REQUIRE(mat_cellmajor && radii && q && counts && n > 0 && W > 0 && R > 0, CASPHY_EARG,
                       "casphy_try_radii_bit_noalloc: bad arguments");
        CASPHY_REQUIRE(nradii >= 1 && nradii <= MAX_ROUND_RADII, CASPHY_EARG,
                       "casphy_try_radii_bit_noalloc: 1..64 radii");
        CASPHY_REQUIRE(mfma_enabled(), CASPHY_EARG,
                       "casphy_try_radii_bit_noalloc: runs on the matrix cores (CASPHY_MFMA=0 is set)");
        const u32 ldw = round_up(W, KC_BIT);
        const u32 padN = stage_rows(n), padR = stage_rows(R), stages = ldw / 4;
        const size_t rows_b = align256((size_t)n * ldw * 4), p_b = align256((size_t)padN * ldw * 4),
                     e_b = align256(fp4_bytes(padN, stages)), pq_b = align256((size_t)padR * ldw * 4),
                     eq_b = align256(fp4_bytes(padR, stages)), s_b = align256((size_t)R * n * 4);
        const size_t need = rows_b + p_b + e_b + pq_b + eq_b + s_b;
        CASPHY_REQUIRE(need <= NOALLOC_SCRATCH_BYTES, CASPHY_EARG,
                       "casphy_try_radii_bit_noalloc: these sizes need " + std::to_string(need >> 20) +
                           " MiB of scratch; the no-allocation entry point holds " +
                           std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
        unsigned char *base = scratch_ptr();
        u32 *rows = reinterpret_cast<u32 *>(base);
        u32 *P = reinterpret_cast<u32 *>(base + rows_b);
        u32 *E = reinterpret_cast<u32 *>(base + rows_b + p_b);
        u32 *PQ = reinterpret_cast<u32 *>(base + rows_b + p_b + e_b);
        u32 *EQ = reinterpret_cast<u32 *>(base + rows_b + p_b + e_b + pq_b);
        int32_t *S = reinterpret_cast<int32_t *>(base + rows_b + p_b + e_b + pq_b + eq_b);
        hipStream_t s = static_cast<hipStream_t>(stream);
        launch_pad_rows(mat_cellmajor, W, n, rows, ldw, s);
        launch_to_stage_major(rows, n, ldw, ldw, P, s);
        launch_expand_fp4(P, padN, stages, E, s);
        // query rows gathered by id (ids of n or more give zero rows)
        launch_to_stage_major(rows, R, ldw, ldw, PQ, s, q, n);
        launch_expand_fp4(PQ, padR, stages, EQ, s);
        MfmaArgs m;
        m.PA = PQ; m.PB = P; m.EA = EQ; m.EB = E;
        m.padA = padR; m.padB = padN; m.stages = stages;
        m.nA = R; m.nB = n; m.triangle = 0;
        m.S = S; m.ldS = n;
        launch_mfma_tiles_untuned(E_SCORES, m, s);
        launch_count_radii(S, R, n, n, radii, nradii, counts, s);
        CASPHY_HIP(hipGetLastError());
    });
}
