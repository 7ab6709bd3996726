// calib.hip -- instruction-rate ceilings of this part, measured rather than
// taken from the data sheet: the denominators the benchmark divides achieved
// rates by.  Built into libcasphy_calib.so, loaded only by bench.py; not part
// of the product library.
//
//   casphy_calib_valu: 32-bit integer VALU issue rate.  Every thread runs 8
//     independent chains of x ^= y; y += popcount(x & K) (v_xor_b32,
//     v_and_b32, v_bcnt_u32_b32 with its fused add: 3 VALU instructions per
//     chain step, no memory traffic in the loop), 8 waves per SIMD.
//   casphy_calib_mfma_fp4: v_mfma_scale_f32_32x32x64_f8f6f4 with fp4
//     operands held in registers, 4 independent accumulators per wave,
//     4 waves per SIMD (131,072 ops per instruction).
//   casphy_calib_mfma_fp4_data: the same with operands that change before
//     every instruction, as a tile kernel's do: each MFMA's A and B words are
//     XORed with a fresh mask of the 1.0 bits (0b0010 per nibble; about half
//     of them flip), 2 VALU instructions per MFMA.  Switching operands draw
//     more power than repeated ones, so this is the rate a matrix-core kernel
//     on real data can reach if the part lowers its clock under that load.
// Each reports the kernel's duration from HIP events on the given stream and
// the work it did; the result words are stored so the loops are not removed.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int VALU_CHAINS = 8;
constexpr int VALU_UNROLL = 8;

__global__ __launch_bounds__(256) void k_valu_rate(uint32_t *out, uint32_t iters, uint32_t seed)
{
    uint32_t x[VALU_CHAINS], y[VALU_CHAINS];
#pragma unroll
    for (int k = 0; k < VALU_CHAINS; ++k) {
        x[k] = (threadIdx.x + 977u * k) ^ seed;
        y[k] = blockIdx.x * 131u + k;
    }
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < VALU_UNROLL; ++u) {
#pragma unroll
            for (int k = 0; k < VALU_CHAINS; ++k) {
                x[k] ^= y[k];
                y[k] += __builtin_popcount(x[k] & 0x5A5A5A5Au);
            }
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < VALU_CHAINS; ++k) r += x[k] + y[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

template <bool VARY>
__global__ __launch_bounds__(256, 4) void k_mfma_fp4_rate(float *out, uint32_t iters, uint32_t seed)
{
    const int lane = threadIdx.x & 63;
    v8i a = v8i{(int)(0x22222222u ^ (seed & lane)), 0x02020202, 0x20202020, 0x22002200, 0, 0, 0, 0};
    v8i b = v8i{0x02200220, (int)(0x22222222u >> (lane & 3)), 0x00222200, 0x20022002, 0, 0, 0, 0};
    v16f acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = v16f{};
    uint32_t x = seed * 2654435761u + (uint32_t)lane * 40503u;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (VARY) {
                x = x * 1664525u + 1013904223u;
                a[k] ^= (int)(x & 0x22222222u);
                b[k] ^= (int)((x >> 1) & 0x22222222u);
            }
            acc[k] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[k], 4, 4, 0, 127, 0, 127);
        }
    }
    float r = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < 16; ++e) r += acc[k][e];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

template <class F> int timed_launch(F launch, hipStream_t s, float *ms)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 1;
    }
    int rc = hipEventRecord(e0, s) != hipSuccess;
    if (!rc) {
        launch();
        rc = hipGetLastError() != hipSuccess;
    }
    if (!rc) rc = hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(ms, e0, e1) != hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" {

// out: blocks x 256 u32 of device memory.  *lane_ops = VALU instructions x 64 lanes.
int casphy_calib_valu(uint32_t *out, uint32_t blocks, uint32_t iters, void *stream, float *ms, uint64_t *lane_ops)
{
    if (!out || !ms || !lane_ops || blocks == 0 || iters == 0) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = timed_launch([&] { hipLaunchKernelGGL(k_valu_rate, dim3(blocks), dim3(256), 0, s, out, iters, 7u); },
                                s, ms);
    *lane_ops = (uint64_t)blocks * 256 * iters * VALU_UNROLL * VALU_CHAINS * 3;
    return rc;
}

// out: blocks x 256 floats of device memory.  *ops = 2 x 32 x 32 x 64 per MFMA.
int casphy_calib_mfma_fp4(float *out, uint32_t blocks, uint32_t iters, void *stream, float *ms, uint64_t *ops)
{
    if (!out || !ms || !ops || blocks == 0 || iters == 0) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = timed_launch(
        [&] { hipLaunchKernelGGL(k_mfma_fp4_rate<false>, dim3(blocks), dim3(256), 0, s, out, iters, 5u); }, s, ms);
    *ops = (uint64_t)blocks * 4 /* waves */ * iters * 4 * (2ull * 32 * 32 * 64);
    return rc;
}

// As casphy_calib_mfma_fp4, with operands that change before every MFMA.
int casphy_calib_mfma_fp4_data(float *out, uint32_t blocks, uint32_t iters, void *stream, float *ms, uint64_t *ops)
{
    if (!out || !ms || !ops || blocks == 0 || iters == 0) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = timed_launch(
        [&] { hipLaunchKernelGGL(k_mfma_fp4_rate<true>, dim3(blocks), dim3(256), 0, s, out, iters, 5u); }, s, ms);
    *ops = (uint64_t)blocks * 4 /* waves */ * iters * 4 * (2ull * 32 * 32 * 64);
    return rc;
}

}  // extern "C"
