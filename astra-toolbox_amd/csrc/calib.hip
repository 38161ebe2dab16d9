// Gather-rate calibration: the rate at which a CU serves the projectors'
// tap loads when the data is resident in its L1 cache.
//
// The 3D projectors issue one global_load_dwordx4 per bilinear sample, the
// 2D ones one 8-byte load per linear sample, with the 64 lanes of a wave
// reading scattered addresses inside a small patch.  Their limiter is the
// vector memory path of the CU (address processing and the return of the
// lanes' data), not HBM.  This kernel measures that path on this device:
// every wave issues nothing but such loads, with the lane patterns of the
// projectors (below), over a 4 KiB window per workgroup that stays in L1.
#include <hip/hip_runtime.h>

#include "devutil.hpp"
#include "internal.hpp"

namespace amx {

namespace {

constexpr unsigned kWinPairs = 512;  // 8 rows x 64 pairs (4 KiB) per window
constexpr unsigned kWindows = 256;   // windows (1 MiB), workgroup w uses w % kWindows

__device__ __forceinline__ unsigned mix(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Lane patterns (every offset stays inside the workgroup's 8 x 64-pair
// window; the base moves with the iteration, with a period of 16 so that
// no load is invariant across the 4-fold unrolled loop):
// MODE 0 (16-byte, the 3D FP's ray bundle): lane (i, j) = (lane & 15,
//   lane >> 4) reads the pair at row br + floor(2 j / 3), column bc +
//   floor(2 i / 3 + f): 16 x 4 rays two thirds of a texel apart (a cone
//   magnification of 1.5), f in {0, 1/4, 1/2, 3/4}; rows <= 4 + 2 = 6,
//   columns <= 42 + 10 + 1 = 53 (pairs 53 and 54).
// MODE 1 (16-byte, coalesced): 1 KiB per wave, rows (i & 3) * 2, (i & 3) * 2 + 1
//   and a column shift of ((i >> 2) & 3) * 0 (pairs <= 3 * 128 + 127 = 511).
// MODE 2 (8-byte at 4-byte alignment, the 2D projectors' linear fetch):
//   the 64 lanes of a wave read one row, 1.2 texels apart: floats
//   floor(1.2 lane + f) + fb, <= 75 + 1 + 48 = 124 (+1) of the row's 128.
// MODE 3 (16-byte at 4-byte alignment, two 2D taps per load): one row,
//   1.8 texels apart: floats floor(1.8 lane + f) + fb', fb' in {0, 8},
//   <= 113 + 1 + 8 = 122 (+3) of 128.
// MODE 4: MODE 3's addresses rounded down to 8-byte alignment.
template <int MODE>
__global__ __launch_bounds__(256) void gather_calib_kernel(const float2 *__restrict__ buf,
                                                           unsigned iters, float *__restrict__ out) {
    const float *win =
        reinterpret_cast<const float *>(buf + (size_t)(blockIdx.x % kWindows) * kWinPairs);
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = threadIdx.x >> 6;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll 4
    for (unsigned i = 0; i < iters; ++i) {
        const unsigned k = (i + wave * 5u) & 15u;  // waves start at different phases
        const float f = 0.25f * (float)(k & 3u);
        if (MODE == 0) {
            const unsigned br = (k & 1u) * 4u, bc = ((k >> 1) & 7u) * 6u;
            const unsigned row = br + (2u * (lane >> 4)) / 3u;
            const unsigned col = bc + (unsigned)(0.6666667f * (float)(lane & 15u) + f);
            const float4 q = ld4b(win, (row * 64u + col) << 3);
            a0 += q.x;
            a1 += q.y;
            a2 += q.z;
            a3 += q.w;
        } else if (MODE == 1) {
            const unsigned pair = (k & 3u) * 128u + 2u * lane;
            const float4 q = ld4b(win, pair << 3);
            a0 += q.x;
            a1 += q.y;
            a2 += q.z;
            a3 += q.w;
        } else if (MODE == 2) {
            const unsigned row = k & 7u, fb = ((k >> 3) & 1u) * 48u;
            const unsigned fl = fb + (unsigned)(1.2f * (float)lane + f);
            const float2 q = ld2b(win, (row * 128u + fl) << 2);
            a0 += q.x;
            a1 += q.y;
        } else {
            const unsigned row = k & 7u, fb = ((k >> 3) & 1u) * 8u;
            unsigned fl = fb + (unsigned)(1.8f * (float)lane + f);
            if (MODE == 4) fl &= ~1u;
            const float4 q = ld4u(win, (row * 128u + fl) << 2);
            a0 += q.x;
            a1 += q.y;
            a2 += q.z;
            a3 += q.w;
        }
    }
    out[blockIdx.x * 256u + threadIdx.x] = (a0 + a1) + (a2 + a3);
}

}  // namespace

double gather_calibrate(int mode, unsigned iters, hipStream_t s) {
    AMX_REQUIRE(mode >= 0 && mode <= 4,
                "calibrate_gather: mode must be 0 (16-byte ray-bundle taps), 1 (16-byte "
                "coalesced), 2 (8-byte 2D taps), 3 (16-byte 2D tap pairs, 4-byte aligned) or "
                "4 (the same, 8-byte aligned)");
    AMX_REQUIRE(iters >= 64 && iters <= (1u << 20), "calibrate_gather: iterations in [64, 2^20]");
    int dev = 0, cus = 0;
    AMX_HIP(hipGetDevice(&dev));
    AMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = (unsigned)cus * 8u;  // 8 workgroups (32 waves) per CU
    DevArray buf((size_t)kWindows * kWinPairs * sizeof(float2));
    DevArray out((size_t)grid * 256u * sizeof(float));
    AMX_HIP(hipMemsetAsync(buf.as<void>(), 0, buf.bytes(), s));
    hipEvent_t e0, e1;
    AMX_HIP(hipEventCreate(&e0));
    AMX_HIP(hipEventCreate(&e1));
    float best = 0.f;
    for (int rep = 0; rep < 4; ++rep) {  // first launch warms up
        AMX_HIP(hipEventRecord(e0, s));
        if (mode == 0)
            gather_calib_kernel<0><<<grid, 256, 0, s>>>(buf.as<float2>(), iters, out.as<float>());
        else if (mode == 1)
            gather_calib_kernel<1><<<grid, 256, 0, s>>>(buf.as<float2>(), iters, out.as<float>());
        else if (mode == 2)
            gather_calib_kernel<2><<<grid, 256, 0, s>>>(buf.as<float2>(), iters, out.as<float>());
        else if (mode == 3)
            gather_calib_kernel<3><<<grid, 256, 0, s>>>(buf.as<float2>(), iters, out.as<float>());
        else
            gather_calib_kernel<4><<<grid, 256, 0, s>>>(buf.as<float2>(), iters, out.as<float>());
        AMX_LAUNCHED();
        AMX_HIP(hipEventRecord(e1, s));
        AMX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        AMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && (best == 0.f || ms < best)) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    AMX_REQUIRE(best > 0.f, "calibrate_gather: no timing");
    return (double)grid * 256.0 * (double)iters / (best * 1e-3);
}

}  // namespace amx
