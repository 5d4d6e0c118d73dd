// Cross-lane reductions of 64-wide wavefronts (DPP moves inside 16-lane rows,
// permutes across rows) shared by the eigensolver kernels.  Every reduction
// adds in a fixed order, so its result does not depend on timing.
#pragma once

#include <hip/hip_runtime.h>

namespace gfn1 {
namespace wave {

// Value of v held by the lane that the DPP control CTRL selects (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// partners: lane ^ 1, lane ^ 2 (quad permutes), 7 - lane within each group of
// eight (opposite bit 2), lane ^ 8 (rotation of a 16-lane row by 8)
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppRor8 = 0x128;

// wavefront sum, identical in every lane (butterfly: each stage adds the same
// two partial sums in every lane)
__device__ __forceinline__ double wsum(double v) {
    v += dpp<kDppXor1>(v);
    v += dpp<kDppXor2>(v);
    v += dpp<kDppHalfMirror>(v);
    v += dpp<kDppRor8>(v);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// deterministic work-group sum; every thread gets the result
__device__ inline double bsum(double v, double* red) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];
    __syncthreads();
    return s;
}

// Transposed wavefront sums: N per-lane values are halved at every exchange
// (each lane keeps the half its partner sends away), so one pass of log2(64)
// exchanges leaves each lane with one finished total.  The first four
// exchanges are DPP moves inside a 16-lane row; only bits 4 and 5 of the lane
// index go through the cross-row permute.
//
// reduce16: lane L receives the total of s[q], q = 8 b2 + 4 b1 + 2 b0 + b3; every
// lane with the same low four bits holds the same total.  s is overwritten.
__device__ __forceinline__ double reduce16(double (&s)[16], int lane) {
    const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4, b3 = lane & 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double send = b2 ? s[q] : s[q + 8];
        s[q] = (b2 ? s[q + 8] : s[q]) + dpp<kDppHalfMirror>(send);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double send = b1 ? s[q] : s[q + 4];
        s[q] = (b1 ? s[q + 4] : s[q]) + dpp<kDppXor2>(send);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double send = b0 ? s[q] : s[q + 2];
        s[q] = (b0 ? s[q + 2] : s[q]) + dpp<kDppXor1>(send);
    }
    const double send = b3 ? s[0] : s[1];
    double t = (b3 ? s[1] : s[0]) + dpp<kDppRor8>(send);
    t += __shfl_xor(t, 16, 64);
    return t + __shfl_xor(t, 32, 64);
}
__device__ __forceinline__ int reduce16_index(int lane) {
    return ((lane >> 2) & 1) * 8 + ((lane >> 1) & 1) * 4 + (lane & 1) * 2 + ((lane >> 3) & 1);
}

}  // namespace wave
}  // namespace gfn1
