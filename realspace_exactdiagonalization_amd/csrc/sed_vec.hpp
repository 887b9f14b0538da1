// Device helpers of the Lanczos vector kernels (complex FP32/FP64 vectors with FP64
// accumulation), shared by the per-sector solver and the packed engine.
#pragma once

#include <algorithm>
#include <cstdint>

#include "sed_device.hpp"

namespace sed {

constexpr int kMaxBasis = 64;   // largest Krylov basis (m)
// Second classical Gram-Schmidt pass only when the first one left |w'|^2 <= eta^2 |w|^2
// (Daniel-Gragg-Kaufman-Stewart test, eta = 0.1): a single pass loses at most a factor
// 1/eta of orthogonality relative to the vector precision.
constexpr double kReorthEta2 = 0.01;

// Basis vectors of one Gram-Schmidt pass of step j = *step: lo..j with lo = *lo (0
// without a range pointer).  Two kernel families are launched for each pass and one of
// them works: disp 1 (matrix-core kernels) runs when the range holds at least
// kMfmaMinVectors vectors, disp 2 (row-wise loops) when it holds fewer, disp 0 always.
// A one-tile MFMA pass costs a fixed latency-bound time that a row-wise pass over a few
// vectors undercuts.
constexpr int kMfmaMinVectors = 5;
struct PassRange {
    int j, lo, nv;
};
__device__ __forceinline__ bool pass_range(const int* step, const int* lo, int disp, PassRange& r) {
    r.j = *step;
    r.lo = lo ? *lo : 0;
    r.nv = r.j + 1 - r.lo;
    if (disp == 1) return r.nv >= kMfmaMinVectors;
    if (disp == 2) return r.nv < kMfmaMinVectors;
    return true;
}

// Krylov basis size for a sector of dimension n and nev wanted eigenpairs.
inline int krylov_dim(int64_t n, int nev) {
    int m = std::max(2 * nev + 24, 40);
    return (int)std::min<int64_t>(m, std::min<int64_t>(n - 1, kMaxBasis));
}

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
template <typename T>
__device__ __forceinline__ T from_d2(double2 v);
template <>
__device__ __forceinline__ double2 from_d2<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 from_d2<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

__device__ __forceinline__ double2 wave_sum_d2(double2 v) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, kWave);
        v.y += __shfl_down(v.y, off, kWave);
    }
    return v;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Seeded complex Gaussian entry r, Re and Im ~ N(0, 1/2) (Box-Muller), a function of
// (seed, r) only, so every solver path draws the same start vector.
__device__ __forceinline__ double2 gauss_entry(uint64_t seed, int64_t r) {
    const uint64_t a = splitmix64(seed * 0x100000001b3ull + 2 * (uint64_t)r);
    const uint64_t b = splitmix64(seed * 0x100000001b3ull + 2 * (uint64_t)r + 1);
    const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740993.0);
    const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    const double rad = sqrt(-log(u1));    // sqrt(-2 ln u1) * sqrt(1/2)
    const double th = 6.283185307179586 * u2;
    return make_double2(rad * cos(th), rad * sin(th));
}

}  // namespace sed
