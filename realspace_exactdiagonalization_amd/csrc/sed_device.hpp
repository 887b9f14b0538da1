// Device-side helpers shared by the libsed kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sed {

constexpr int kWave = 64;
constexpr uint32_t kNoRow = 0xffffffffu;

template <typename M>
__device__ __forceinline__ int ctz_m(M m);
template <>
__device__ __forceinline__ int ctz_m<uint32_t>(uint32_t m) { return __builtin_ctz(m); }
template <>
__device__ __forceinline__ int ctz_m<uint64_t>(uint64_t m) { return __builtin_ctzll(m); }

template <typename M>
__device__ __forceinline__ int popc_m(M m);
template <>
__device__ __forceinline__ int popc_m<uint32_t>(uint32_t m) { return __builtin_popcount(m); }
template <>
__device__ __forceinline__ int popc_m<uint64_t>(uint64_t m) { return __builtin_popcountll(m); }

// Colex rank sum_j C(b_j, j+1) from a binomial table bin[n * bs + j] in LDS.
template <typename M>
__device__ __forceinline__ uint32_t colex_rank(M m, const uint32_t* bin, int bs) {
    uint32_t r = 0;
    int j = 1;
    while (m) {
        int b = ctz_m<M>(m);
        r += bin[b * bs + j];
        ++j;
        m &= m - 1;
    }
    return r;
}

// Inverse of colex_rank for k-subsets of {0..n-1}: the mask of rank r.
template <typename M>
__device__ __forceinline__ M colex_unrank(uint32_t r, int n, int k, const uint32_t* bin, int bs) {
    M m = 0;
    int b = n - 1;
    for (int j = k; j >= 1; --j) {
        while (bin[b * bs + j] > r) --b;
        m |= (M)1 << b;
        r -= bin[b * bs + j];
        --b;
    }
    return m;
}

// Image of mask s under a site permutation p (0-based targets); parity = number of
// transpositions (mod 2) needed to reorder the created operators, counted as the
// reference's apply_operation_to_mask does (bits already placed above the target).
template <typename M, bool FERMI>
__device__ __forceinline__ M apply_perm(M s, const uint8_t* p, int& parity) {
    M img = 0;
    int par = 0;
    while (s) {
        int i = ctz_m<M>(s);
        s &= s - 1;
        int q = p[i];
        if (FERMI) par ^= popc_m<M>((img >> q) >> 1) & 1;
        img |= (M)1 << q;
    }
    parity = par;
    return img;
}

// complex helpers on float2 / double2
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename R>
struct C2;
template <>
struct C2<double> { using T = double2; };
template <>
struct C2<float> { using T = float2; };

// Logical block of this workgroup.  Blocks are dealt round-robin to the 8 XCDs
// (physical block b runs on XCD b % 8); here XCD x takes the logical blocks in runs of
// `chunk`: chunk = 1 keeps the natural order (all XCDs on neighbouring rows at once,
// sharing the MALL), chunk = ceil(n / 8) gives each XCD one contiguous range (sharing
// its own L2 only).  The grid is a multiple of 8 * chunk; padding blocks get -1.
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks_logical, int64_t chunk) {
    const int64_t b = blockIdx.x;
    const int64_t x = b & 7, s = b >> 3;
    const int64_t lb = (s / chunk) * (8 * chunk) + x * chunk + (s % chunk);
    return lb < nblocks_logical ? lb : -1;
}

}  // namespace sed
