// Tall-skinny FP64 MFMA GEMMs of the Lanczos loop (gfx950): the re-orthogonalisation
// of a new vector against the stored Krylov basis and the thick-restart rotation.
//
// Complex vectors (FP32 or FP64 storage) are read as real arrays of interleaved
// (re, im) pairs, converted to double and fed to the FP64 matrix cores:
//   * projections  h = V^H w   -> (2n x (j+2))^T x (2n x 2) GEMM, K = 2n (k_proj_mfma);
//   * update       w - V h     -> (n x 2(j+1)) x (2(j+1) x 2) GEMM  (k_orth_update_mfma);
//   * restart      V S         -> (2n x m) x (m x kk) GEMM          (k_rotate_mfma).
// Every GEMM streams the basis once; the new vector w is read once per pass (a
// row-wise VALU loop reads it once per group of basis vectors).  All sums run in a
// fixed order (fixed grid, fixed row partition, fixed lane layout): bitwise
// reproducible.
//
// Lane layouts, measured on gfx950:
//   v_mfma_f64_4x4x4f64 (4 blocks b): A_b[i][k] in lane 16k + 4b + i,
//     B_b[k][j] in lane 16k + 4b + j, D_b[i][j] in lane 16i + 4b + j;
//   v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
//     D[i][j] in lane j + 16 (i % 4), register i / 4.
#pragma once

#include <algorithm>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "sed_device.hpp"
#include "sed_vec.hpp"

namespace sed {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double mfma4x4(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ dbl4 mfma16x16(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kProjTiles = (kMaxBasis + 2 + 15) / 16;   // 16 vectors per tile
constexpr int kTileRows = 64;                            // complex rows per staged tile
constexpr int kTilePad = kTileRows + 1;                  // LDS row pitch (bank spread)

// Rows of block b for a fixed grid G: [b * chunk, min(n, (b + 1) * chunk)), chunk a
// multiple of the staged tile so no tile straddles two blocks.
__device__ __forceinline__ int64_t proj_chunk(int64_t n, int G) {
    return (((n + G - 1) / G) + kTileRows - 1) / kTileRows * kTileRows;
}

// Coalesced load of vector tile tau (vectors 16 tau + 4 wv + u, u < 4) at rows
// t0 + lane: vector index nv stands for w itself when with_w (the norm column of the
// projection), indices beyond and rows at or beyond r1 read as zero.  The loads are
// unconditional (clamped row; slots without a vector read w, which is cached) and their
// raw values are kept as loaded, with bit u of `keep` saying whether slot u counts:
// nothing consumes them before the tile is staged, so the loads of later tiles stay in
// flight meanwhile.
template <typename T>
__device__ __forceinline__ void load_vec_tile(const T* __restrict__ V, const T* __restrict__ w, int64_t n,
                                              int64_t t0, int64_t r1, int tau, int nv, int with_w, int wv,
                                              int lane, T (&x)[4], int& keep) {
    const int64_t row = t0 + lane;
    const bool live = row < r1;
    const int64_t rc = live ? row : r1 - 1;
    keep = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int vi = 16 * tau + 4 * wv + u;
        const bool isv = vi < nv;
        keep |= (live && (isv || (with_w && vi == nv))) ? 1 << u : 0;
        const T* src = isv ? V + (int64_t)vi * n : w;
        x[u] = src[rc];
    }
}

// Staged value of slot u: the loaded element when it counts, else zero.
template <typename T>
__device__ __forceinline__ double2 kept(const T (&x)[4], int keep, int u) {
    return (keep >> u) & 1 ? to_d2(x[u]) : make_double2(0.0, 0.0);
}

// The (row tile, vector tile) pairs of a block, row tile outer: iteration it covers rows
// r0 + 64 (it / ntile) .. + 63 and vector tile it % ntile.  Two iterations are in flight:
// while iteration it is multiplied out of LDS, the loads of it + 1 (issued one iteration
// earlier) and of it + 2 (issued right after it was staged) are outstanding.
struct TileWalk {
    int64_t r0, r1;
    int ntile, niter;
    __device__ TileWalk(int64_t r0_, int64_t r1_, int ntile_) : r0(r0_), r1(r1_), ntile(ntile_) {
        const int64_t nrow = r1 > r0 ? (r1 - r0 + kTileRows - 1) / kTileRows : 0;
        niter = (int)nrow * ntile;
    }
    __device__ int64_t t0(int it) const { return r0 + (int64_t)(it / ntile) * kTileRows; }
    __device__ int tau(int it) const { return it % ntile; }
};

// part[i * G + b] = sum over block b's rows of conj(V_i[r]) w[r] for lo <= i <= j, and
// part[(j+1) * G + b] = sum |w[r]|^2, j = *step, lo = *lo (pass_range).  w is
// v_{j+1} itself when w_next.  A zero gate skips the pass.
//
// The block walks its rows in tiles of 64; each tile of 16 basis vectors is loaded with
// one coalesced 64-row load per vector (lane = row) and staged in LDS.  Product on the
// 4x4x4 MFMA (4 blocks): D_b[i][c] = sum_k A_b[i][k] B_b[k][c], i = vector
// 16 tau + 4b + i, k = a real component (wave wv takes rows 16 wv .. 16 wv + 15 of the
// tile), c = 0: sum vr wr + vi wi (real part), c = 1: sum vr wi - vi wr (imaginary part);
// columns 2, 3 are zero.
template <typename T>
__global__ void __launch_bounds__(256) k_proj_mfma(const T* __restrict__ V0, int64_t n, const int* __restrict__ step,
                                                   const T* __restrict__ w0, double2* __restrict__ part0,
                                                   const int* __restrict__ gate, const int* __restrict__ lo,
                                                   int w_next, int disp) {
    __shared__ double2 sv[16 * kTilePad];
    __shared__ double2 sw[kTileRows];
    __shared__ double red[256 / kWave][kProjTiles][kWave];
    if (gate && *gate == 0) return;
    PassRange pr;
    if (!pass_range(step, lo, disp, pr)) return;
    const int G = gridDim.x;
    const T* __restrict__ V = V0 + (int64_t)pr.lo * n;   // the range's first vector
    const T* __restrict__ w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    double2* __restrict__ part = part0 + (int64_t)pr.lo * G;
    const int nv = pr.nv;                   // range vectors 0..nv-1, slot nv = w itself
    const int ntile = (nv + 1 + 15) / 16;
    const int64_t chunk = proj_chunk(n, G);
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const TileWalk tw(r0, r0 + chunk < n ? r0 + chunk : n, ntile);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int k = lane >> 4, b = (lane >> 2) & 3, q = lane & 3;
    double acc[kProjTiles];
#pragma unroll
    for (int t = 0; t < kProjTiles; ++t) acc[t] = 0.0;
    auto load = [&](int it, T (&x)[4], int& keep, T& xw) {
        const int64_t t0 = tw.t0(it);
        const int tau = tw.tau(it);
        load_vec_tile(V, w, n, t0, tw.r1, tau, nv, 1, wv, lane, x, keep);
        xw = w[t0 + lane < tw.r1 ? t0 + lane : tw.r1 - 1];
    };
    auto stage = [&](int it, T (&x)[4], int& keep, T& xw) {
        const int tau = tw.tau(it);
        __syncthreads();   // the previous iteration's products are done with the LDS
#pragma unroll
        for (int u = 0; u < 4; ++u) sv[(4 * wv + u) * kTilePad + lane] = kept(x, keep, u);
        if (tau == 0 && wv == 0) sw[lane] = tw.t0(it) + lane < tw.r1 ? to_d2(xw) : make_double2(0.0, 0.0);
        __syncthreads();
        load(it + 2, x, keep, xw);   // past the last tile: a clamped dummy, kept as zero
        double a = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int rr = 16 * wv + 4 * c + k;
            const double2 v = sv[(4 * b + q) * kTilePad + rr];
            const double2 y = sw[rr];
            const double bre = q == 0 ? y.x : (q == 1 ? y.y : 0.0);
            const double bim = q == 0 ? y.y : (q == 1 ? -y.x : 0.0);
            a = mfma4x4(v.x, bre, a);
            a = mfma4x4(v.y, bim, a);
        }
#pragma unroll
        for (int t = 0; t < kProjTiles; ++t) acc[t] += t == tau ? a : 0.0;
    };
    // an even number of iterations (a padded one stages zeros): every load is issued on
    // every path, so the wait before staging a tile leaves the next tile's loads in flight
    const int nit = (tw.niter + 1) & ~1;
    if (nit > 0) {
        T xa[4], xb[4], wa, wb;
        int ka = 0, kb = 0;
        load(0, xa, ka, wa);
        load(1, xb, kb, wb);
        for (int it = 0; it < nit; it += 2) {
            stage(it, xa, ka, wa);
            stage(it + 1, xb, kb, wb);
        }
    }
#pragma unroll
    for (int tau = 0; tau < kProjTiles; ++tau) red[wv][tau][lane] = acc[tau];
    __syncthreads();
    // vector v = 16 tau + 4b + i: real part in lane 16i + 4b, imaginary part in the next lane
    for (int o = threadIdx.x; o < ntile * 16; o += blockDim.x) {
        const int tau = o >> 4, vl = o & 15;
        const int v = 16 * tau + vl;
        if (v > nv) continue;
        const int l0 = 16 * (vl & 3) + 4 * (vl >> 2);
        double re = 0.0, im = 0.0;
        for (int x2 = 0; x2 < 256 / kWave; ++x2) {
            re += red[x2][tau][l0];
            im += red[x2][tau][l0 + 1];
        }
        part[(int64_t)v * G + blockIdx.x] = make_double2(re, im);
    }
}

// out[r] = w[r] - sum_{lo<=i<=j} V_i[r] h[i] with out = w (to_next = 0) or V_{j+1}
// (to_next = 1) and w = v_{j+1} itself when w_next (in place: every row is read before
// the thread that owns it writes it); each block also writes the partial |out|^2 of its
// rows (fixed grid).  A zero gate, or a range that the other kernel family takes, leaves
// everything untouched.
//
// Tiles of 64 rows x 16 vectors are staged through LDS as in k_proj_mfma.  Product on
// the 4x4x4 MFMA (4 blocks): D_b[i][c] over complex rows 16 wv + 4b + c of the tile,
// i = 0: real part, i = 1: imaginary part of sum_i V_i h_i (rows 2, 3 of A are zero);
// for vectors 4s .. 4s + 3 of the tile, lane (k, b, c) takes vector 4s + k: one
// instruction its real part (A = (Re h, Im h)), the next its imaginary part
// (A = (-Im h, Re h)).
template <typename T>
__global__ void __launch_bounds__(256) k_orth_update_mfma(T* V0, int64_t n, const int* __restrict__ step,
                                                          const double2* __restrict__ h, T* w0, int to_next,
                                                          double* __restrict__ normpart,
                                                          const int* __restrict__ gate, const int* __restrict__ lo,
                                                          int w_next, int disp) {
    __shared__ double2 sv[16 * kTilePad];
    __shared__ double2 lh[kProjTiles * 16];
    __shared__ double red[256 / kWave];
    if (gate && *gate == 0) return;
    PassRange pr;
    if (!pass_range(step, lo, disp, pr)) return;
    const int nv = pr.nv;
    const int ntile = (nv + 15) / 16;
    for (int i = threadIdx.x; i < 16 * ntile; i += blockDim.x)
        lh[i] = i < nv ? h[pr.lo + i] : make_double2(0.0, 0.0);
    const T* V = V0 + (int64_t)pr.lo * n;   // the range's first vector
    T* out = to_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    const T* w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int k = lane >> 4, b = (lane >> 2) & 3, q = lane & 3;
    const int G = gridDim.x;
    const int64_t chunk = proj_chunk(n, G);
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
    double nrm = 0.0;
    {
        const TileWalk tw(r0, r1, ntile);
        double acc = 0.0;
        // tile loads plus this wave's 16 rows of w for the row tile's final subtraction;
        // slots without a vector re-read the range's first vector (never w, which may be
        // the output)
        auto load = [&](int it, T (&x)[4], int& keep, T& xe) {
            load_vec_tile(V, V, n, tw.t0(it), r1, tw.tau(it), nv, 0, wv, lane, x, keep);
            const int64_t re = tw.t0(it) + 16 * wv + (lane & 15);
            xe = w[re < r1 ? re : r1 - 1];
        };
        auto stage = [&](int it, T (&x)[4], int& keep, T& xe) {
            const int tau = tw.tau(it);
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) sv[(4 * wv + u) * kTilePad + lane] = kept(x, keep, u);
            const T we = xe;
            __syncthreads();
            load(it + 2, x, keep, xe);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int vl = 4 * s4 + k;
                const double2 y = sv[vl * kTilePad + 16 * wv + 4 * b + q];
                const double2 c = lh[16 * tau + vl];
                const double are = q == 0 ? c.x : (q == 1 ? c.y : 0.0);
                const double aim = q == 0 ? -c.y : (q == 1 ? c.x : 0.0);
                acc = mfma4x4(are, y.x, acc);
                acc = mfma4x4(aim, y.y, acc);
            }
            if (tau == ntile - 1) {
                // lanes 0..15 hold the real parts of rows t0 + 16 wv + lane, lanes 16..31
                // the imaginary parts
                const double im = __shfl_down(acc, 16, kWave);
                if (lane < 16) {
                    const int64_t r = tw.t0(it) + 16 * wv + lane;
                    if (r < r1) {
                        double2 a = to_d2(we);
                        a.x -= acc;
                        a.y -= im;
                        const T o = from_d2<T>(a);
                        out[r] = o;
                        const double2 od = to_d2(o);
                        nrm += od.x * od.x + od.y * od.y;
                    }
                }
                acc = 0.0;
            }
        };
        const int nit = (tw.niter + 1) & ~1;   // even, as in k_proj_mfma
        if (nit > 0) {
            T xa[4], xb[4], ea, eb;
            int ka = 0, kb = 0;
            load(0, xa, ka, ea);
            load(1, xb, kb, eb);
            for (int it = 0; it < nit; it += 2) {
                stage(it, xa, ka, ea);
                stage(it + 1, xb, kb, eb);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, kWave);
    __syncthreads();
    if (lane == 0) red[wv] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int x2 = 0; x2 < 256 / kWave; ++x2) sm += red[x2];
        normpart[blockIdx.x] = sm;
    }
}

// In place, V_i[r] <- sum_{j<m} V_j[r] S[j * kk + i] for i < kk (S real, m <= kMaxBasis).
//
// 16x16x4 MFMA over slabs of 16 complex rows, U slabs per wave and iteration.  Lane
// e + 16 k loads the whole complex V_{4t+k}[s0 + e] (one 16- or 8-byte load: a load
// instruction covers 16 consecutive rows of 4 vectors), and each step t runs two MFMAs on
// the same B[k][i] = S[4t+k][16 tau + i]: A[e][k] = Re V_{4t+k}[s0 + e] and the imaginary
// parts, so D[e][i] is the real (imaginary) part of output 16 tau + i at row s0 + e; lane
// l then holds rows s0 + (l >> 4) + 4 r (register r) of output 16 tau + (l & 15) and
// stores both parts as one complex.  NS = steps of the instantiation (>= ceil(m / 4));
// vectors j >= m are loaded clamped to vector m - 1 (addresses already read, so cached)
// and enter as zeros, rows >= n clamped to row n - 1 and never stored.  Each output
// element accumulates the same products in the same order whatever NS and U are.  A wave
// loads all m vectors of its slabs before it writes any output; S comes in dynamic LDS
// (m * kk doubles).
template <typename T, int NS, int U>
__global__ void __launch_bounds__(256) k_rotate_mfma(T* V, int64_t ld, int m, const double* __restrict__ S, int kk,
                                                     int64_t n) {
    extern __shared__ double lS[];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int e = lane & 15, k = lane >> 4;
    const int ntile = (kk + 15) / 16;
    const int64_t stride = (int64_t)gridDim.x * (256 / kWave) * 16 * U;
    for (int i = threadIdx.x; i < m * kk; i += blockDim.x) lS[i] = S[i];
    __syncthreads();
    for (int64_t s0 = ((int64_t)blockIdx.x * (256 / kWave) + wv) * 16 * U; s0 < n; s0 += stride) {
        double2 a[U][NS];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = s0 + 16 * u + e;
            const int64_t ra = r < n ? r : n - 1;
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const int j = 4 * t + k;
                a[u][t] = to_d2(V[(int64_t)(j < m ? j : m - 1) * ld + ra]);
            }
        }
        for (int tau = 0; tau < ntile; ++tau) {
            const int io = 16 * tau + e;   // output vector of the B column
#pragma unroll
            for (int u = 0; u < U; ++u) {
                dbl4 ar = {0.0, 0.0, 0.0, 0.0}, ai = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int t = 0; t < NS; ++t) {
                    const int j = 4 * t + k;
                    const bool ok = j < m;
                    const double sv = (ok && io < kk) ? lS[j * kk + io] : 0.0;
                    ar = mfma16x16(ok ? a[u][t].x : 0.0, sv, ar);
                    ai = mfma16x16(ok ? a[u][t].y : 0.0, sv, ai);
                }
                if (io < kk) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int64_t rw = s0 + 16 * u + k + 4 * r;
                        if (rw < n) V[(int64_t)io * ld + rw] = from_d2<T>(make_double2(ar[r], ai[r]));
                    }
                }
            }
        }
    }
}

// Launch of the rotation on stream st: the instantiation with the fewest steps that cover
// m among 40, 48 and 64 vectors (two slabs per wave up to 40 vectors, one above: register
// budget), one 4-wave block per 64 U rows up to 16384 blocks.
template <typename T>
inline hipError_t launch_rotate(hipStream_t st, T* V, int64_t ld, int m, const double* S, int kk, int64_t n) {
    const size_t lds = (size_t)m * kk * sizeof(double);
    auto go = [&](auto kern, int u) {
        const unsigned g = (unsigned)std::min<int64_t>(16384, (n + 64 * u - 1) / (64 * u));
        hipLaunchKernelGGL(kern, dim3(g), dim3(256), lds, st, V, ld, m, S, kk, n);
        return hipGetLastError();
    };
    if (m <= 40) return go(k_rotate_mfma<T, 10, 2>, 2);
    if (m <= 48) return go(k_rotate_mfma<T, 12, 1>, 1);
    return go(k_rotate_mfma<T, 16, 1>, 1);
}

}  // namespace sed
