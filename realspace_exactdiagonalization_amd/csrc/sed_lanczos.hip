// Lanczos eigensolver around the sector H|psi> on gfx950.
//
// Replaces diagonalize_block_matrixfree (src/symmetry_resolved_ed.jl:998-1030), which
// hands the operator to KrylovKit.eigsolve(H, x0, nev, :SR; tol=1e-10, maxiter=300)
// (Lanczos with restarts and full re-orthogonalisation) and solves sectors of
// dimension <= 500 densely.  Here:
//   * thick-restart Lanczos (Wu & Simon) with a basis of at most m vectors resident in
//     HBM and partial re-orthogonalisation (Simon): each new vector is orthogonalised
//     against the two previous ones, and against the whole basis (classical
//     Gram-Schmidt, a second pass when the first cancelled most of it) only when
//     Simon's recurrence estimates a loss of orthogonality beyond sqrt(eps), for the
//     two steps after a restart, and once more after each such step; the alpha/beta
//     recurrence and every projection coefficient are carried in FP64 whatever the
//     vector precision;
//   * one Lanczos step is a fixed sequence of kernels whose step index j lives in device
//     memory: coefficients, norms, alpha_j and beta_j never leave the device inside a
//     step, so steps run back to back without host synchronisation and the sequence is
//     captured once into a hipGraph and replayed;
//   * all reductions run over a fixed grid in a fixed order (bitwise reproducible);
//   * the host reads alpha/beta every few steps, diagonalises the projected matrix
//     (order <= m), checks the residuals and drives the thick restarts;
//   * SED_PREC_FP32_REFINE runs the loop with complex-FP32 vectors and matvec, then
//     restarts an FP64 loop from the sum of the FP32 Ritz vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>

#include "sed_device.hpp"
#include "sed_gemm.hpp"
#include "sed_internal.hpp"
#include "sed_vec.hpp"

namespace sed {

int launch_dense(sed_sector* sec, double* H_dev, hipStream_t s);
int launch_matvec_xhat(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s);

constexpr int kRedGrid = 1024;    // fixed reduction grid (determinism)
constexpr int kRedBlock = 256;
constexpr int kDotGroup = 8;
constexpr int kRitzRows = 32;
constexpr int kUpdGrid = 1024;  // fixed grid of the update / norm kernels

// ---- one Lanczos step, step index j = *step on the device ----------------------
// The step reads v_j through xhat = sqrt(stab) v_j (written by the previous step),
// w = H v_j, then a classical Gram-Schmidt pass against v_lo..v_j into v_{j+1} (lo set by
// the partial re-orthogonalisation state: j - 1, or 0 where the whole basis is due), a
// second pass against V[0..j] in place when pro_decide asks for it, then
// v_{j+1} = w' / |w'| and its xhat.

// part[i * G + b] = sum over block b's rows of conj(V_i[r]) w[r] for lo <= i <= j, and
// part[(j+1) * G + b] = sum |w[r]|^2 (range and dispatch as pass_range; w = v_{j+1}
// when w_next).
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_dots(const T* __restrict__ V0, int64_t n, const int* __restrict__ step,
                                                    const T* __restrict__ w0, double2* __restrict__ part0,
                                                    const int* __restrict__ gate, const int* __restrict__ lo,
                                                    int w_next, int disp) {
    __shared__ double2 red[kDotGroup][kRedBlock / kWave];
    if (gate && *gate == 0) return;   // second pass not needed this step
    PassRange pr;
    if (!pass_range(step, lo, disp, pr)) return;
    const T* __restrict__ V = V0 + (int64_t)pr.lo * n;
    const T* __restrict__ w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    double2* __restrict__ part = part0 + (int64_t)pr.lo * gridDim.x;
    const int nv = pr.nv;
    const int total = nv + 1;
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = std::min<int64_t>(n, r0 + chunk);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (int g0 = 0; g0 < total; g0 += kDotGroup) {
        double2 acc[kDotGroup];
#pragma unroll
        for (int q = 0; q < kDotGroup; ++q) acc[q] = make_double2(0.0, 0.0);
        for (int64_t r = r0 + threadIdx.x; r < r1; r += kRedBlock) {
            const double2 wr = to_d2(w[r]);
#pragma unroll
            for (int q = 0; q < kDotGroup; ++q) {
                const int i = g0 + q;
                if (i < nv) {
                    const double2 v = to_d2(V[(int64_t)i * n + r]);
                    acc[q].x += v.x * wr.x + v.y * wr.y;
                    acc[q].y += v.x * wr.y - v.y * wr.x;
                } else if (i == nv) {
                    acc[q].x += wr.x * wr.x + wr.y * wr.y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kDotGroup; ++q) {
            const double2 sm = wave_sum_d2(acc[q]);
            if (lane == 0) red[q][wv] = sm;
        }
        __syncthreads();
        if (threadIdx.x < kDotGroup && g0 + (int)threadIdx.x < total) {
            double2 sm = make_double2(0.0, 0.0);
            for (int i = 0; i < kRedBlock / kWave; ++i) {
                sm.x += red[threadIdx.x][i].x;
                sm.y += red[threadIdx.x][i].y;
            }
            part[(int64_t)(g0 + threadIdx.x) * gridDim.x + blockIdx.x] = sm;
        }
        __syncthreads();
    }
}

// out[i] = sum_b part[i * G + b] for lo <= i <= j + 1 (fixed order); launched with
// m + 2 blocks.
__global__ void __launch_bounds__(kRedBlock) k_reduce(const double2* __restrict__ part, int G,
                                                      const int* __restrict__ step, double2* __restrict__ out,
                                                      const int* __restrict__ gate, const int* __restrict__ lo) {
    __shared__ double2 red[kRedBlock / kWave];
    if (gate && *gate == 0) return;
    const int i = blockIdx.x;
    if (i > *step + 1 || i < (lo ? *lo : 0)) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < G; b += kRedBlock) {
        acc.x += part[(int64_t)i * G + b].x;
        acc.y += part[(int64_t)i * G + b].y;
    }
    acc = wave_sum_d2(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 sm = make_double2(0.0, 0.0);
        for (int k = 0; k < kRedBlock / kWave; ++k) {
            sm.x += red[k].x;
            sm.y += red[k].y;
        }
        out[i] = sm;
    }
}

// out[r] = w[r] - sum_{lo<=i<=j} V_i[r] h[i]; out = w (to_next = 0) or V_{j+1}
// (to_next = 1), w = v_{j+1} itself when w_next (in place, row by row).  Each block also
// writes the partial |out|^2 of its rows (fixed grid).  A zero gate, or a range that the
// other kernel family takes, leaves everything untouched.
template <typename T>
__global__ void __launch_bounds__(256) k_update(T* V0, int64_t n, const int* __restrict__ step,
                                                const double2* __restrict__ h, T* w0, int to_next,
                                                double* __restrict__ normpart, const int* __restrict__ gate,
                                                const int* __restrict__ lo, int w_next, int disp) {
    __shared__ double2 lh[kMaxBasis + 2];
    __shared__ double red[256 / kWave];
    if (gate && *gate == 0) return;
    PassRange pr;
    if (!pass_range(step, lo, disp, pr)) return;
    const int nv = pr.nv;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) lh[i] = h[pr.lo + i];
    __syncthreads();
    const T* V = V0 + (int64_t)pr.lo * n;
    T* out = to_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    const T* w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    double nrm = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 a = to_d2(w[r]);
        for (int i = 0; i < nv; ++i) {
            const double2 v = to_d2(V[(int64_t)i * n + r]);
            const double2 c = lh[i];
            a.x -= v.x * c.x - v.y * c.y;
            a.y -= v.x * c.y + v.y * c.x;
        }
        const T o = from_d2<T>(a);
        out[r] = o;
        const double2 od = to_d2(o);
        nrm += od.x * od.x + od.y * od.y;
    }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int k = 0; k < 256 / kWave; ++k) sm += red[k];
        normpart[blockIdx.x] = sm;
    }
}

// The row-wise kernels of a range of fewer than kMfmaMinVectors vectors beside the
// matrix-core kernels (disp 2; the local pass v_{j-1}, v_j of most steps): the same sums
// in the same order as k_dots / k_update (bitwise equal), with U rows of a thread in
// flight: every load of U rows (w and kFew vectors, slots past the range re-reading
// vector lo, which is cached) is issued before the first product.  Complex-FP32 vectors
// otherwise keep too few bytes in flight per wave to stream at full rate.
constexpr int kFew = kMfmaMinVectors - 1;

template <typename T, int U>
__global__ void __launch_bounds__(kRedBlock) k_dots_few(const T* __restrict__ V0, int64_t n,
                                                        const int* __restrict__ step, const T* __restrict__ w0,
                                                        double2* __restrict__ part0, const int* __restrict__ gate,
                                                        const int* __restrict__ lo, int w_next) {
    __shared__ double2 red[kFew + 1][kRedBlock / kWave];
    if (gate && *gate == 0) return;
    PassRange pr;
    if (!pass_range(step, lo, 2, pr)) return;
    const T* __restrict__ V = V0 + (int64_t)pr.lo * n;
    const T* __restrict__ w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    double2* __restrict__ part = part0 + (int64_t)pr.lo * gridDim.x;
    const int nv = pr.nv;
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = std::min<int64_t>(n, r0 + chunk);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    double2 acc[kFew + 1];
#pragma unroll
    for (int q = 0; q <= kFew; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int64_t rb = r0 + threadIdx.x; rb < r1; rb += (int64_t)kRedBlock * U) {
        T wr[U], vv[U][kFew];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = rb + (int64_t)u * kRedBlock;
            const int64_t rc = r < r1 ? r : rb;
            wr[u] = w[rc];
#pragma unroll
            for (int q = 0; q < kFew; ++q) vv[u][q] = V[(int64_t)(q < nv ? q : 0) * n + rc];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (rb + (int64_t)u * kRedBlock >= r1) break;
            const double2 wd = to_d2(wr[u]);
#pragma unroll
            for (int q = 0; q <= kFew; ++q) {
                if (q < nv) {
                    const double2 v = to_d2(vv[u][q < kFew ? q : 0]);
                    acc[q].x += v.x * wd.x + v.y * wd.y;
                    acc[q].y += v.x * wd.y - v.y * wd.x;
                } else if (q == nv) {
                    acc[q].x += wd.x * wd.x + wd.y * wd.y;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q <= kFew; ++q) {
        const double2 sm = wave_sum_d2(acc[q]);
        if (lane == 0) red[q][wv] = sm;
    }
    __syncthreads();
    if ((int)threadIdx.x <= nv) {
        double2 sm = make_double2(0.0, 0.0);
        for (int i = 0; i < kRedBlock / kWave; ++i) {
            sm.x += red[threadIdx.x][i].x;
            sm.y += red[threadIdx.x][i].y;
        }
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = sm;
    }
}

template <typename T, int U>
__global__ void __launch_bounds__(256) k_update_few(T* V0, int64_t n, const int* __restrict__ step,
                                                    const double2* __restrict__ h, T* w0, int to_next,
                                                    double* __restrict__ normpart, const int* __restrict__ gate,
                                                    const int* __restrict__ lo, int w_next) {
    __shared__ double2 lh[kFew];
    __shared__ double red[256 / kWave];
    if (gate && *gate == 0) return;
    PassRange pr;
    if (!pass_range(step, lo, 2, pr)) return;
    const int nv = pr.nv;
    if ((int)threadIdx.x < nv) lh[threadIdx.x] = h[pr.lo + threadIdx.x];
    __syncthreads();
    const T* V = V0 + (int64_t)pr.lo * n;
    T* out = to_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    const T* w = w_next ? V0 + (int64_t)(pr.j + 1) * n : w0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double nrm = 0.0;
    for (int64_t rb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rb < n; rb += stride * U) {
        T wr[U], vv[U][kFew];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = rb + u * stride;
            const int64_t rc = r < n ? r : rb;
            wr[u] = w[rc];
#pragma unroll
            for (int i = 0; i < kFew; ++i) vv[u][i] = V[(int64_t)(i < nv ? i : 0) * n + rc];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = rb + u * stride;
            if (r >= n) break;
            double2 a = to_d2(wr[u]);
#pragma unroll
            for (int i = 0; i < kFew; ++i) {
                if (i < nv) {
                    const double2 v = to_d2(vv[u][i]);
                    const double2 c = lh[i];
                    a.x -= v.x * c.x - v.y * c.y;
                    a.y -= v.x * c.y + v.y * c.x;
                }
            }
            const T o = from_d2<T>(a);
            out[r] = o;
            const double2 od = to_d2(o);
            nrm += od.x * od.x + od.y * od.y;
        }
    }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int k = 0; k < 256 / kWave; ++k) sm += red[k];
        normpart[blockIdx.x] = sm;
    }
}
// rows per thread in flight: measured on a 15.7M-row sector (local pass, two vectors),
// complex FP32 dots 127 -> 84 us and update 125 -> 100 us at two rows; complex FP64 dots
// 147 -> 137 us at four rows, its update fastest as k_update (one row)
template <typename T>
struct FewRows;
template <>
struct FewRows<float2> {
    static constexpr int dots = 2, update = 2;
};
template <>
struct FewRows<double2> {
    static constexpr int dots = 4, update = 1;
};

// Single block, after the first pass: |w'|^2 from the update's partials; pro_decide
// sets the gate of the second pass (and clears h2[j] when it does not run).
__global__ void __launch_bounds__(kRedBlock) k_reorth_check(const int* __restrict__ step,
                                                            const double* __restrict__ normpart, int G,
                                                            const double2* __restrict__ h1, double2* __restrict__ h2,
                                                            int* __restrict__ gate, ProState* __restrict__ pro,
                                                            const double* __restrict__ alpha,
                                                            const double* __restrict__ beta) {
    __shared__ double red[kRedBlock / kWave];
    __shared__ int sh_need;
    double acc = 0.0;
    for (int b = threadIdx.x; b < G; b += kRedBlock) acc += normpart[b];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    double nrm2 = 0.0;
    for (int k = 0; k < kRedBlock / kWave; ++k) nrm2 += red[k];   // the same order on every thread
    __syncthreads();
    const int j = *step;
    const bool need = pro_decide_wg<kRedBlock>(pro, j, nrm2, h1[j + 1].x, h1[j].x, alpha, beta, red, &sh_need);
    if (threadIdx.x != 0) return;
    *gate = need ? 1 : 0;
    if (!need) h2[j] = make_double2(0.0, 0.0);
}

// Single block: |v_{j+1}|^2 from the update's partials, alpha_j, beta_j, 1/beta_j, and
// the step counter advances to j + 1.  A vanishing residual (invariant subspace) is
// flagged with the step at which it happened and scales the new vector by zero.
__global__ void __launch_bounds__(kRedBlock) k_finalize(int* step, const double2* __restrict__ h1,
                                                        const double2* __restrict__ h2,
                                                        const double* __restrict__ normpart, int G,
                                                        double* __restrict__ alpha, double* __restrict__ beta,
                                                        double* __restrict__ scale, int* __restrict__ flag,
                                                        ProState* __restrict__ pro) {
    __shared__ double red[kRedBlock / kWave];
    double acc = 0.0;
    for (int b = threadIdx.x; b < G; b += kRedBlock) acc += normpart[b];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double nrm2 = 0.0;
    for (int k = 0; k < kRedBlock / kWave; ++k) nrm2 += red[k];
    const int j = *step;
    const double a = h1[j].x + h2[j].x;
    const double hv = h1[j + 1].x;   // |H v_j|^2 before orthogonalisation
    const double b = sqrt(nrm2);
    const bool brk = !(b > 1e-12 * sqrt(hv)) || !(b > 0.0);
    alpha[j] = a;
    beta[j] = b;
    *scale = brk ? 0.0 : 1.0 / b;
    if (brk && *flag == 0) *flag = j + 1;
    *step = j + 1;
    pro_advance(pro, j, a, b, beta);
}

// The whole step after the matvec in one workgroup, for sectors of at most kTailMaxRows
// rows (their passes read tens of kilobytes, so one kernel with workgroup barriers beats
// the eight launches of the multi-workgroup sequence: XXZ N=16 and 18 solves 17 % and 9 %
// faster; from about 9,000 rows one CU's bandwidth makes it slower): the first Gram-Schmidt pass
// over v_lo..v_j on w in place, pro_decide, the second pass over the whole basis when it
// asks for it, alpha_j, beta_j and the breakdown flag, the step counter, and
// v_{j+1} = w' / beta_j with its xhat.  The same arithmetic as k_dots .. k_scale, with the
// reductions in one workgroup's fixed order.
constexpr int kTailThreads = 1024;
constexpr int64_t kTailMaxRows = 1 << 12;
template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_step_tail(T* __restrict__ V, int64_t n, int* step, T* __restrict__ w,
                                                            double* __restrict__ alpha, double* __restrict__ beta,
                                                            int* __restrict__ flag, ProState* __restrict__ pro,
                                                            const uint8_t* __restrict__ stab, T* __restrict__ xhat) {
    constexpr int NT = kTailThreads;
    __shared__ double2 sh_h[kMaxBasis + 2];
    __shared__ double2 sh_red[kBlockDots][NT / kWave];
    __shared__ double sh_redd[NT / kWave];
    __shared__ double sh_nrm;
    __shared__ int sh_j, sh_lo, sh_need;
    if (threadIdx.x == 0) {
        sh_j = *step;
        sh_lo = pro->lo;
    }
    __syncthreads();
    const int j = sh_j, lo = sh_lo, nv = j + 1;
    block_dots<T, NT>(V + (int64_t)lo * n, n, nv - lo, w, sh_h + lo, &sh_nrm, sh_red);
    const double nw = sh_nrm;
    double a = sh_h[j].x;
    double nw2 = block_update<T, NT>(V + (int64_t)lo * n, n, nv - lo, sh_h + lo, w, sh_redd);
    if (pro_decide_wg<NT>(pro, j, nw2, nw, a, alpha, beta, sh_redd, &sh_need)) {
        block_dots<T, NT>(V, n, nv, w, sh_h, &sh_nrm, sh_red);
        a += sh_h[j].x;
        nw2 = block_update<T, NT>(V, n, nv, sh_h, w, sh_redd);
    }
    const double b = sqrt(nw2);
    const bool brk = !(b > 1e-12 * sqrt(nw)) || !(b > 0.0);
    const double sc = brk ? 0.0 : 1.0 / b;
    if (threadIdx.x == 0) {
        alpha[j] = a;
        beta[j] = b;
        if (brk && *flag == 0) *flag = j + 1;
        *step = j + 1;
        pro_advance(pro, j, a, b, beta);
    }
    T* vn = V + (int64_t)(j + 1) * n;
    for (int64_t r = threadIdx.x; r < n; r += NT) {
        const double2 d = to_d2(w[r]);
        const T o = from_d2<T>(make_double2(d.x * sc, d.y * sc));
        vn[r] = o;
        const double2 od = to_d2(o);
        const double s = sqrt((double)stab[r]);
        xhat[r] = from_d2<T>(make_double2(od.x * s, od.y * s));
    }
}

// v_j *= scale and xhat = sqrt(stab) v_j for j = *step (the vector the next step applies H to).
template <typename T>
__global__ void __launch_bounds__(256) k_scale(T* V, int64_t n, const int* __restrict__ step,
                                               const double* __restrict__ scale, const uint8_t* __restrict__ stab,
                                               T* __restrict__ xhat) {
    T* v = V + (int64_t)(*step) * n;
    const double sc = *scale;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(v[r]);
        const T o = from_d2<T>(make_double2(a.x * sc, a.y * sc));
        v[r] = o;
        const double2 od = to_d2(o);
        const double s = sqrt((double)stab[r]);
        xhat[r] = from_d2<T>(make_double2(od.x * s, od.y * s));
    }
}

// |v_0|^2 partials for the start vector (fixed grid), then k_start_finalize
template <typename T>
__global__ void __launch_bounds__(256) k_normsq(const T* __restrict__ v, int64_t n, double* __restrict__ normpart) {
    __shared__ double red[256 / kWave];
    double nrm = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(v[r]);
        nrm += a.x * a.x + a.y * a.y;
    }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int k = 0; k < 256 / kWave; ++k) sm += red[k];
        normpart[blockIdx.x] = sm;
    }
}

__global__ void __launch_bounds__(kRedBlock) k_start_finalize(int* step, int start, const double* __restrict__ normpart,
                                                              int G, double* __restrict__ scale) {
    __shared__ double red[kRedBlock / kWave];
    double acc = 0.0;
    for (int b = threadIdx.x; b < G; b += kRedBlock) acc += normpart[b];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double nrm2 = 0.0;
    for (int k = 0; k < kRedBlock / kWave; ++k) nrm2 += red[k];
    *scale = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 0.0;
    *step = start;
}

// In place, row by row: V_i[r] <- sum_{j<m} V_j[r] S[j * kk + i] for i < kk (S real).
template <typename T>
__global__ void __launch_bounds__(kRitzRows) k_ritz(T* V, int64_t ld, int m, const double* __restrict__ S, int kk,
                                                    int64_t n) {
    extern __shared__ __align__(16) uint8_t smem[];
    double* lS = reinterpret_cast<double*>(smem);
    T* tile = reinterpret_cast<T*>(smem + (((size_t)m * kk * 8 + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < m * kk; i += blockDim.x) lS[i] = S[i];
    const int64_t r = (int64_t)blockIdx.x * kRitzRows + threadIdx.x;
    if (r < n)
        for (int j = 0; j < m; ++j) tile[j * kRitzRows + threadIdx.x] = V[(int64_t)j * ld + r];
    __syncthreads();
    if (r >= n) return;
    for (int i = 0; i < kk; ++i) {
        double ax = 0.0, ay = 0.0;
        for (int j = 0; j < m; ++j) {
            const double2 v = to_d2(tile[j * kRitzRows + threadIdx.x]);
            const double s = lS[j * kk + i];
            ax += v.x * s;
            ay += v.y * s;
        }
        V[(int64_t)i * ld + r] = from_d2<T>(make_double2(ax, ay));
    }
}

// Seeded complex Gaussian start vector (gauss_entry per row).
template <typename T>
__global__ void __launch_bounds__(256) k_random(T* __restrict__ v, int64_t n, uint64_t seed) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        v[r] = from_d2<T>(gauss_entry(seed, r));
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_convert(const Tin* __restrict__ a, Tout* __restrict__ b, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        b[r] = from_d2<Tout>(to_d2(a[r]));
}

// Partial sums of |w - theta v|^2 (the true residual of a Ritz pair, w = H v).
template <typename T>
__global__ void __launch_bounds__(256) k_resid_part(const T* __restrict__ w, const T* __restrict__ v, double theta,
                                                    int64_t n, double* __restrict__ part) {
    __shared__ double red[256 / kWave];
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(w[r]), b = to_d2(v[r]);
        const double x = a.x - theta * b.x, y = a.y - theta * b.y;
        acc += x * x + y * y;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int k = 0; k < 256 / kWave; ++k) sm += red[k];
        part[blockIdx.x] = sm;
    }
}

// sum of nv vectors (complex double out)
template <typename T>
__global__ void __launch_bounds__(256) k_sum_vectors(const T* __restrict__ V, int64_t ld, int nv,
                                                     double2* __restrict__ out, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 a = make_double2(0.0, 0.0);
        for (int i = 0; i < nv; ++i) {
            const double2 v = to_d2(V[(int64_t)i * ld + r]);
            a.x += v.x;
            a.y += v.y;
        }
        out[r] = a;
    }
}

static inline unsigned grid_for(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 20));
}

// ------------------------------------------------------------------------
// host driver
// ------------------------------------------------------------------------
// The pinned host block of a workspace is allocated once, at the largest size any solve
// uses (the alpha/beta/flag block of the largest basis, or the fixed reduction grid), and
// kept until release(): it is never re-allocated between or during solves, so no copy in
// flight can land in freed pinned memory.
constexpr size_t kHostDoubles = std::max<size_t>(4 * (kMaxBasis + 2) + 2 * kMaxBasis + 1 + 8, kUpdGrid);

int Workspace::ensure_host(size_t n_doubles) {
    if (n_doubles > kHostDoubles) return fail(SED_EINVAL, "pinned host block too small");
    if (hcoef) return SED_OK;
    hipError_t e = hipHostMalloc((void**)&hcoef, kHostDoubles * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        hcoef = nullptr;
        hcoef_n = 0;
        return fail(SED_ENOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    hcoef_n = kHostDoubles;
    return SED_OK;
}

// A captured step is destroyed only after the stream has drained, so no replay of it is
// still queued or running when its executable graph goes away.
void Workspace::drop_graph() {
    if (graph_exec) {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipGraphExecDestroy(graph_exec);
    }
    graph_exec = nullptr;
    graph_key = GraphKey();
}

void Workspace::release() {
    drop_graph();
    V64.release(); w64.release(); xh64.release();
    V32.release(); w32.release(); xh32.release();
    part.release(); coef.release(); S.release(); v0.release(); state.release();
    if (hcoef) (void)hipHostFree(hcoef);
    hcoef = nullptr;
    hcoef_n = 0;
}

size_t Workspace::device_bytes() const {
    return V64.bytes() + w64.bytes() + xh64.bytes() + V32.bytes() + w32.bytes() + xh32.bytes() + part.bytes() +
           coef.bytes() + S.bytes() + v0.bytes() + state.bytes();
}

template <typename T>
struct VecTraits;
template <>
struct VecTraits<double2> {
    static constexpr bool fp64 = true;
    static double* V(Workspace& w) { return w.V64.p; }
    static double* W(Workspace& w) { return w.w64.p; }
    static double* X(Workspace& w) { return w.xh64.p; }
    static int reserve(Workspace& w, int64_t n, int m) {
        SED_TRY(grow(w.V64, (size_t)2 * (m + 1) * n));
        SED_TRY(grow(w.w64, (size_t)2 * n));
        SED_TRY(grow(w.xh64, (size_t)2 * n));
        return SED_OK;
    }
};
template <>
struct VecTraits<float2> {
    static constexpr bool fp64 = false;
    static float* V(Workspace& w) { return w.V32.p; }
    static float* W(Workspace& w) { return w.w32.p; }
    static float* X(Workspace& w) { return w.xh32.p; }
    static int reserve(Workspace& w, int64_t n, int m) {
        SED_TRY(grow(w.V32, (size_t)2 * (m + 1) * n));
        SED_TRY(grow(w.w32, (size_t)2 * n));
        SED_TRY(grow(w.xh32, (size_t)2 * n));
        return SED_OK;
    }
};

struct SolveOut {
    std::vector<double> evals;
    std::vector<double> resid;
    int converged = 0;
    int iters = 0;
    int restarts = 0;
    int full_steps = 0;   // steps orthogonalised against the whole basis
    int64_t matvecs = 0;
};

// Re-orthogonalisation and restart rotation on the FP64 matrix cores for sectors of at
// least kMfmaMinRows rows, as row-wise VALU loops below that: there the basis sits in the
// caches and a step is bound by its launch count, which a single kernel family keeps
// lowest.  SED_REORTH=valu / SED_REORTH=mfma force either form (read once per solve).
constexpr int64_t kMfmaMinRows = 1 << 17;
static bool reorth_mfma(int64_t n) {
    const char* e = std::getenv("SED_REORTH");
    if (e && std::strcmp(e, "valu") == 0) return false;
    if (e && std::strcmp(e, "mfma") == 0) return true;
    return n >= kMfmaMinRows;
}

// Partial re-orthogonalisation (default), or a pass over the whole basis at every step
// with SED_PRO=0 (kept for comparison; read once per solve).
static bool pro_enabled() {
    const char* e = std::getenv("SED_PRO");
    return !(e && e[0] == '0');
}

// Largest nev solved with partial re-orthogonalisation in the per-sector path
// (kProMaxNev; SED_PRO_MAX_NEV overrides it, e.g. to exercise the residual check).
static int pro_max_nev() {
    const char* e = std::getenv("SED_PRO_MAX_NEV");
    return e && *e ? std::atoi(e) : kProMaxNev;
}

// Captured steps are replayed unless SED_LANCZOS_GRAPH=0, and never under rocprofv3's
// kernel tracing (it sets ROCPROF_KERNEL_TRACE): traced runs whose solve launched a step
// graph after an earlier solve in the process (the 48-site checkerboard sector solved
// twice; a flux scan of 24-site sectors) ended with SIGSEGV in a host memcpy from an
// unmapped address, called from the profiler's library below hipGraphLaunch.  The same
// solves without the profiler, and eager launches of the same kernels under it, run
// without fault.
static bool graphs_enabled() {
    const char* t = std::getenv("ROCPROF_KERNEL_TRACE");
    if (t && *t && t[0] != '0') return false;
    const char* e = std::getenv("SED_LANCZOS_GRAPH");
    return !(e && e[0] == '0');
}

// SED_SOLVE_EVENTS=1: every Lanczos loop launches its steps eagerly with HIP events between
// the parts of a step and prints, at its end, the GPU time of each part summed over the
// steps (H|psi>, first Gram-Schmidt pass with its check, gated second pass, finalisation and
// scaling, gaps between consecutive steps; sectors of at most kTailMaxRows rows count their
// one-workgroup step tail as the first pass), and the wall time of the thick restarts and of
// the host checks, to stderr.  Read once per loop.
static bool solve_events_enabled() {
    const char* e = std::getenv("SED_SOLVE_EVENTS");
    return e && e[0] == '1';
}

// Device state block (doubles): h1 (2*(m+2)), h2 (2*(m+2)), alpha (m), beta (m), scale (1);
// ints after it: step, flag, gate; then the partial re-orthogonalisation state.
template <typename T>
class ThickLanczos {
   public:
    ThickLanczos(sed_sector* sec, int m, Workspace& ws)
        : sec_(sec), n_(sec->dim), m_(m), ws_(ws), st_(ws.stream), mfma_(reorth_mfma(sec->dim)), pro_on_(pro_enabled()),
          ev_on_(solve_events_enabled()) {}
    ~ThickLanczos() {
        for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    }
    ThickLanczos(const ThickLanczos&) = delete;
    ThickLanczos& operator=(const ThickLanczos&) = delete;

    int init() {
        SED_TRY(VecTraits<T>::reserve(ws_, n_, m_));
        G_ = (int)std::max<int64_t>(1, std::min<int64_t>(kRedGrid, (n_ + kRedBlock - 1) / kRedBlock));
        SED_TRY(grow(ws_.part, (size_t)2 * (m_ + 2) * G_ + kUpdGrid));
        SED_TRY(grow(ws_.S, (size_t)(m_ + 1) * (m_ + 1)));
        const size_t nd = 4 * (size_t)(m_ + 2) + 2 * (size_t)m_ + 1;
        const size_t npro = (sizeof(ProState) + sizeof(double) - 1) / sizeof(double);
        SED_TRY(grow(ws_.state, nd + 4 + npro));
        SED_TRY(ws_.ensure_host(std::max<size_t>(nd + 8, kUpdGrid)));
        V_ = reinterpret_cast<T*>(VecTraits<T>::V(ws_));
        w_ = reinterpret_cast<T*>(VecTraits<T>::W(ws_));
        xh_ = reinterpret_cast<T*>(VecTraits<T>::X(ws_));
        part_ = reinterpret_cast<double2*>(ws_.part.p);
        normpart_ = ws_.part.p + (size_t)2 * (m_ + 2) * G_;
        double* st = ws_.state.p;
        h1_ = reinterpret_cast<double2*>(st);
        h2_ = reinterpret_cast<double2*>(st + 2 * (m_ + 2));
        alpha_ = st + 4 * (m_ + 2);
        beta_ = alpha_ + m_;
        scale_ = beta_ + m_;
        step_ = reinterpret_cast<int*>(scale_ + 1);
        flag_ = step_ + 1;
        gate_ = step_ + 2;
        pro_ = reinterpret_cast<ProState*>(st + nd + 4);
        int zero[2] = {0, 0};
        SED_HIP(hipMemcpyAsync(step_, zero, sizeof(zero), hipMemcpyHostToDevice, st_));
        ProState p{};
        p.enabled = pro_on_ ? 1 : 0;
        p.eps = VecTraits<T>::fp64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
        p.delta = std::sqrt(p.eps);
        SED_HIP(hipMemcpyAsync(pro_, &p, sizeof(p), hipMemcpyHostToDevice, st_));
        SED_HIP(hipStreamSynchronize(st_));  // `zero` and `p` are stack temporaries
        return SED_OK;
    }
    // steps of the solve whose orthogonalisation spanned the whole basis
    int read_full_steps(int& out) {
        SED_HIP(hipMemcpyAsync(&ws_.hcoef[0], &pro_->nfull, sizeof(int), hipMemcpyDeviceToHost, st_));
        SED_HIP(hipStreamSynchronize(st_));
        std::memcpy(&out, &ws_.hcoef[0], sizeof(int));
        return SED_OK;
    }
    T* vec(int i) { return V_ + (size_t)i * n_; }

    int start_random(uint64_t seed) {
        hipLaunchKernelGGL(k_random<T>, dim3(grid_for(n_)), dim3(256), 0, st_, vec(0), n_, seed);
        SED_HIP(hipGetLastError());
        return normalize_start(0);
    }
    int start_from(const double2* v0) {
        hipLaunchKernelGGL((k_convert<double2, T>), dim3(grid_for(n_)), dim3(256), 0, st_, v0, vec(0), n_);
        SED_HIP(hipGetLastError());
        return normalize_start(0);
    }

    int run(int nev, double tol, int maxiter, SolveOut& out) {
        std::vector<double> T_((size_t)m_ * m_, 0.0);  // row-major projected matrix
        int k = 0;
        double beta = 0.0;
        out = SolveOut();
        std::vector<double> theta, Sm, last;
        int jm = 0;
        bool done = false, breakdown = false;
        // first matvec eagerly: it fixes the launch shape (autotuner) before any capture
        const bool use_graph = graphs_enabled() && !ev_on_;
        if (ev_on_) SED_TRY(events_init());
        const auto t_run = std::chrono::steady_clock::now();
        while (!done) {
            int j = k;
            while (j < m_) {
                int chunk = std::min(m_ - j, j < nev + 2 ? nev + 2 - j : kCheckEvery);
                for (int c = 0; c < chunk; ++c) SED_TRY(step(use_graph && out.iters > 0));
                out.iters += chunk;
                out.matvecs += chunk;
                j += chunk;
                // read alpha/beta/flag of this cycle
                const size_t na = (size_t)m_;
                SED_HIP(hipMemcpyAsync(ws_.hcoef, alpha_, (2 * na + 1) * sizeof(double), hipMemcpyDeviceToHost, st_));
                SED_HIP(hipMemcpyAsync(ws_.hcoef + 2 * na + 2, step_, 2 * sizeof(int), hipMemcpyDeviceToHost, st_));
                SED_HIP(hipStreamSynchronize(st_));
                const auto t_check = std::chrono::steady_clock::now();
                if (ev_on_) SED_TRY(events_collect());
                const double* ha = ws_.hcoef;
                const double* hb = ws_.hcoef + na;
                int flag = 0;
                std::memcpy(&flag, reinterpret_cast<const char*>(ws_.hcoef + 2 * na + 2) + sizeof(int), sizeof(int));
                jm = j;
                if (flag > 0 && flag <= j) {
                    breakdown = true;
                    jm = flag;   // v_flag could not be formed: basis V[0..flag-1]
                }
                for (int i = k; i < jm; ++i) {
                    T_[(size_t)i * m_ + i] = ha[i];
                    if (i + 1 < jm) {
                        T_[(size_t)(i + 1) * m_ + i] = hb[i];
                        T_[(size_t)i * m_ + i + 1] = hb[i];
                    }
                }
                beta = breakdown ? 0.0 : hb[jm - 1];
                // eigenvalues and the last row of the eigenvectors decide convergence; the
                // whole eigenvector matrix only where Ritz vectors are formed
                eig_last(jm, T_, theta, last);
                int conv = 0;
                for (int i = 0; i < std::min(nev, jm); ++i) {
                    const double res = beta * std::fabs(last[i]);
                    if (res <= tol) ++conv; else break;
                }
                if (ev_on_) ev_host_ms_ += ms_since(t_check);
                if (breakdown || conv >= nev || (j == m_ && out.restarts >= maxiter)) {
                    done = true;
                    break;
                }
            }
            const auto t_restart = std::chrono::steady_clock::now();
            eig(jm, T_, theta, Sm);
            if (done) break;
            // thick restart: kk Ritz vectors + the residual direction v_m
            const int kk = std::min(m_ - 2, nev + std::max(4, (m_ - nev) / 3));
            std::vector<double> Sk((size_t)jm * kk);
            for (int jj = 0; jj < jm; ++jj)
                for (int i = 0; i < kk; ++i) Sk[(size_t)jj * kk + i] = Sm[(size_t)jj * jm + i];
            SED_TRY(ritz(jm, Sk, kk));
            SED_HIP(hipMemcpyAsync(vec(kk), vec(jm), (size_t)n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
            std::fill(T_.begin(), T_.end(), 0.0);
            std::vector<double> arrow(kk);
            for (int i = 0; i < kk; ++i) {
                T_[(size_t)i * m_ + i] = theta[i];
                const double sc = beta * Sm[(size_t)(jm - 1) * jm + i];
                T_[(size_t)kk * m_ + i] = sc;
                T_[(size_t)i * m_ + kk] = sc;
                arrow[i] = sc;
            }
            SED_TRY(pro_restart(kk, theta, arrow));
            SED_TRY(set_step_and_xhat(kk));
            k = kk;
            ++out.restarts;
            if (ev_on_) ev_restart_ms_ += ms_since(t_restart);
        }
        const int nk = std::min(nev, jm);
        out.evals.assign(theta.begin(), theta.begin() + nk);
        out.resid.resize(nk);
        out.converged = 0;
        for (int i = 0; i < nk; ++i) {   // the estimates the last check decided on
            out.resid[i] = beta * std::fabs(last[i]);
            if (out.resid[i] <= tol && out.converged == i) ++out.converged;
        }
        std::vector<double> Sk((size_t)jm * nk);
        for (int jj = 0; jj < jm; ++jj)
            for (int i = 0; i < nk; ++i) Sk[(size_t)jj * nk + i] = Sm[(size_t)jj * jm + i];
        SED_TRY(ritz(jm, Sk, nk));
        nritz_ = nk;
        SED_TRY(read_full_steps(out.full_steps));
        if (ev_on_) {
            const double tot = ms_since(t_run);
            std::fprintf(stderr,
                         "[sed-solve-events] %s dim=%lld m=%d steps=%d restarts=%d full_orth_steps=%d wall_ms=%.3f "
                         "matvec_ms=%.3f pass1_ms=%.3f pass2_ms=%.3f finalize_scale_ms=%.3f step_gaps_ms=%.3f "
                         "restart_wall_ms=%.3f host_check_wall_ms=%.3f\n",
                         VecTraits<T>::fp64 ? "fp64" : "fp32", (long long)n_, m_, out.iters, out.restarts,
                         out.full_steps, tot, ev_ms_[0], ev_ms_[1], ev_ms_[2], ev_ms_[3], ev_ms_[4], ev_restart_ms_,
                         ev_host_ms_);
            std::fflush(stderr);
        }
        return SED_OK;
    }

    // True residuals |H x_i - theta_i x_i| of the nk Ritz pairs in V[0..nk) (FP64 vectors:
    // one FP64 matvec each, through the step's w and x-hat buffers).
    int true_residuals(int nk, const std::vector<double>& theta, std::vector<double>& res) {
        res.assign(nk, 0.0);
        for (int i = 0; i < nk; ++i) {
            SED_TRY(launch_matvec64(sec_, reinterpret_cast<const double*>(vec(i)), reinterpret_cast<double*>(w_),
                                    reinterpret_cast<double*>(xh_), st_));
            hipLaunchKernelGGL(k_resid_part<T>, dim3(kUpdGrid), dim3(256), 0, st_, w_, vec(i), theta[i], (int64_t)n_,
                               normpart_);
            SED_HIP(hipGetLastError());
            SED_HIP(hipMemcpyAsync(ws_.hcoef, normpart_, kUpdGrid * sizeof(double), hipMemcpyDeviceToHost, st_));
            SED_HIP(hipStreamSynchronize(st_));
            double sm = 0.0;
            for (int b = 0; b < kUpdGrid; ++b) sm += ws_.hcoef[b];
            res[i] = std::sqrt(sm);
        }
        return SED_OK;
    }
    bool pro_on() const { return pro_on_; }
    // every step orthogonalised against the whole basis (before init())
    void disable_pro() { pro_on_ = false; }

    int nritz_ = 0;
    T* V_ = nullptr;

   private:
    static constexpr int kCheckEvery = 8;

    // one Lanczos step as plain launches
    int launch_step() {
        hipEvent_t* ev = nullptr;   // 5 events of this step when timing, else none
        if (ev_on_ && (ev_n_ + 1) * kEvPerStep <= (int)ev_.size()) ev = &ev_[(size_t)(ev_n_++) * kEvPerStep];
        if (ev) SED_HIP(hipEventRecord(ev[0], st_));
        SED_TRY(launch_matvec_xhat(sec_, xh_, w_, VecTraits<T>::fp64, st_));
        if (ev) SED_HIP(hipEventRecord(ev[1], st_));
        if (n_ <= kTailMaxRows) {
            hipLaunchKernelGGL(k_step_tail<T>, dim3(1), dim3(kTailThreads), 0, st_, V_, (int64_t)n_, step_, w_, alpha_,
                               beta_, flag_, pro_, sec_->d_stab.p, xh_);
            if (ev) {
                SED_HIP(hipEventRecord(ev[2], st_));
                SED_HIP(hipEventRecord(ev[3], st_));
                SED_HIP(hipEventRecord(ev[4], st_));
            }
            SED_HIP(hipGetLastError());
            return SED_OK;
        }
        // first classical Gram-Schmidt pass over v_lo..v_j (lo decided on the device) into
        // v_{j+1}; a second pass over the whole basis, in place, only when the device
        // decides so (partial re-orthogonalisation or cancellation)
        SED_TRY(launch_orth_pass(h1_, &pro_->lo, 0, nullptr));
        hipLaunchKernelGGL(k_reorth_check, dim3(1), dim3(kRedBlock), 0, st_, step_, normpart_, kUpdGrid, h1_, h2_,
                           gate_, pro_, alpha_, beta_);
        if (ev) SED_HIP(hipEventRecord(ev[2], st_));
        SED_TRY(launch_orth_pass(h2_, nullptr, 1, gate_));
        if (ev) SED_HIP(hipEventRecord(ev[3], st_));
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kRedBlock), 0, st_, step_, h1_, h2_, normpart_, kUpdGrid, alpha_,
                           beta_, scale_, flag_, pro_);
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n_)), dim3(256), 0, st_, V_, (int64_t)n_, step_, scale_,
                           sec_->d_stab.p, xh_);
        if (ev) SED_HIP(hipEventRecord(ev[4], st_));
        SED_HIP(hipGetLastError());
        return SED_OK;
    }

    // event timing of the steps (SED_SOLVE_EVENTS=1)
    static constexpr int kEvPerStep = 5;
    static constexpr int kEvSteps = 64;   // more than the steps between two host checks
    int events_init() {
        if (!ev_.empty()) return SED_OK;
        ev_.resize((size_t)kEvSteps * kEvPerStep, nullptr);
        for (hipEvent_t& e : ev_) SED_HIP(hipEventCreate(&e));
        return SED_OK;
    }
    // after a stream synchronisation: add the recorded steps' part times, then reuse the pool
    int events_collect() {
        for (int s = 0; s < ev_n_; ++s) {
            const hipEvent_t* e = &ev_[(size_t)s * kEvPerStep];
            for (int p = 0; p < 4; ++p) {
                float ms = 0.f;
                SED_HIP(hipEventElapsedTime(&ms, e[p], e[p + 1]));
                ev_ms_[p] += ms;
            }
            if (s + 1 < ev_n_) {
                float ms = 0.f;
                SED_HIP(hipEventElapsedTime(&ms, e[4], e[kEvPerStep]));
                ev_ms_[4] += ms;
            }
        }
        ev_n_ = 0;
        return SED_OK;
    }
    static double ms_since(std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // One Gram-Schmidt pass: h = V[lo..j]^H w (and |w|^2), then w - V h into v_{j+1},
    // lo = *lo (0 without a pointer), w = the matvec output (w_next = 0) or v_{j+1}
    // itself (w_next = 1, in place).  With the matrix-core kernels both kernel families
    // are launched and the range size picks the one that works (pass_range).
    int launch_orth_pass(double2* h, const int* lo, int w_next, const int* gate) {
        if (mfma_) {
            hipLaunchKernelGGL(k_proj_mfma<T>, dim3(G_), dim3(256), 0, st_, V_, (int64_t)n_, step_, w_, part_, gate,
                               lo, w_next, 1);
            hipLaunchKernelGGL((k_dots_few<T, FewRows<T>::dots>), dim3(G_), dim3(kRedBlock), 0, st_, V_, (int64_t)n_,
                               step_, w_, part_, gate, lo, w_next);
        } else {
            hipLaunchKernelGGL(k_dots<T>, dim3(G_), dim3(kRedBlock), 0, st_, V_, (int64_t)n_, step_, w_, part_, gate,
                               lo, w_next, 0);
        }
        hipLaunchKernelGGL(k_reduce, dim3(m_ + 2), dim3(kRedBlock), 0, st_, part_, G_, step_, h, gate, lo);
        if (mfma_) {
            hipLaunchKernelGGL(k_orth_update_mfma<T>, dim3(kUpdGrid), dim3(256), 0, st_, V_, (int64_t)n_, step_, h, w_,
                               1, normpart_, gate, lo, w_next, 1);
            if constexpr (FewRows<T>::update > 1)
                hipLaunchKernelGGL((k_update_few<T, FewRows<T>::update>), dim3(kUpdGrid), dim3(256), 0, st_, V_,
                                   (int64_t)n_, step_, h, w_, 1, normpart_, gate, lo, w_next);
            else
                hipLaunchKernelGGL(k_update<T>, dim3(kUpdGrid), dim3(256), 0, st_, V_, (int64_t)n_, step_, h, w_, 1,
                                   normpart_, gate, lo, w_next, 2);
        } else {
            hipLaunchKernelGGL(k_update<T>, dim3(kUpdGrid), dim3(256), 0, st_, V_, (int64_t)n_, step_, h, w_, 1,
                               normpart_, gate, lo, w_next, 0);
        }
        SED_HIP(hipGetLastError());
        return SED_OK;
    }

    // a step through the workspace's captured graph (captured on first use for this
    // sector / precision / basis size)
    int step(bool graph) {
        if (!graph) return launch_step();
        GraphKey key;
        key.sec = sec_;
        key.version = sec_->version;
        key.prec = VecTraits<T>::fp64 ? 0 : 1;
        key.m = m_;
        key.shape = sec_->mv_shape[key.prec];
        key.chunk = sec_->mv_chunk[key.prec];
        key.reorth = mfma_ ? 1 : 0;
        key.n = n_;
        key.ptrs[0] = V_;
        key.ptrs[1] = w_;
        key.ptrs[2] = xh_;
        key.ptrs[3] = part_;
        key.ptrs[4] = ws_.state.p;
        if (!ws_.graph_exec || !(ws_.graph_key == key)) {
            ws_.drop_graph();
            hipGraph_t g = nullptr;
            SED_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
            const int rc = launch_step();
            hipError_t e = hipStreamEndCapture(st_, &g);
            if (rc != SED_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(SED_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&ws_.graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ws_.graph_exec = nullptr;
                return fail(SED_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            ws_.graph_key = key;
        }
        SED_HIP(hipGraphLaunch(ws_.graph_exec, st_));
        return SED_OK;
    }

    int normalize_start(int j) {
        hipLaunchKernelGGL(k_normsq<T>, dim3(kUpdGrid), dim3(256), 0, st_, vec(j), (int64_t)n_, normpart_);
        hipLaunchKernelGGL(k_start_finalize, dim3(1), dim3(kRedBlock), 0, st_, step_, j, normpart_, kUpdGrid, scale_);
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n_)), dim3(256), 0, st_, V_, (int64_t)n_, step_, scale_,
                           sec_->d_stab.p, xh_);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }
    // after a thick restart to kk kept Ritz vectors: their Ritz values on the diagonal of
    // T, no coupling among them, the arrowhead T[kk][i]; step kk's first pass spans the
    // whole basis (the arrowhead), and steps kk and kk + 1 run the second pass.
    int pro_restart(int kk, const std::vector<double>& theta, const std::vector<double>& arrow) {
        std::vector<double> zeros((size_t)kk, 0.0);
        const int hdr[3] = {0, pro_on_ ? 2 : 0, kk};   // lo, force, kk
        SED_HIP(hipMemcpyAsync(alpha_, theta.data(), (size_t)kk * sizeof(double), hipMemcpyHostToDevice, st_));
        SED_HIP(hipMemcpyAsync(beta_, zeros.data(), (size_t)kk * sizeof(double), hipMemcpyHostToDevice, st_));
        SED_HIP(hipMemcpyAsync(pro_->arrow, arrow.data(), (size_t)kk * sizeof(double), hipMemcpyHostToDevice, st_));
        SED_HIP(hipMemcpyAsync(&pro_->lo, hdr, sizeof(hdr), hipMemcpyHostToDevice, st_));
        SED_HIP(hipStreamSynchronize(st_));  // host temporaries
        return SED_OK;
    }
    // after a restart: step = j and xhat from v_j (v_j already unit norm)
    int set_step_and_xhat(int j) {
        const double one = 1.0;
        SED_HIP(hipMemcpyAsync(step_, &j, sizeof(int), hipMemcpyHostToDevice, st_));
        SED_HIP(hipMemcpyAsync(scale_, &one, sizeof(double), hipMemcpyHostToDevice, st_));
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n_)), dim3(256), 0, st_, V_, (int64_t)n_, step_, scale_,
                           sec_->d_stab.p, xh_);
        SED_HIP(hipGetLastError());
        SED_HIP(hipStreamSynchronize(st_));  // the host values above are stack temporaries
        return SED_OK;
    }
    int ritz(int m, const std::vector<double>& S, int kk) {
        SED_HIP(hipMemcpyAsync(ws_.S.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        if (mfma_) {
            SED_HIP(launch_rotate<T>(st_, V_, (int64_t)n_, m, ws_.S.p, kk, (int64_t)n_));
        } else {
            const size_t lds = (((size_t)m * kk * 8 + 15) & ~(size_t)15) + (size_t)m * kRitzRows * sizeof(T);
            const unsigned g = (unsigned)((n_ + kRitzRows - 1) / kRitzRows);
            hipLaunchKernelGGL(k_ritz<T>, dim3(g), dim3(kRitzRows), lds, st_, V_, (int64_t)n_, m, ws_.S.p, kk,
                               (int64_t)n_);
        }
        SED_HIP(hipGetLastError());
        SED_HIP(hipStreamSynchronize(st_));  // S is host-pageable
        return SED_OK;
    }
    static void eig(int jm, const std::vector<double>& T_, std::vector<double>& theta, std::vector<double>& S) {
        const int m = (int)std::lround(std::sqrt((double)T_.size()));
        std::vector<double> a((size_t)jm * jm);
        for (int i = 0; i < jm; ++i)
            for (int j = 0; j < jm; ++j) a[(size_t)i * jm + j] = T_[(size_t)i * m + j];
        jacobi_eigh(jm, a, theta, S);
    }
    static void eig_last(int jm, const std::vector<double>& T_, std::vector<double>& theta, std::vector<double>& last) {
        const int m = (int)std::lround(std::sqrt((double)T_.size()));
        std::vector<double> a((size_t)jm * jm);
        for (int i = 0; i < jm; ++i)
            for (int j = 0; j < jm; ++j) a[(size_t)i * jm + j] = T_[(size_t)i * m + j];
        eigh_last_row(jm, a, theta, last);
    }

    sed_sector* sec_;
    int64_t n_;
    int m_;
    Workspace& ws_;
    hipStream_t st_;
    int G_ = 1;
    T* w_ = nullptr;
    T* xh_ = nullptr;
    double2* part_ = nullptr;
    double* normpart_ = nullptr;
    double2 *h1_ = nullptr, *h2_ = nullptr;
    double *alpha_ = nullptr, *beta_ = nullptr, *scale_ = nullptr;
    int *step_ = nullptr, *flag_ = nullptr, *gate_ = nullptr;
    ProState* pro_ = nullptr;
    bool mfma_ = true;
    bool pro_on_ = true;
    bool ev_on_ = false;
    std::vector<hipEvent_t> ev_;
    int ev_n_ = 0;
    double ev_ms_[5] = {0, 0, 0, 0, 0};
    double ev_restart_ms_ = 0, ev_host_ms_ = 0;
};

int64_t solve_footprint(const sed_sector* sec, int nev, int precision) {
    const int64_t n = sec->dim;
    if (n <= 500) return n * n * 16 + n * 16 * std::max(nev, 1);
    const int m = krylov_dim(n, nev);
    int64_t b = (int64_t)(m + 1) * n * 16 + 2 * n * 16 + n * 16;          // FP64 basis, w, xhat, v0
    if (precision != SED_PREC_FP64) b += (int64_t)(m + 1) * n * 8 + 2 * n * 8;  // FP32 basis, w, xhat
    b += (int64_t)(m + 2) * kRedGrid * 16 + (int64_t)(m + 1) * (m + 1) * 8;
    b += n * 16 * nev;                                                    // eigenvectors
    return b;
}

// Dense path (dim <= 500), as the reference's n <= 500 branch.
static int solve_dense(sed_sector* sec, int nev, std::vector<double>& evals, std::vector<cd>* evecs, Workspace& ws) {
    const int64_t n = sec->dim;
    hipStream_t s = ws.stream;
    SED_TRY(grow(ws.V64, (size_t)2 * n * n));
    SED_TRY(launch_dense(sec, ws.V64.p, s));
    std::vector<cd> A((size_t)n * n);
    SED_HIP(hipMemcpyAsync(A.data(), ws.V64.p, (size_t)n * n * 16, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    // device wrote row-major H'[r][c]; hermitian_eigh takes column-major
    std::vector<cd> Ac((size_t)n * n);
    for (int64_t r = 0; r < n; ++r)
        for (int64_t c = 0; c < n; ++c) Ac[(size_t)c * n + r] = A[(size_t)r * n + c];
    for (int64_t r = 0; r < n; ++r)
        for (int64_t c = 0; c < r; ++c) {
            const cd avg = 0.5 * (Ac[(size_t)c * n + r] + std::conj(Ac[(size_t)r * n + c]));
            Ac[(size_t)c * n + r] = avg;
            Ac[(size_t)r * n + c] = std::conj(avg);
        }
    for (int64_t r = 0; r < n; ++r) Ac[(size_t)r * n + r] = cd(Ac[(size_t)r * n + r].real(), 0.0);
    std::vector<double> ev;
    const int k = (int)std::min<int64_t>(nev, n);
    hermitian_eigh((int)n, Ac, ev, evecs, k);
    evals.assign(ev.begin(), ev.begin() + k);
    return SED_OK;
}

// The convergence estimates beta |s_i| hold for the Lanczos relation of a semi-orthogonal
// basis.  With many kept Ritz vectors the relation of a partially re-orthogonalised run can
// be off by more than tol (true residuals up to 1e-7 measured on a spin-degenerate Hubbard
// sector at nev = 20; within tol at nev <= 5), so the FP64 phase's pairs are checked with
// one FP64 matvec each.  If one exceeds tol, the FP64 solve is repeated with a pass over the
// whole basis at every step, started from the sum of the Ritz vectors (from the seeded random
// vector when the wanted values are degenerate).  On return o holds the final pairs (true
// residuals when they were checked and passed) and vecs the device address of their vectors.
static int verify_fp64(sed_sector* sec, int m, int nev, double tol, uint64_t seed, int maxiter, Workspace& ws,
                       ThickLanczos<double2>& L, SolveOut& o, sed_solve_stats& st, double2*& vecs) {
    vecs = L.V_;
    if (!L.pro_on() || o.converged < nev) return SED_OK;
    const int nk = (int)o.evals.size();
    std::vector<double> res;
    SED_TRY(L.true_residuals(nk, o.evals, res));
    st.matvecs += nk;
    double rmax = 0.0;
    for (double r : res) rmax = std::max(rmax, r);
    if (rmax <= tol) {
        o.resid = res;
        return SED_OK;
    }
    const int64_t n = sec->dim;
    hipStream_t s = ws.stream;
    ThickLanczos<double2> F(sec, m, ws);
    F.disable_pro();
    if (fp32_cluster(o.evals, nev)) {
        SED_TRY(F.init());
        SED_TRY(F.start_random(seed));
    } else {
        SED_TRY(grow(ws.v0, (size_t)2 * n));
        double2* v0 = reinterpret_cast<double2*>(ws.v0.p);
        hipLaunchKernelGGL(k_sum_vectors<double2>, dim3(grid_for(n)), dim3(256), 0, s, L.V_, (int64_t)n, L.nritz_, v0,
                           (int64_t)n);
        SED_HIP(hipGetLastError());
        SED_TRY(F.init());
        SED_TRY(F.start_from(v0));
    }
    // the first run's work is counted here, the repeated run's by the caller
    st.repeated = 1;
    st.iters_fp64 += o.iters;
    st.matvecs += o.matvecs;
    st.restarts += o.restarts;
    st.full_orth_steps += o.full_steps;
    SED_TRY(F.run(nev, tol, maxiter, o));
    vecs = F.V_;
    return SED_OK;
}

static void fill_resid(sed_solve_stats& st, const SolveOut& o) {
    for (double r : o.resid) st.residual_max = std::max(st.residual_max, r);
}

// Full solve on ws.stream.  evals ascending; eigenvectors (dim x nev, primed basis,
// complex double) into out_vecs when non-null.
int solve_sector(sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter, int precision,
                 std::vector<double>& evals, DevBuf<double2>* out_vecs, sed_solve_stats* stats, Workspace& ws) {
    const auto t0 = std::chrono::steady_clock::now();
    sed_solve_stats st{};
    hipStream_t s = ws.stream;
    const int64_t n = sec->dim;
    auto finish = [&](int rc) {
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return rc;
    };
    if (n == 0) {
        evals.clear();
        return finish(SED_OK);
    }
    if (n <= 500) {
        std::vector<cd> X;
        SED_TRY(solve_dense(sec, nev, evals, out_vecs ? &X : nullptr, ws));
        if (out_vecs) {
            SED_TRY(out_vecs->alloc((size_t)n * evals.size()));
            SED_HIP(hipMemcpyAsync(out_vecs->p, X.data(), X.size() * 16, hipMemcpyHostToDevice, s));
            SED_HIP(hipStreamSynchronize(s));
        }
        st.converged = (int)evals.size();
        st.matvecs = n;
        return finish(SED_OK);
    }
    const int m = krylov_dim(n, nev);
    if (nev > m - 2) return fail(SED_EINVAL, "nev too large for the Krylov basis");
    SolveOut o;
    if (precision == SED_PREC_FP32 || precision == SED_PREC_FP32_REFINE) {
        ThickLanczos<float2> L32(sec, m, ws);
        if (nev > pro_max_nev()) L32.disable_pro();
        SED_TRY(L32.init());
        SED_TRY(L32.start_random(seed));
        // FP32 vectors resolve residuals down to ~1e-6 |H|; stop well above that floor.
        const double tol32 = precision == SED_PREC_FP32 ? std::max(tol, 1e-5) : 1e-4;
        SED_TRY(L32.run(nev, tol32, maxiter, o));
        st.iters_fp32 = o.iters;
        st.matvecs += o.matvecs;
        st.restarts += o.restarts;
        st.full_orth_steps += o.full_steps;
        if (precision == SED_PREC_FP32) {
            evals = o.evals;
            st.converged = o.converged;
            fill_resid(st, o);
            if (out_vecs) {
                SED_TRY(out_vecs->alloc((size_t)n * o.evals.size()));
                hipLaunchKernelGGL((k_convert<float2, double2>), dim3(grid_for(n * o.evals.size())), dim3(256), 0, s,
                                   L32.V_, out_vecs->p, (int64_t)(n * o.evals.size()));
                SED_HIP(hipGetLastError());
            }
            SED_HIP(hipStreamSynchronize(s));
            if (st.converged < nev) return finish(fail(SED_ENOCONV, "Lanczos (FP32) did not converge the requested eigenpairs"));
            return finish(SED_OK);
        }
        // FP64 refinement restarted from the sum of the FP32 Ritz vectors (from the seeded
        // random vector when the wanted FP32 Ritz values are degenerate)
        ThickLanczos<double2> L64(sec, m, ws);
        if (nev > pro_max_nev()) L64.disable_pro();
        if (fp32_cluster(o.evals, nev)) {
            SED_TRY(L64.init());
            SED_TRY(L64.start_random(seed));
        } else {
            SED_TRY(grow(ws.v0, (size_t)2 * n));
            double2* v0 = reinterpret_cast<double2*>(ws.v0.p);
            hipLaunchKernelGGL(k_sum_vectors<float2>, dim3(grid_for(n)), dim3(256), 0, s, L32.V_, (int64_t)n,
                               L32.nritz_, v0, (int64_t)n);
            SED_HIP(hipGetLastError());
            SED_TRY(L64.init());
            SED_TRY(L64.start_from(v0));
        }
        SED_TRY(L64.run(nev, tol, maxiter, o));
        double2* vecs = nullptr;
        SED_TRY(verify_fp64(sec, m, nev, tol, seed, maxiter, ws, L64, o, st, vecs));
        st.iters_fp64 += o.iters;
        if (out_vecs) {
            SED_TRY(out_vecs->alloc((size_t)n * o.evals.size()));
            SED_HIP(hipMemcpyAsync(out_vecs->p, vecs, (size_t)n * o.evals.size() * 16, hipMemcpyDeviceToDevice, s));
        }
    } else {
        ThickLanczos<double2> L64(sec, m, ws);
        if (nev > pro_max_nev()) L64.disable_pro();
        SED_TRY(L64.init());
        SED_TRY(L64.start_random(seed));
        SED_TRY(L64.run(nev, tol, maxiter, o));
        double2* vecs = nullptr;
        SED_TRY(verify_fp64(sec, m, nev, tol, seed, maxiter, ws, L64, o, st, vecs));
        st.iters_fp64 += o.iters;
        if (out_vecs) {
            SED_TRY(out_vecs->alloc((size_t)n * o.evals.size()));
            SED_HIP(hipMemcpyAsync(out_vecs->p, vecs, (size_t)n * o.evals.size() * 16, hipMemcpyDeviceToDevice, s));
        }
    }
    st.matvecs += o.matvecs;
    st.restarts += o.restarts;
    st.full_orth_steps += o.full_steps;
    evals = o.evals;
    st.converged = o.converged;
    fill_resid(st, o);
    SED_HIP(hipStreamSynchronize(s));
    if (st.converged < nev)
        return finish(fail(SED_ENOCONV, "Lanczos did not converge the requested " + std::to_string(nev) + " eigenpairs"));
    return finish(SED_OK);
}

}  // namespace sed
