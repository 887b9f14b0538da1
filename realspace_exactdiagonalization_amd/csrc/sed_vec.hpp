// Device helpers of the Lanczos vector kernels (complex FP32/FP64 vectors with FP64
// accumulation), shared by the per-sector solver and the packed engine.
#pragma once

#include <algorithm>
#include <cstdint>

#include "sed_device.hpp"

namespace sed {

constexpr int kMaxBasis = 64;   // largest Krylov basis (m)
// Second classical Gram-Schmidt pass on cancellation (Daniel-Gragg-Kaufman-Stewart test,
// |w'|^2 <= eta^2 |w|^2).  After a first pass over the whole basis eta = 1/sqrt(2), the
// "twice is enough" bound: a single classical pass against a basis that is orthonormal only
// to some level delta leaves overlaps of about delta |w|/|w'| in the new vector, and with
// eta = 0.1 those compound from step to step and over thick restarts (on a spin-degenerate
// Hubbard sector with nev = 20 the basis lost orthogonality completely within six restarts).
// After the local first pass (v_{j-1}, v_j only) large cancellation is the Lanczos
// recurrence itself, so there eta = 0.1 only guards against a pass that cancelled almost
// all of w; the partial re-orthogonalisation estimates decide the rest.
constexpr double kReorthEta2 = 0.01;
constexpr double kReorthEta2Whole = 0.5;

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

// ---------------------------------------------------------------------------
// Partial re-orthogonalisation (Simon): state in device memory, one per Krylov process.
// omega[r % 3][k] estimates |v_r^H v_k| for the rows r = j-1, j, j+1 around step j; T is
// the projected matrix: alpha[k] on the diagonal (the kept Ritz values after a thick
// restart), beta[k] = T[k+1][k] for k >= kk, and the restart's arrowhead
// T[kk][k] = arrow[k] for k < kk.
// ---------------------------------------------------------------------------
struct ProState {
    int lo;        // first basis vector of the current step's first pass
    int force;     // steps, from the current one, that run the second pass in any case
    int kk;        // kept Ritz vectors of this cycle (the arrowhead row)
    int nfull;     // steps with a pass over the whole basis
    int enabled;   // 0: every pass spans the whole basis
    int pad[3];
    double eps;     // unit roundoff of the stored vectors
    double delta;   // semi-orthogonality threshold, sqrt(eps)
    double anorm;   // running estimate of |T|
    double pad2;
    double arrow[kMaxBasis];
    double omega[3][kMaxBasis + 2];
};

// Largest nev solved with partial re-orthogonalisation; more wanted pairs get a pass over
// the whole basis at every step.  With many kept Ritz vectors after a thick restart, their
// Ritz relation holds only to the previous basis's orthogonality level (up to sqrt(eps)),
// which the recurrence does not model: in the model of the step driven by pro_decide the
// true residuals of the converged pairs stayed within 1e-10 for nev <= 5 and reached
// 1.4e-10 (nev = 6) to 1.4e-7 (nev = 20) on a spin-degenerate Hubbard sector.
constexpr int kProMaxNev = 5;

// Start of a Krylov process (step 0): first passes span the basis until step 2.
__host__ __device__ inline void pro_start(ProState* pro, bool fp64, bool enabled) {
    pro->lo = 0;
    pro->force = 0;
    pro->kk = 0;
    pro->enabled = enabled ? 1 : 0;
    pro->eps = fp64 ? 1.1102230246251565e-16 : 5.9604644775390625e-08;   // 2^-53, 2^-24
    pro->delta = sqrt(pro->eps);
    pro->anorm = 0.0;
}

// After the first Gram-Schmidt pass of step j (over v_lo..v_j, lo = pro->lo): whether the
// second pass, over the whole basis, must run.  It runs on cancellation (the test above),
// when the step is forced, or
// when the first pass was local (lo > 0) and the recurrence
//   beta_j omega_{j+1,k} = (T omega_j)_k - beta_{j-1} omega_{j-1,k} - alpha_j omega_{j,k}
//                          +- eps |T|   (rounding of the step)
// puts some |omega_{j+1,k}|, k <= j-2, above sqrt(eps).  The step after such a step is
// forced too (v_j itself is only semi-orthogonal, and its large coefficient in the next
// vector carries that into it), as are the two steps after a thick restart.  A forced
// step keeps its local first pass, which removes the large components, so that the
// second pass over the semi-orthogonal basis leaves eps-level overlaps.  After a second
// pass (or a first pass over an orthonormal basis, steps 0 and 1) omega_{j+1,.} is reset
// to eps; the two vectors a local pass orthogonalises against get eps.
// nrm2 = |w'|^2 after the first pass, nw = |w|^2 before it, aj = Re h_j of the first pass.
// One thread (host-callable, so the decision logic can be driven from a CPU model).
__host__ __device__ inline bool pro_decide(ProState* pro, int j, double nrm2, double nw, double aj,
                                  const double* alpha, const double* beta) {
    bool need = !(nrm2 > (pro->lo == 0 ? kReorthEta2Whole : kReorthEta2) * nw);
    if (pro->force > 0) {
        need = true;
        --pro->force;
    }
    const int lo = pro->lo;
    double* wn = pro->omega[(j + 1) % 3];
    const double eps = pro->eps;
    if (pro->enabled && lo > 0 && !need) {
        const double* wj = pro->omega[j % 3];
        const double* wm = pro->omega[(j + 2) % 3];   // row j - 1
        const int kk = pro->kk;
        const double bj = sqrt(nrm2);
        const double bjm = beta[j - 1];
        const double noise = eps * fmax(pro->anorm, fabs(aj) + bj + bjm);
        double mx = 0.0;
        for (int k = 0; k <= j - 2; ++k) {
            double tw = alpha[k] * wj[k];
            if (k < kk) {
                tw += pro->arrow[k] * wj[kk];
            } else {
                tw += beta[k] * wj[k + 1];
                if (k > kk) {
                    tw += beta[k - 1] * wj[k - 1];
                } else {
                    for (int b = 0; b < kk; ++b) tw += pro->arrow[b] * wj[b];
                }
            }
            double x = tw - bjm * wm[k] - aj * wj[k];
            x += x < 0.0 ? -noise : noise;
            x /= bj;
            wn[k] = x;
            mx = fmax(mx, fabs(x));
        }
        if (!(mx <= pro->delta)) {   // NaN (bj = 0) included
            need = true;
            pro->force = pro->force > 1 ? pro->force : 1;
        } else {
            wn[j - 1] = eps;
            wn[j] = eps;
        }
    }
    if (need || lo == 0) {
        for (int k = 0; k <= j; ++k) wn[k] = eps;
        ++pro->nfull;
    }
    wn[j + 1] = 1.0;
    return need;
}

// pro_decide spread over the NT threads of one workgroup, all of which call it with the
// same arguments: each k of the omega recurrence runs on its own thread with pro_decide's
// arithmetic, and the maximum over k is order-free (fmax), so the state left and the
// decision returned (to every thread) are pro_decide's, bit for bit.  The serial form
// is a chain of dependent loads of alpha, beta, omega and the arrowhead on one thread.
// red: NT / kWave doubles of shared memory; sh: one shared int.
template <int NT>
__device__ inline bool pro_decide_wg(ProState* pro, int j, double nrm2, double nw, double aj, const double* alpha,
                                     const double* beta, double* red, int* sh) {
    const int lo = pro->lo;
    const int force = pro->force;
    const bool need0 = !(nrm2 > (lo == 0 ? kReorthEta2Whole : kReorthEta2) * nw) || force > 0;
    double* wn = pro->omega[(j + 1) % 3];
    const double eps = pro->eps;
    const bool rec = pro->enabled && lo > 0 && !need0;
    double mx = 0.0;
    if (rec) {
        const double* wj = pro->omega[j % 3];
        const double* wm = pro->omega[(j + 2) % 3];   // row j - 1
        const int kk = pro->kk;
        const double bj = sqrt(nrm2);
        const double bjm = beta[j - 1];
        const double noise = eps * fmax(pro->anorm, fabs(aj) + bj + bjm);
        for (int k = threadIdx.x; k <= j - 2; k += NT) {
            double tw = alpha[k] * wj[k];
            if (k < kk) {
                tw += pro->arrow[k] * wj[kk];
            } else {
                tw += beta[k] * wj[k + 1];
                if (k > kk) {
                    tw += beta[k - 1] * wj[k - 1];
                } else {
                    for (int b = 0; b < kk; ++b) tw += pro->arrow[b] * wj[b];
                }
            }
            double x = tw - bjm * wm[k] - aj * wj[k];
            x += x < 0.0 ? -noise : noise;
            x /= bj;
            wn[k] = x;
            mx = fmax(mx, fabs(x));
        }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < NT / kWave; ++k) m = fmax(m, red[k]);
        bool need = need0;
        if (force > 0) --pro->force;
        if (rec) {
            if (!(m <= pro->delta)) {   // NaN (bj = 0) included
                need = true;
                pro->force = pro->force > 1 ? pro->force : 1;
            } else {
                wn[j - 1] = eps;
                wn[j] = eps;
            }
        }
        if (need || lo == 0) ++pro->nfull;
        wn[j + 1] = 1.0;
        *sh = need ? 1 : 0;
    }
    __syncthreads();
    const bool need = *sh != 0;
    if (need || lo == 0)
        for (int k = threadIdx.x; k <= j; k += NT) wn[k] = eps;
    __syncthreads();
    return need;
}

// End of step j (alpha_j = a, beta_j = b final): the next step's first pass covers
// v_j, v_{j+1} (the whole basis when partial re-orthogonalisation is off).  One thread.
__host__ __device__ inline void pro_advance(ProState* pro, int j, double a, double b, const double* beta) {
    pro->anorm = fmax(pro->anorm, fabs(a) + b + (j > 0 ? beta[j - 1] : 0.0));
    pro->lo = pro->enabled ? j : 0;
}

// After a thick restart to kk kept Ritz vectors (values theta, arrowhead couplings arrow):
// the diagonal of T, no coupling among them, step kk's first pass over the whole basis
// (the arrowhead), and the second pass at steps kk and kk + 1.  One thread.
__host__ __device__ inline void pro_restart(ProState* pro, int kk, const double* theta, const double* arrow,
                                            double* alpha, double* beta) {
    for (int i = 0; i < kk; ++i) {
        alpha[i] = theta[i];
        beta[i] = 0.0;
        pro->arrow[i] = arrow[i];
    }
    pro->lo = 0;
    pro->force = pro->enabled ? 2 : 0;
    pro->kk = kk;
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

// ---------------------------------------------------------------------------
// Gram-Schmidt passes inside one workgroup (the packed engine, and the fused step tail of
// sectors small enough for one workgroup)
// ---------------------------------------------------------------------------
constexpr int kBlockDots = 4;   // basis vectors per sweep of the dot products

// Sum of v over a workgroup of NT threads, returned to every thread (fixed order).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NT / kWave; ++k) s += red[k];
    __syncthreads();
    return s;
}

// One workgroup of NT threads: h[i] = V_i^dag w for i < nv and *nrm = |w|^2, kBlockDots
// vectors per sweep over the rows (results in LDS for every thread).
template <typename T, int NT>
__device__ __forceinline__ void block_dots(const T* __restrict__ V, int64_t n, int nv, const T* __restrict__ w, double2* h,
                          double* nrm, double2 (*red)[NT / kWave]) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (int g0 = 0; g0 <= nv; g0 += kBlockDots) {
        double2 acc[kBlockDots];
#pragma unroll
        for (int q = 0; q < kBlockDots; ++q) acc[q] = make_double2(0.0, 0.0);
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            const double2 wr = to_d2(w[r]);
            double2 v[kBlockDots];
#pragma unroll
            for (int q = 0; q < kBlockDots; ++q)
                v[q] = g0 + q < nv ? to_d2(V[(int64_t)(g0 + q) * n + r]) : make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < kBlockDots; ++q) {
                if (g0 + q < nv) {
                    acc[q].x += v[q].x * wr.x + v[q].y * wr.y;
                    acc[q].y += v[q].x * wr.y - v[q].y * wr.x;
                } else if (g0 + q == nv) {
                    acc[q].x += wr.x * wr.x + wr.y * wr.y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kBlockDots; ++q) {
            const double2 sm = wave_sum_d2(acc[q]);
            if (lane == 0) red[q][wv] = sm;
        }
        __syncthreads();
        if (threadIdx.x < kBlockDots) {
            const int i = g0 + threadIdx.x;
            double2 sm = make_double2(0.0, 0.0);
            for (int k = 0; k < NT / kWave; ++k) {
                sm.x += red[threadIdx.x][k].x;
                sm.y += red[threadIdx.x][k].y;
            }
            if (i < nv) h[i] = sm;
            else if (i == nv) *nrm = sm.x;
        }
        __syncthreads();
    }
}

// One workgroup: w <- w - sum_{i<nv} V_i h_i (four basis vectors per group of loads);
// returns |w|^2 to every thread.
template <typename T, int NT>
__device__ __forceinline__ double block_update(const T* __restrict__ V, int64_t n, int nv, const double2* h, T* w, double* red) {
    double nrm = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += NT) {
        double2 a = to_d2(w[r]);
        int i = 0;
        for (; i + 4 <= nv; i += 4) {
            double2 v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = to_d2(V[(int64_t)(i + q) * n + r]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 c = h[i + q];
                a.x -= v[q].x * c.x - v[q].y * c.y;
                a.y -= v[q].x * c.y + v[q].y * c.x;
            }
        }
        for (; i < nv; ++i) {
            const double2 v = to_d2(V[(int64_t)i * n + r]);
            const double2 c = h[i];
            a.x -= v.x * c.x - v.y * c.y;
            a.y -= v.x * c.y + v.y * c.x;
        }
        const T o = from_d2<T>(a);
        w[r] = o;
        const double2 od = to_d2(o);
        nrm += od.x * od.x + od.y * od.y;
    }
    return block_sum<NT>(nrm, red);
}

}  // namespace sed
