// Lanczos eigensolver around the sector H|psi> on gfx950.
//
// Replaces diagonalize_block_matrixfree (src/symmetry_resolved_ed.jl:998-1030), which
// hands the operator to KrylovKit.eigsolve(H, x0, nev, :SR; tol=1e-10, maxiter=300)
// (Lanczos with restarts and full re-orthogonalisation) and solves sectors of
// dimension <= 500 densely.  Here:
//   * thick-restart Lanczos (Wu & Simon) with a basis of at most m vectors resident in
//     HBM; every new vector is orthogonalised twice against the whole basis
//     (classical Gram-Schmidt, "twice is enough"), and the alpha/beta recurrence and
//     every projection coefficient are carried in FP64 whatever the vector precision;
//   * all reductions run over a fixed grid in a fixed order (bitwise reproducible);
//   * the projected matrix is diagonalised on the host (order <= m);
//   * SED_PREC_FP32_REFINE runs the loop with complex-FP32 vectors and matvec, then
//     restarts an FP64 loop from the sum of the FP32 Ritz vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "sed_device.hpp"
#include "sed_internal.hpp"

namespace sed {

int launch_dense(sed_sector* sec, double* H_dev, hipStream_t s);

constexpr int kRedGrid = 512;     // fixed reduction grid (determinism)
constexpr int kRedBlock = 256;
constexpr int kDotGroup = 8;
constexpr int kRitzRows = 32;

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

// partial[i * gridDim + b] = sum over block b's rows of conj(V_i[r]) w[r], i < nv;
// partial[nv * gridDim + b].x = sum |w[r]|^2 when with_norm.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_dots(const T* __restrict__ V, int64_t ld, int nv,
                                                    const T* __restrict__ w, int64_t n, int with_norm,
                                                    double2* __restrict__ part) {
    __shared__ double2 red[kDotGroup][kRedBlock / kWave];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = std::min<int64_t>(n, r0 + chunk);
    const int total = nv + (with_norm ? 1 : 0);
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
                    const double2 v = to_d2(V[(int64_t)i * ld + r]);
                    acc[q].x += v.x * wr.x + v.y * wr.y;
                    acc[q].y += v.x * wr.y - v.y * wr.x;
                } else if (i == nv && with_norm) {
                    acc[q].x += wr.x * wr.x + wr.y * wr.y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kDotGroup; ++q) {
            const double2 s = wave_sum_d2(acc[q]);
            if (lane == 0) red[q][wv] = s;
        }
        __syncthreads();
        if (threadIdx.x < kDotGroup && g0 + (int)threadIdx.x < total) {
            double2 s = make_double2(0.0, 0.0);
            for (int i = 0; i < kRedBlock / kWave; ++i) {
                s.x += red[threadIdx.x][i].x;
                s.y += red[threadIdx.x][i].y;
            }
            part[(int64_t)(g0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

// out[i] = sum_b part[i * G + b], fixed order
__global__ void __launch_bounds__(kRedBlock) k_reduce(const double2* __restrict__ part, int G,
                                                      double2* __restrict__ out) {
    __shared__ double2 red[kRedBlock / kWave];
    const int i = blockIdx.x;
    double2 acc = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < G; b += kRedBlock) {
        acc.x += part[(int64_t)i * G + b].x;
        acc.y += part[(int64_t)i * G + b].y;
    }
    acc = wave_sum_d2(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 s = make_double2(0.0, 0.0);
        for (int k = 0; k < kRedBlock / kWave; ++k) {
            s.x += red[k].x;
            s.y += red[k].y;
        }
        out[i] = s;
    }
}

// out[r] = (w[r] - sum_{i<nv} V_i[r] h[i]) * scale ; out may alias w
template <typename T>
__global__ void __launch_bounds__(256) k_update(const T* __restrict__ V, int64_t ld, int nv,
                                                const double2* __restrict__ h, const T* w, int64_t n,
                                                T* out, double scale) {
    extern __shared__ double2 lh[];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) lh[i] = h[i];
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 a = to_d2(w[r]);
        for (int i = 0; i < nv; ++i) {
            const double2 v = to_d2(V[(int64_t)i * ld + r]);
            const double2 c = lh[i];
            a.x -= v.x * c.x - v.y * c.y;
            a.y -= v.x * c.y + v.y * c.x;
        }
        a.x *= scale;
        a.y *= scale;
        out[r] = from_d2<T>(a);
    }
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

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Seeded complex Gaussian, Re and Im ~ N(0, 1/2), a function of (seed, r) only.
template <typename T>
__global__ void __launch_bounds__(256) k_random(T* __restrict__ v, int64_t n, uint64_t seed) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = splitmix64(seed * 0x100000001b3ull + 2 * (uint64_t)r);
        const uint64_t b = splitmix64(seed * 0x100000001b3ull + 2 * (uint64_t)r + 1);
        const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740993.0);
        const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
        const double rad = sqrt(-log(u1));    // sqrt(-2 ln u1) * sqrt(1/2)
        const double th = 6.283185307179586 * u2;
        v[r] = from_d2<T>(make_double2(rad * cos(th), rad * sin(th)));
    }
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_convert(const Tin* __restrict__ a, Tout* __restrict__ b, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        b[r] = from_d2<Tout>(to_d2(a[r]));
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
int Workspace::ensure_host(size_t n_doubles) {
    if (hcoef_n >= n_doubles) return SED_OK;
    if (hcoef) (void)hipHostFree(hcoef);
    hcoef = nullptr;
    hipError_t e = hipHostMalloc((void**)&hcoef, n_doubles * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        hcoef = nullptr;
        hcoef_n = 0;
        return fail(SED_ENOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    hcoef_n = n_doubles;
    return SED_OK;
}

void Workspace::release() {
    V64.release(); w64.release(); xh64.release();
    V32.release(); w32.release(); xh32.release();
    part.release(); coef.release(); S.release(); v0.release();
    if (hcoef) (void)hipHostFree(hcoef);
    hcoef = nullptr;
    hcoef_n = 0;
}

template <typename T>
struct VecTraits;
template <>
struct VecTraits<double2> {
    static double* V(Workspace& w) { return w.V64.p; }
    static double* W(Workspace& w) { return w.w64.p; }
    static int reserve(Workspace& w, int64_t n, int m) {
        SED_TRY(grow(w.V64, (size_t)2 * (m + 1) * n));
        SED_TRY(grow(w.w64, (size_t)2 * n));
        SED_TRY(grow(w.xh64, (size_t)2 * n));
        return SED_OK;
    }
    static int matvec(sed_sector* s, const double2* x, double2* y, Workspace& w) {
        return launch_matvec64(s, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y), w.xh64.p,
                               w.stream);
    }
};
template <>
struct VecTraits<float2> {
    static float* V(Workspace& w) { return w.V32.p; }
    static float* W(Workspace& w) { return w.w32.p; }
    static int reserve(Workspace& w, int64_t n, int m) {
        SED_TRY(grow(w.V32, (size_t)2 * (m + 1) * n));
        SED_TRY(grow(w.w32, (size_t)2 * n));
        SED_TRY(grow(w.xh32, (size_t)2 * n));
        return SED_OK;
    }
    static int matvec(sed_sector* s, const float2* x, float2* y, Workspace& w) {
        return launch_matvec32(s, reinterpret_cast<const float*>(x), reinterpret_cast<float*>(y), w.xh32.p,
                               w.stream);
    }
};

struct SolveOut {
    std::vector<double> evals;
    std::vector<double> resid;
    int converged = 0;
    int iters = 0;
    int restarts = 0;
    int64_t matvecs = 0;
};

template <typename T>
class ThickLanczos {
   public:
    ThickLanczos(sed_sector* sec, int m, Workspace& ws) : sec_(sec), n_(sec->dim), m_(m), ws_(ws), st_(ws.stream) {}

    int init() {
        SED_TRY(VecTraits<T>::reserve(ws_, n_, m_));
        SED_TRY(grow(ws_.part, (size_t)2 * (m_ + 2) * kRedGrid));
        SED_TRY(grow(ws_.coef, (size_t)2 * (m_ + 2)));
        SED_TRY(grow(ws_.S, (size_t)(m_ + 1) * (m_ + 1)));
        SED_TRY(ws_.ensure_host((size_t)2 * (m_ + 2)));
        V_ = reinterpret_cast<T*>(VecTraits<T>::V(ws_));
        w_ = reinterpret_cast<T*>(VecTraits<T>::W(ws_));
        part_ = reinterpret_cast<double2*>(ws_.part.p);
        coef_ = reinterpret_cast<double2*>(ws_.coef.p);
        return SED_OK;
    }
    T* vec(int i) { return V_ + (size_t)i * n_; }

    int start_random(uint64_t seed) {
        hipLaunchKernelGGL(k_random<T>, dim3(grid_for(n_)), dim3(256), 0, st_, vec(0), n_, seed);
        SED_HIP(hipGetLastError());
        return normalize0();
    }
    int start_from(const double2* v0) {
        hipLaunchKernelGGL((k_convert<double2, T>), dim3(grid_for(n_)), dim3(256), 0, st_, v0, vec(0), n_);
        SED_HIP(hipGetLastError());
        return normalize0();
    }

    int run(int nev, double tol, int maxiter, SolveOut& out) {
        std::vector<double> T_((size_t)m_ * m_, 0.0);  // row-major projected matrix
        std::vector<double2> h1(m_ + 2), h2(m_ + 2);
        int k = 0;
        double beta = 0.0;
        out = SolveOut();
        std::vector<double> theta, Sm;
        int jm = 0;
        bool done = false, breakdown = false;
        while (!done) {
            for (int j = k; j < m_; ++j) {
                SED_TRY(VecTraits<T>::matvec(sec_, vec(j), w_, ws_));
                ++out.matvecs;
                ++out.iters;
                SED_TRY(dots(j + 1, w_, false, h1.data()));
                SED_TRY(update(j + 1, w_, w_, 1.0));
                SED_TRY(dots(j + 1, w_, true, h2.data()));
                const double nrm2 = h2[j + 1].x;
                double p2 = 0.0;
                for (int i = 0; i <= j; ++i) p2 += h2[i].x * h2[i].x + h2[i].y * h2[i].y;
                double b2 = nrm2 - p2;
                T_[(size_t)j * m_ + j] = h1[j].x + h2[j].x;
                if (!(b2 > 0.5 * nrm2)) {
                    // heavy cancellation: apply the second correction and measure directly
                    SED_TRY(update(j + 1, w_, w_, 1.0));
                    std::vector<double2> h3(2);
                    SED_TRY(dots(0, w_, true, h3.data()));
                    b2 = h3[0].x;
                    std::fill(h2.begin(), h2.end(), make_double2(0.0, 0.0));
                    SED_TRY(upload_coef(j + 1, h2.data()));
                }
                beta = std::sqrt(std::max(b2, 0.0));
                double tnorm = 0.0;
                for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, std::fabs(T_[(size_t)i * m_ + i]));
                jm = j + 1;
                if (beta <= 1e-13 * std::max(1.0, tnorm)) {
                    breakdown = true;
                    break;
                }
                // coef_ still holds h2: v_{j+1} = (w - V h2) / beta
                SED_TRY(update(j + 1, w_, vec(j + 1), 1.0 / beta));
                if (j + 1 < m_) {
                    T_[(size_t)(j + 1) * m_ + j] = beta;
                    T_[(size_t)j * m_ + j + 1] = beta;
                }
                const bool check = (j + 1 == m_) || ((j + 1 - k) % 8 == 0 && j + 1 >= nev + 2) ||
                                   out.iters >= maxiter;
                if (check) {
                    eig(jm, T_, theta, Sm);
                    int conv = 0;
                    for (int i = 0; i < std::min(nev, jm); ++i) {
                        const double res = beta * std::fabs(Sm[(size_t)(jm - 1) * jm + i]);
                        if (res <= tol) ++conv; else break;
                    }
                    if (conv >= nev || out.iters >= maxiter) {
                        done = true;
                        break;
                    }
                }
            }
            if (breakdown) {
                eig(jm, T_, theta, Sm);
                beta = 0.0;
                done = true;
            }
            if (done) break;
            // thick restart: kk Ritz vectors + the residual direction
            eig(jm, T_, theta, Sm);
            const int kk = std::min(m_ - 2, nev + std::max(4, (m_ - nev) / 3));
            std::vector<double> Sk((size_t)jm * kk);
            for (int jj = 0; jj < jm; ++jj)
                for (int i = 0; i < kk; ++i) Sk[(size_t)jj * kk + i] = Sm[(size_t)jj * jm + i];
            SED_TRY(ritz(jm, Sk, kk));
            SED_HIP(hipMemcpyAsync(vec(kk), vec(jm), (size_t)n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
            std::fill(T_.begin(), T_.end(), 0.0);
            for (int i = 0; i < kk; ++i) {
                T_[(size_t)i * m_ + i] = theta[i];
                const double sc = beta * Sm[(size_t)(jm - 1) * jm + i];
                T_[(size_t)kk * m_ + i] = sc;
                T_[(size_t)i * m_ + kk] = sc;
            }
            k = kk;
            ++out.restarts;
        }
        const int nk = std::min(nev, jm);
        out.evals.assign(theta.begin(), theta.begin() + nk);
        out.resid.resize(nk);
        out.converged = 0;
        for (int i = 0; i < nk; ++i) {
            out.resid[i] = beta * std::fabs(Sm[(size_t)(jm - 1) * jm + i]);
            if (out.resid[i] <= tol && out.converged == i) ++out.converged;
        }
        std::vector<double> Sk((size_t)jm * nk);
        for (int jj = 0; jj < jm; ++jj)
            for (int i = 0; i < nk; ++i) Sk[(size_t)jj * nk + i] = Sm[(size_t)jj * jm + i];
        SED_TRY(ritz(jm, Sk, nk));
        nritz_ = nk;
        SED_HIP(hipStreamSynchronize(st_));
        return SED_OK;
    }

    int nritz_ = 0;
    T* V_ = nullptr;

   private:
    int normalize0() {
        std::vector<double2> h(2);
        SED_TRY(dots(0, vec(0), true, h.data()));
        const double nrm = std::sqrt(h[0].x);
        if (!(nrm > 0)) return fail(SED_EINVAL, "zero start vector");
        hipLaunchKernelGGL(k_update<T>, dim3(grid_for(n_)), dim3(256), 16, st_, vec(0), (int64_t)n_, 0, coef_,
                           vec(0), (int64_t)n_, vec(0), 1.0 / nrm);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }
    int dots(int nv, const T* w, bool with_norm, double2* host_out) {
        const int total = nv + (with_norm ? 1 : 0);
        hipLaunchKernelGGL(k_dots<T>, dim3(kRedGrid), dim3(kRedBlock), 0, st_, V_, (int64_t)n_, nv, w, (int64_t)n_,
                           with_norm ? 1 : 0, part_);
        SED_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_reduce, dim3(total), dim3(kRedBlock), 0, st_, part_, kRedGrid, coef_);
        SED_HIP(hipGetLastError());
        SED_HIP(hipMemcpyAsync(ws_.hcoef, coef_, total * sizeof(double2), hipMemcpyDeviceToHost, st_));
        SED_HIP(hipStreamSynchronize(st_));
        std::memcpy(host_out, ws_.hcoef, total * sizeof(double2));
        return SED_OK;
    }
    int upload_coef(int nv, const double2* h) {
        std::memcpy(ws_.hcoef, h, nv * sizeof(double2));
        SED_HIP(hipMemcpyAsync(coef_, ws_.hcoef, nv * sizeof(double2), hipMemcpyHostToDevice, st_));
        return SED_OK;
    }
    // out <- (w - V coef_) * scale, coef_ already on device
    int update(int nv, const T* w, T* out, double scale) {
        hipLaunchKernelGGL(k_update<T>, dim3(grid_for(n_)), dim3(256), nv * sizeof(double2), st_, V_, (int64_t)n_,
                           nv, coef_, w, (int64_t)n_, out, scale);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }
    int ritz(int m, const std::vector<double>& S, int kk) {
        SED_HIP(hipMemcpyAsync(ws_.S.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        const size_t lds = (((size_t)m * kk * 8 + 15) & ~(size_t)15) + (size_t)m * kRitzRows * sizeof(T);
        const unsigned g = (unsigned)((n_ + kRitzRows - 1) / kRitzRows);
        hipLaunchKernelGGL(k_ritz<T>, dim3(g), dim3(kRitzRows), lds, st_, V_, (int64_t)n_, m, ws_.S.p, kk,
                           (int64_t)n_);
        SED_HIP(hipGetLastError());
        // S is host-pageable: make sure the copy has been consumed before it goes out of scope
        SED_HIP(hipStreamSynchronize(st_));
        return SED_OK;
    }
    static void eig(int jm, const std::vector<double>& T_, std::vector<double>& theta, std::vector<double>& S) {
        const int m = (int)std::lround(std::sqrt((double)T_.size()));
        std::vector<double> a((size_t)jm * jm);
        for (int i = 0; i < jm; ++i)
            for (int j = 0; j < jm; ++j) a[(size_t)i * jm + j] = T_[(size_t)i * m + j];
        jacobi_eigh(jm, a, theta, S);
    }

    sed_sector* sec_;
    int64_t n_;
    int m_;
    Workspace& ws_;
    hipStream_t st_;
    T* w_ = nullptr;
    double2* part_ = nullptr;
    double2* coef_ = nullptr;
};

static int krylov_dim(int64_t n, int nev) {
    int m = std::max(2 * nev + 24, 40);
    return (int)std::min<int64_t>(m, std::min<int64_t>(n - 1, 64));
}

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
    std::vector<cd> X;
    hermitian_eigh((int)n, Ac, ev, evecs ? &X : nullptr);
    const int k = (int)std::min<int64_t>(nev, n);
    evals.assign(ev.begin(), ev.begin() + k);
    if (evecs) {
        evecs->assign((size_t)n * k, cd(0.0));
        for (int j = 0; j < k; ++j)
            for (int64_t i = 0; i < n; ++i) (*evecs)[(size_t)j * n + i] = X[(size_t)j * n + i];
    }
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
        SED_TRY(L32.init());
        SED_TRY(L32.start_random(seed));
        // FP32 vectors resolve residuals down to ~1e-6 |H|; stop well above that floor.
        const double tol32 = precision == SED_PREC_FP32 ? std::max(tol, 1e-5) : 1e-4;
        SED_TRY(L32.run(nev, tol32, maxiter, o));
        st.iters_fp32 = o.iters;
        st.matvecs += o.matvecs;
        st.restarts += o.restarts;
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
        // FP64 refinement restarted from the sum of the FP32 Ritz vectors
        SED_TRY(grow(ws.v0, (size_t)2 * n));
        double2* v0 = reinterpret_cast<double2*>(ws.v0.p);
        hipLaunchKernelGGL(k_sum_vectors<float2>, dim3(grid_for(n)), dim3(256), 0, s, L32.V_, (int64_t)n, L32.nritz_,
                           v0, (int64_t)n);
        SED_HIP(hipGetLastError());
        ThickLanczos<double2> L64(sec, m, ws);
        SED_TRY(L64.init());
        SED_TRY(L64.start_from(v0));
        SED_TRY(L64.run(nev, tol, maxiter, o));
        st.iters_fp64 = o.iters;
        if (out_vecs) {
            SED_TRY(out_vecs->alloc((size_t)n * o.evals.size()));
            SED_HIP(hipMemcpyAsync(out_vecs->p, L64.V_, (size_t)n * o.evals.size() * 16, hipMemcpyDeviceToDevice, s));
        }
    } else {
        ThickLanczos<double2> L64(sec, m, ws);
        SED_TRY(L64.init());
        SED_TRY(L64.start_random(seed));
        SED_TRY(L64.run(nev, tol, maxiter, o));
        st.iters_fp64 = o.iters;
        if (out_vecs) {
            SED_TRY(out_vecs->alloc((size_t)n * o.evals.size()));
            SED_HIP(hipMemcpyAsync(out_vecs->p, L64.V_, (size_t)n * o.evals.size() * 16, hipMemcpyDeviceToDevice, s));
        }
    }
    st.matvecs += o.matvecs;
    st.restarts += o.restarts;
    evals = o.evals;
    st.converged = o.converged;
    fill_resid(st, o);
    SED_HIP(hipStreamSynchronize(s));
    if (st.converged < nev)
        return finish(fail(SED_ENOCONV, "Lanczos did not converge the requested " + std::to_string(nev) + " eigenpairs"));
    return finish(SED_OK);
}

}  // namespace sed
