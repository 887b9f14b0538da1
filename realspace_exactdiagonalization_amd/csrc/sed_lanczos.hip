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
#include <cstring>

#include "sed_device.hpp"
#include "sed_internal.hpp"

namespace sed {

int launch_dense(sed_sector* sec, double* H_dev, hipStream_t s);
int launch_matvec_xhat(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s);

constexpr int kRedGrid = 512;     // fixed reduction grid (determinism)
constexpr int kRedBlock = 256;
constexpr int kDotGroup = 8;
constexpr int kRitzRows = 32;
constexpr int kMaxBasis = 64;   // largest Krylov basis (m)
constexpr int kUpdGrid = 1024;  // fixed grid of the update / norm kernels

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


// ---- one Lanczos step, step index j = *step on the device ----------------------
// The step reads v_j through xhat = sqrt(stab) v_j (written by the previous step),
// w = H v_j, then two classical Gram-Schmidt passes against V[0..j], then
// v_{j+1} = w' / |w'| and its xhat.

// part[i * G + b] = sum over block b's rows of conj(V_i[r]) w[r] for i <= j, and
// part[(j+1) * G + b] = sum |w[r]|^2.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_dots(const T* __restrict__ V, int64_t n, const int* __restrict__ step,
                                                    const T* __restrict__ w, double2* __restrict__ part) {
    __shared__ double2 red[kDotGroup][kRedBlock / kWave];
    const int nv = *step + 1;
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

// out[i] = sum_b part[i * G + b] for i <= j + 1 (fixed order); launched with m + 2 blocks.
__global__ void __launch_bounds__(kRedBlock) k_reduce(const double2* __restrict__ part, int G,
                                                      const int* __restrict__ step, double2* __restrict__ out) {
    __shared__ double2 red[kRedBlock / kWave];
    const int i = blockIdx.x;
    if (i > *step + 1) return;
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

// out[r] = w[r] - sum_{i<=j} V_i[r] h[i]; out = w (to_next = 0) or V_{j+1} (to_next = 1).
// With to_next, each block also writes the partial |out|^2 of its rows (fixed grid).
template <typename T>
__global__ void __launch_bounds__(256) k_update(T* V, int64_t n, const int* __restrict__ step,
                                                const double2* __restrict__ h, T* w, int to_next,
                                                double* __restrict__ normpart) {
    __shared__ double2 lh[kMaxBasis + 2];
    __shared__ double red[256 / kWave];
    const int nv = *step + 1;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) lh[i] = h[i];
    __syncthreads();
    T* out = to_next ? V + (int64_t)nv * n : w;
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
    if (!to_next) return;
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int k = 0; k < 256 / kWave; ++k) sm += red[k];
        normpart[blockIdx.x] = sm;
    }
}

// Single block: |v_{j+1}|^2 from the update's partials, alpha_j, beta_j, 1/beta_j, and
// the step counter advances to j + 1.  A vanishing residual (invariant subspace) is
// flagged with the step at which it happened and scales the new vector by zero.
__global__ void __launch_bounds__(kRedBlock) k_finalize(int* step, const double2* __restrict__ h1,
                                                        const double2* __restrict__ h2,
                                                        const double* __restrict__ normpart, int G,
                                                        double* __restrict__ alpha, double* __restrict__ beta,
                                                        double* __restrict__ scale, int* __restrict__ flag) {
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

void Workspace::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
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
    int64_t matvecs = 0;
};

static bool graphs_enabled() {
    const char* e = std::getenv("SED_LANCZOS_GRAPH");
    return !(e && e[0] == '0');
}

// Device state block (doubles): h1 (2*(m+2)), h2 (2*(m+2)), alpha (m), beta (m), scale (1);
// ints after it: step, flag.
template <typename T>
class ThickLanczos {
   public:
    ThickLanczos(sed_sector* sec, int m, Workspace& ws) : sec_(sec), n_(sec->dim), m_(m), ws_(ws), st_(ws.stream) {}

    int init() {
        SED_TRY(VecTraits<T>::reserve(ws_, n_, m_));
        G_ = (int)std::max<int64_t>(1, std::min<int64_t>(kRedGrid, (n_ + kRedBlock - 1) / kRedBlock));
        SED_TRY(grow(ws_.part, (size_t)2 * (m_ + 2) * G_ + kUpdGrid));
        SED_TRY(grow(ws_.S, (size_t)(m_ + 1) * (m_ + 1)));
        const size_t nd = 4 * (size_t)(m_ + 2) + 2 * (size_t)m_ + 1;
        SED_TRY(grow(ws_.state, nd + 4));
        SED_TRY(ws_.ensure_host(nd + 8));
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
        int zero[2] = {0, 0};
        SED_HIP(hipMemcpyAsync(step_, zero, sizeof(zero), hipMemcpyHostToDevice, st_));
        SED_HIP(hipStreamSynchronize(st_));  // `zero` is a stack temporary
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
        std::vector<double> theta, Sm;
        int jm = 0;
        bool done = false, breakdown = false;
        // first matvec eagerly: it fixes the launch shape (autotuner) before any capture
        const bool use_graph = graphs_enabled();
        while (!done) {
            int j = k;
            while (j < m_) {
                int chunk = std::min(m_ - j, j < nev + 2 ? nev + 2 - j : kCheckEvery);
                chunk = std::max(1, std::min(chunk, maxiter - out.iters));
                for (int c = 0; c < chunk; ++c) SED_TRY(step(use_graph && out.iters > 0));
                out.iters += chunk;
                out.matvecs += chunk;
                j += chunk;
                // read alpha/beta/flag of this cycle
                const size_t na = (size_t)m_;
                SED_HIP(hipMemcpyAsync(ws_.hcoef, alpha_, (2 * na + 1) * sizeof(double), hipMemcpyDeviceToHost, st_));
                SED_HIP(hipMemcpyAsync(ws_.hcoef + 2 * na + 2, step_, 2 * sizeof(int), hipMemcpyDeviceToHost, st_));
                SED_HIP(hipStreamSynchronize(st_));
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
                eig(jm, T_, theta, Sm);
                int conv = 0;
                for (int i = 0; i < std::min(nev, jm); ++i) {
                    const double res = beta * std::fabs(Sm[(size_t)(jm - 1) * jm + i]);
                    if (res <= tol) ++conv; else break;
                }
                if (breakdown || conv >= nev || out.iters >= maxiter) {
                    done = true;
                    break;
                }
            }
            if (done) break;
            // thick restart: kk Ritz vectors + the residual direction v_m
            const int kk = std::min(m_ - 2, nev + std::max(4, (m_ - nev) / 3));
            std::vector<double> Sk((size_t)jm * kk);
            for (int jj = 0; jj < jm; ++jj)
                for (int i = 0; i < kk; ++i) Sk[(size_t)jj * kk + i] = Sm[(size_t)jj * jm + i];
            SED_TRY(ritz(jm, Sk, kk));
            SED_HIP(hipMemcpyAsync(vec(kk), vec(jm), (size_t)n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
            SED_TRY(set_step_and_xhat(kk));
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
    static constexpr int kCheckEvery = 8;

    // one Lanczos step as plain launches
    int launch_step() {
        SED_TRY(launch_matvec_xhat(sec_, xh_, w_, VecTraits<T>::fp64, st_));
        hipLaunchKernelGGL(k_dots<T>, dim3(G_), dim3(kRedBlock), 0, st_, V_, (int64_t)n_, step_, w_, part_);
        hipLaunchKernelGGL(k_reduce, dim3(m_ + 2), dim3(kRedBlock), 0, st_, part_, G_, step_, h1_);
        hipLaunchKernelGGL(k_update<T>, dim3(kUpdGrid), dim3(256), 0, st_, V_, (int64_t)n_, step_, h1_, w_, 0,
                           normpart_);
        hipLaunchKernelGGL(k_dots<T>, dim3(G_), dim3(kRedBlock), 0, st_, V_, (int64_t)n_, step_, w_, part_);
        hipLaunchKernelGGL(k_reduce, dim3(m_ + 2), dim3(kRedBlock), 0, st_, part_, G_, step_, h2_);
        hipLaunchKernelGGL(k_update<T>, dim3(kUpdGrid), dim3(256), 0, st_, V_, (int64_t)n_, step_, h2_, w_, 1,
                           normpart_);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kRedBlock), 0, st_, step_, h1_, h2_, normpart_, kUpdGrid, alpha_,
                           beta_, scale_, flag_);
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n_)), dim3(256), 0, st_, V_, (int64_t)n_, step_, scale_,
                           sec_->d_stab.p, xh_);
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
        const size_t lds = (((size_t)m * kk * 8 + 15) & ~(size_t)15) + (size_t)m * kRitzRows * sizeof(T);
        const unsigned g = (unsigned)((n_ + kRitzRows - 1) / kRitzRows);
        hipLaunchKernelGGL(k_ritz<T>, dim3(g), dim3(kRitzRows), lds, st_, V_, (int64_t)n_, m, ws_.S.p, kk,
                           (int64_t)n_);
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
    int *step_ = nullptr, *flag_ = nullptr;
};

static int krylov_dim(int64_t n, int nev) {
    int m = std::max(2 * nev + 24, 40);
    return (int)std::min<int64_t>(m, std::min<int64_t>(n - 1, kMaxBasis));
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

// ============================================================================
// Packed Lanczos: many small sectors in one persistent launch
// ============================================================================
//
// One workgroup owns one sector and runs whole Lanczos steps inside the kernel:
// H|psi> over all of the sector's rows, two Gram-Schmidt passes (the second only when
// the first cancelled more than half the norm, the Daniel-Gragg-Kaufman-Stewart
// criterion), normalisation and the next prescaled vector, separated by workgroup
// barriers instead of kernel boundaries.  All sectors of a launch advance to their
// next check point together; the host then reads every sector's alpha/beta in one
// copy, checks residuals, restarts or retires sectors and admits queued ones before
// the next launch (continuous batching at check-point granularity).

}  // namespace sed

#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "sed_mv.hpp"

namespace sed {

MvConst mv_const(const sed_sector* sec, bool fp64);

constexpr int kPackBlock = 512;
constexpr int kPackCheck = 8;          // steps between host checks
constexpr int64_t kPackMaxDim = 1 << 17;

struct PackItem {
    MvConst c;
    const uint64_t* rep;
    const uint8_t* stab;
    const uint32_t* table;
    void* V;
    void* w;
    void* xhat;
    double* slot;      // alpha[m], beta[m], scale, (int step, int flag)
    int64_t n;
    int m;
    int target;
};

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

// h[i] = V_i^dag w for i < nv and *nrm = |w|^2 (every thread gets the results in LDS).
template <typename T, int NT>
__device__ void block_dots(const T* __restrict__ V, int64_t n, int nv, const T* __restrict__ w, double2* h,
                           double* nrm, double2 (*red)[NT / kWave]) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (int g0 = 0; g0 <= nv; g0 += kDotGroup) {
        double2 acc[kDotGroup];
#pragma unroll
        for (int q = 0; q < kDotGroup; ++q) acc[q] = make_double2(0.0, 0.0);
        for (int64_t r = threadIdx.x; r < n; r += NT) {
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
        if (threadIdx.x < kDotGroup) {
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

// w <- w - sum_{i<nv} V_i h_i; returns |w|^2 (all threads)
template <typename T, int NT>
__device__ double block_update(const T* __restrict__ V, int64_t n, int nv, const double2* h, T* w, double* red) {
    double nrm = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += NT) {
        double2 a = to_d2(w[r]);
        for (int i = 0; i < nv; ++i) {
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

template <typename M, typename R, bool FERMI>
__global__ void __launch_bounds__(kPackBlock) k_pack_lanczos(const PackItem* __restrict__ items) {
    using T = typename C2<R>::T;
    constexpr int NT = kPackBlock;
    __shared__ double2 sh_h[kMaxBasis + 2];
    __shared__ double2 sh_red[kDotGroup][NT / kWave];
    __shared__ double sh_redd[NT / kWave];
    __shared__ double sh_nrm;
    const PackItem& it = items[blockIdx.x];
    const MvLds<R> L = stage_mv<R>(it.c, NT);
    T* V = reinterpret_cast<T*>(it.V);
    T* w = reinterpret_cast<T*>(it.w);
    T* xh = reinterpret_cast<T*>(it.xhat);
    const int64_t n = it.n;
    const int m = it.m;
    double* alpha = it.slot;
    double* beta = it.slot + m;
    int* state = reinterpret_cast<int*>(it.slot + 2 * m + 1);
    int j = state[0];
    const int target = it.target;
    while (j < target) {
        // w = H' v_j from xhat = sqrt(stab) v_j
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            const M mm = (M)it.rep[r];
            R ax = 0, ay = 0, dx = 0, dy = 0;
            row_sums<M, R, FERMI, 1, 4, NT>(L, it.c, mm, 0, L.S + threadIdx.x, it.table, xh, ax, ay, dx, dy);
            const R inv = (R)1 / sqrt((R)it.stab[r]);
            const T xr = xh[r];
            T out;
            out.x = (ax + (dx * xr.x - dy * xr.y)) * inv;
            out.y = (ay + (dx * xr.y + dy * xr.x)) * inv;
            w[r] = out;
        }
        __syncthreads();
        const int nv = j + 1;
        block_dots<T, NT>(V, n, nv, w, sh_h, &sh_nrm, sh_red);
        const double nw = sh_nrm;
        double a = sh_h[j].x;
        double nw2 = block_update<T, NT>(V, n, nv, sh_h, w, sh_redd);
        if (!(nw2 > 0.5 * nw)) {
            // the first pass cancelled most of w: orthogonalise once more
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
            if (brk && state[1] == 0) state[1] = j + 1;
        }
        T* vn = V + (int64_t)(j + 1) * n;
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            const double2 d = to_d2(w[r]);
            const T o = from_d2<T>(make_double2(d.x * sc, d.y * sc));
            vn[r] = o;
            const double2 od = to_d2(o);
            const double s = sqrt((double)it.stab[r]);
            xh[r] = from_d2<T>(make_double2(od.x * s, od.y * s));
        }
        __syncthreads();
        ++j;
        if (brk) break;
    }
    if (threadIdx.x == 0) state[0] = j;
}

// step/flag/scale of a packed sector set by a kernel (no host temporaries in flight)
__global__ void k_set_state(int* state, int step, int reset_flag, double* scale, double sval) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        state[0] = step;
        if (reset_flag) state[1] = 0;
        *scale = sval;
    }
}

// ---------------------------------------------------------------------------
// device arena: one allocation, first-fit sub-allocation (256-byte aligned)
// ---------------------------------------------------------------------------
class Arena {
   public:
    int init(int64_t bytes) {
        bytes_ = (bytes + 255) & ~(int64_t)255;
        SED_TRY(buf_.alloc((size_t)bytes_));
        free_.clear();
        free_.push_back({0, bytes_});
        return SED_OK;
    }
    int64_t alloc(int64_t bytes) {
        bytes = (bytes + 255) & ~(int64_t)255;
        for (size_t i = 0; i < free_.size(); ++i) {
            if (free_[i].second >= bytes) {
                const int64_t off = free_[i].first;
                free_[i].first += bytes;
                free_[i].second -= bytes;
                if (free_[i].second == 0) free_.erase(free_.begin() + i);
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = (bytes + 255) & ~(int64_t)255;
        auto it = free_.begin();
        while (it != free_.end() && it->first < off) ++it;
        it = free_.insert(it, {off, bytes});
        // coalesce with neighbours
        if (it + 1 != free_.end() && it->first + it->second == (it + 1)->first) {
            it->second += (it + 1)->second;
            free_.erase(it + 1);
        }
        if (it != free_.begin() && (it - 1)->first + (it - 1)->second == it->first) {
            (it - 1)->second += it->second;
            free_.erase(it);
        }
    }
    uint8_t* base() { return buf_.p; }
    int64_t capacity() const { return bytes_; }
    void destroy() { buf_.release(); }

   private:
    DevBuf<uint8_t> buf_;
    int64_t bytes_ = 0;
    std::vector<std::pair<int64_t, int64_t>> free_;
};

static int pack_krylov_dim(int64_t n, int nev) { return krylov_dim(n, nev); }

bool pack_eligible(const sed_sector* sec, int nev) {
    return sec->dim > 500 && sec->dim <= kPackMaxDim && nev <= pack_krylov_dim(sec->dim, nev) - 2;
}

static int64_t pack_bytes(int64_t n, int m, bool fp32) {
    int64_t b = (int64_t)(m + 1) * n * 16 + 3 * n * 16;          // V64, w64, xh64, v0
    if (fp32) b += (int64_t)(m + 1) * n * 8 + 2 * n * 8;          // V32, w32, xh32
    b += (int64_t)(m + 1) * (m + 1) * 8 + 4096;                    // Ritz coefficients + alignment
    return b;
}

int64_t pack_footprint(const sed_sector* sec, int nev, int precision) {
    return pack_bytes(sec->dim, pack_krylov_dim(sec->dim, nev), precision != SED_PREC_FP64);
}

struct PackReq {
    PackJob* job = nullptr;
    int64_t n = 0;
    int m = 0;
    int phase = 1;           // 0 = FP32 Lanczos, 1 = FP64 Lanczos
    int slot = -1;
    int64_t off = -1, bytes = 0;
    double *V64 = nullptr, *w64 = nullptr, *xh64 = nullptr, *v0 = nullptr, *S = nullptr;
    float *V32 = nullptr, *w32 = nullptr, *xh32 = nullptr;
    std::vector<double> T;   // m x m projected matrix
    int k = 0;               // kept Ritz vectors after the last restart
    int j = 0;               // steps taken in the current cycle (device step)
    int iters[2] = {0, 0};
    int restarts = 0;
    std::chrono::steady_clock::time_point t0;
};

class PackEngine {
   public:
    PackEngine(sed_context* ctx, int max_active, void (*cb)(void*, PackJob*), void* user)
        : ctx_(ctx), max_active_(max_active), cb_(cb), user_(user) {}

    int start(int64_t arena_bytes) {
        SED_HIP(hipSetDevice(ctx_->device));
        SED_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        SED_TRY(arena_.init(arena_bytes));
        slot_doubles_ = 2 * kMaxBasis + 2;
        SED_TRY(slots_.alloc((size_t)max_active_ * slot_doubles_));
        SED_TRY(items_.alloc((size_t)max_active_ * sizeof(PackItem)));
        SED_HIP(hipHostMalloc((void**)&h_slots_, (size_t)max_active_ * slot_doubles_ * 8, hipHostMallocDefault));
        SED_HIP(hipHostMalloc((void**)&h_items_, (size_t)max_active_ * sizeof(PackItem), hipHostMallocDefault));
        free_slots_.clear();
        for (int i = max_active_ - 1; i >= 0; --i) free_slots_.push_back(i);
        th_ = std::thread([this] { loop(); });
        return SED_OK;
    }

    void submit(PackJob* job) {
        std::lock_guard<std::mutex> lk(mu_);
        queue_.push_back(job);
        cv_.notify_all();
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cv_.notify_all();
        }
        if (th_.joinable()) th_.join();
        (void)hipSetDevice(ctx_->device);
        if (st_) (void)hipStreamDestroy(st_);
        if (h_slots_) (void)hipHostFree(h_slots_);
        if (h_items_) (void)hipHostFree(h_items_);
        slots_.release();
        items_.release();
        arena_.destroy();
    }

   private:
    double* slot_ptr(int s) { return slots_.p + (size_t)s * slot_doubles_; }

    void finish(PackReq& q, int rc, const std::string& err) {
        PackJob* job = q.job;
        job->status = rc == SED_OK ? 0 : rc;
        job->error = err;
        job->stats.iters_fp32 = q.iters[0];
        job->stats.iters_fp64 = q.iters[1];
        job->stats.matvecs = q.iters[0] + q.iters[1];
        job->stats.restarts = q.restarts;
        job->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - q.t0).count();
        if (q.off >= 0) arena_.release(q.off, q.bytes);
        if (q.slot >= 0) free_slots_.push_back(q.slot);
        q.off = -1;
        q.slot = -1;
        cb_(user_, job);
    }

    // start vector of the current phase: random (FP64-only or FP32 phase) or v0 (refinement);
    // only launches (no synchronisation)
    template <typename T>
    int start_phase(PackReq& q, T* V, T* xh, bool from_v0) {
        double* sl = slot_ptr(q.slot);
        int* state = reinterpret_cast<int*>(sl + 2 * q.m + 1);
        double* scale = sl + 2 * q.m;
        double* normpart = q.S;   // scratch: kUpdGrid doubles fit in the Ritz area
        if (from_v0)
            hipLaunchKernelGGL((k_convert<double2, T>), dim3(grid_for(q.n)), dim3(256), 0, st_,
                               reinterpret_cast<const double2*>(q.v0), V, q.n);
        else
            hipLaunchKernelGGL(k_random<T>, dim3(grid_for(q.n)), dim3(256), 0, st_, V, q.n, q.job->seed);
        hipLaunchKernelGGL(k_set_state, dim3(1), dim3(64), 0, st_, state, 0, 1, scale, 1.0);
        const unsigned g = (unsigned)std::min<int64_t>(kUpdGrid, std::max<int64_t>(1, (q.n + 255) / 256));
        hipLaunchKernelGGL(k_normsq<T>, dim3(g), dim3(256), 0, st_, V, q.n, normpart);
        hipLaunchKernelGGL(k_start_finalize, dim3(1), dim3(kRedBlock), 0, st_, state, 0, normpart, (int)g, scale);
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(q.n)), dim3(256), 0, st_, V, q.n, state, scale,
                           q.job->sec->d_stab.p, xh);
        SED_HIP(hipGetLastError());
        q.T.assign((size_t)q.m * q.m, 0.0);
        q.k = 0;
        q.j = 0;
        return SED_OK;
    }

    int admit(PackJob* job, PackReq& q) {
        q.job = job;
        q.t0 = std::chrono::steady_clock::now();
        q.n = job->sec->dim;
        q.m = pack_krylov_dim(q.n, job->nev);
        const bool fp32 = job->precision != SED_PREC_FP64;
        q.bytes = pack_bytes(q.n, q.m, fp32);
        q.off = arena_.alloc(q.bytes);
        if (q.off < 0) return 1;   // no room yet
        q.slot = free_slots_.back();
        free_slots_.pop_back();
        uint8_t* p = arena_.base() + q.off;
        auto take = [&](int64_t bytes) {
            uint8_t* r = p;
            p += (bytes + 255) & ~(int64_t)255;
            return r;
        };
        q.V64 = reinterpret_cast<double*>(take((int64_t)(q.m + 1) * q.n * 16));
        q.w64 = reinterpret_cast<double*>(take(q.n * 16));
        q.xh64 = reinterpret_cast<double*>(take(q.n * 16));
        q.v0 = reinterpret_cast<double*>(take(q.n * 16));
        if (fp32) {
            q.V32 = reinterpret_cast<float*>(take((int64_t)(q.m + 1) * q.n * 8));
            q.w32 = reinterpret_cast<float*>(take(q.n * 8));
            q.xh32 = reinterpret_cast<float*>(take(q.n * 8));
        }
        q.S = reinterpret_cast<double*>(take((int64_t)std::max((q.m + 1) * (q.m + 1), kUpdGrid) * 8));
        q.phase = fp32 ? 0 : 1;
        int rc = q.phase == 0 ? start_phase<float2>(q, reinterpret_cast<float2*>(q.V32), reinterpret_cast<float2*>(q.xh32), false)
                              : start_phase<double2>(q, reinterpret_cast<double2*>(q.V64), reinterpret_cast<double2*>(q.xh64), false);
        return rc == SED_OK ? 0 : rc;
    }

    // Ritz combination in place; Sk must stay alive until the stream is synchronised
    template <typename T>
    int ritz(PackReq& q, T* V, int jm, const std::vector<double>& Sk, int kk) {
        SED_HIP(hipMemcpyAsync(q.S, Sk.data(), Sk.size() * 8, hipMemcpyHostToDevice, st_));
        const size_t lds = (((size_t)jm * kk * 8 + 15) & ~(size_t)15) + (size_t)jm * kRitzRows * sizeof(T);
        hipLaunchKernelGGL(k_ritz<T>, dim3((unsigned)((q.n + kRitzRows - 1) / kRitzRows)), dim3(kRitzRows), lds, st_, V,
                           q.n, jm, q.S, kk, q.n);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }

    // What a check point decided for one request (pure host computation).
    struct Decision {
        int action = 0;          // 0 continue, 1 done, 2 refine (FP32 -> FP64), 3 restart, -1 fail
        int rc = SED_OK;
        std::string err;
        int jm = 0, kk = 0;
        std::vector<double> theta, Sk;   // Ritz values; kept columns of S (jm x kk)
        double beta = 0.0, resmax = 0.0;
        int conv = 0;
    };

    void decide(PackReq& q, Decision& d) {
        PackJob* job = q.job;
        const int m = q.m;
        const double* hs = h_slots_ + (size_t)q.slot * slot_doubles_;
        const double* ha = hs;
        const double* hb = hs + m;
        int st2[2];
        std::memcpy(st2, hs + 2 * m + 1, sizeof(st2));
        const int jdev = st2[0], flag = st2[1];
        q.iters[q.phase] += jdev - q.j;
        q.j = jdev;
        int jm = jdev;
        bool breakdown = false;
        if (flag > 0 && flag <= jdev) {
            breakdown = true;
            jm = flag;
        }
        for (int i = q.k; i < jm; ++i) {
            q.T[(size_t)i * m + i] = ha[i];
            if (i + 1 < jm) {
                q.T[(size_t)(i + 1) * m + i] = hb[i];
                q.T[(size_t)i * m + i + 1] = hb[i];
            }
        }
        const double beta = breakdown ? 0.0 : hb[jm - 1];
        std::vector<double> a((size_t)jm * jm), Sm;
        for (int r = 0; r < jm; ++r)
            for (int c = 0; c < jm; ++c) a[(size_t)r * jm + c] = q.T[(size_t)r * m + c];
        jacobi_eigh(jm, a, d.theta, Sm);
        const int nev = std::min(job->nev, jm);
        const double tol = q.phase == 0 ? (job->precision == SED_PREC_FP32 ? std::max(job->tol, 1e-5) : 1e-4)
                                        : job->tol;
        for (int i = 0; i < nev; ++i) {
            const double res = beta * std::fabs(Sm[(size_t)(jm - 1) * jm + i]);
            if (res <= tol && d.conv == i) ++d.conv;
            d.resmax = std::max(d.resmax, res);
        }
        d.jm = jm;
        d.beta = beta;
        auto keep_cols = [&](int kk) {
            d.kk = kk;
            d.Sk.resize((size_t)jm * kk);
            for (int r = 0; r < jm; ++r)
                for (int i = 0; i < kk; ++i) d.Sk[(size_t)r * kk + i] = Sm[(size_t)r * jm + i];
        };
        if (d.conv >= job->nev || breakdown) {
            if (d.conv < job->nev) {
                d.action = -1;
                d.rc = SED_ENOCONV;
                d.err = "Lanczos (packed) did not converge the requested eigenpairs";
            } else if (q.phase == 0 && job->precision == SED_PREC_FP32_REFINE) {
                d.action = 2;
                keep_cols(nev);
            } else {
                d.action = 1;
            }
            return;
        }
        if (q.iters[q.phase] >= job->maxiter) {
            d.action = -1;
            d.rc = SED_ENOCONV;
            d.err = "Lanczos (packed) reached maxiter";
            return;
        }
        if (jm == m) {
            d.action = 3;
            keep_cols(std::min(m - 2, job->nev + std::max(4, (m - job->nev) / 3)));
            // the restart's projected matrix: Ritz values and their couplings to v_m
            std::fill(q.T.begin(), q.T.end(), 0.0);
            for (int i = 0; i < d.kk; ++i) {
                q.T[(size_t)i * m + i] = d.theta[i];
                const double sc = beta * Sm[(size_t)(jm - 1) * jm + i];
                q.T[(size_t)d.kk * m + i] = sc;
                q.T[(size_t)i * m + d.kk] = sc;
            }
        }
    }

    // GPU side of a decision (launches only); returns 1 when the request left the engine
    template <typename T>
    int apply(PackReq& q, T* V, T* xh, Decision& d) {
        PackJob* job = q.job;
        switch (d.action) {
            case -1:
                finish(q, d.rc, d.err);
                return 1;
            case 1:
                job->evals.assign(d.theta.begin(), d.theta.begin() + std::min<int>(job->nev, (int)d.theta.size()));
                job->stats.converged = d.conv;
                job->stats.residual_max = d.resmax;
                finish(q, SED_OK, "");
                return 1;
            case 2: {
                SED_TRY(ritz<T>(q, V, d.jm, d.Sk, d.kk));
                hipLaunchKernelGGL(k_sum_vectors<T>, dim3(grid_for(q.n)), dim3(256), 0, st_, V, q.n, d.kk,
                                   reinterpret_cast<double2*>(q.v0), q.n);
                SED_HIP(hipGetLastError());
                q.phase = 1;
                SED_TRY(start_phase<double2>(q, reinterpret_cast<double2*>(q.V64), reinterpret_cast<double2*>(q.xh64),
                                             true));
                return 0;
            }
            case 3: {
                const int m = q.m, kk = d.kk;
                SED_TRY(ritz<T>(q, V, d.jm, d.Sk, kk));
                SED_HIP(hipMemcpyAsync(V + (size_t)kk * q.n, V + (size_t)d.jm * q.n, (size_t)q.n * sizeof(T),
                                       hipMemcpyDeviceToDevice, st_));
                double* sl = slot_ptr(q.slot);
                int* state = reinterpret_cast<int*>(sl + 2 * m + 1);
                double* scale = sl + 2 * m;
                hipLaunchKernelGGL(k_set_state, dim3(1), dim3(64), 0, st_, state, kk, 0, scale, 1.0);
                hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(q.n)), dim3(256), 0, st_, V, q.n, state, scale,
                                   job->sec->d_stab.p, xh);
                SED_HIP(hipGetLastError());
                q.k = kk;
                q.j = kk;
                ++q.restarts;
                return 0;
            }
            default:
                return 0;
        }
    }

    template <typename M, typename R, bool F>
    int launch_group(const std::vector<int>& idx, std::vector<PackReq>& act, int offset) {
        if (idx.empty()) return SED_OK;
        size_t lds = 0;
        for (size_t t = 0; t < idx.size(); ++t) {
            PackReq& q = act[idx[t]];
            PackItem& it = h_items_[offset + t];
            it.c = mv_const(q.job->sec, sizeof(R) == 8);
            it.rep = q.job->sec->d_rep.p;
            it.stab = q.job->sec->d_stab.p;
            it.table = q.job->sec->d_table.p;
            if (sizeof(R) == 8) {
                it.V = q.V64;
                it.w = q.w64;
                it.xhat = q.xh64;
            } else {
                it.V = q.V32;
                it.w = q.w32;
                it.xhat = q.xh32;
            }
            it.slot = slot_ptr(q.slot);
            it.n = q.n;
            it.m = q.m;
            const int first = q.job->nev + 2;
            it.target = std::min(q.m, q.j < first ? first : q.j + kPackCheck);
            size_t off[9];
            lds = std::max(lds, mv_lds_layout<R>(it.c, off, kPackBlock));
        }
        PackItem* d_items = reinterpret_cast<PackItem*>(items_.p) + offset;
        SED_HIP(hipMemcpyAsync(d_items, h_items_ + offset, idx.size() * sizeof(PackItem), hipMemcpyHostToDevice, st_));
        hipLaunchKernelGGL((k_pack_lanczos<M, R, F>), dim3((unsigned)idx.size()), dim3(kPackBlock), lds, st_, d_items);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }

    int cycle(std::vector<PackReq>& act) {
        // group by kernel instantiation: mask width, statistics, precision phase
        std::vector<int> grp[8];
        for (int i = 0; i < (int)act.size(); ++i) {
            const sed_sector* s = act[i].job->sec;
            const int key = (s->cat->n_site <= 32 ? 0 : 4) + (s->cat->stats == SED_FERMIONIC ? 2 : 0) +
                            (act[i].phase == 1 ? 1 : 0);
            grp[key].push_back(i);
        }
        int off = 0;
        SED_TRY((launch_group<uint32_t, float, false>(grp[0], act, off))); off += (int)grp[0].size();
        SED_TRY((launch_group<uint32_t, double, false>(grp[1], act, off))); off += (int)grp[1].size();
        SED_TRY((launch_group<uint32_t, float, true>(grp[2], act, off))); off += (int)grp[2].size();
        SED_TRY((launch_group<uint32_t, double, true>(grp[3], act, off))); off += (int)grp[3].size();
        SED_TRY((launch_group<uint64_t, float, false>(grp[4], act, off))); off += (int)grp[4].size();
        SED_TRY((launch_group<uint64_t, double, false>(grp[5], act, off))); off += (int)grp[5].size();
        SED_TRY((launch_group<uint64_t, float, true>(grp[6], act, off))); off += (int)grp[6].size();
        SED_TRY((launch_group<uint64_t, double, true>(grp[7], act, off))); off += (int)grp[7].size();
        SED_HIP(hipMemcpyAsync(h_slots_, slots_.p, (size_t)max_active_ * slot_doubles_ * 8, hipMemcpyDeviceToHost, st_));
        SED_HIP(hipStreamSynchronize(st_));
        return SED_OK;
    }

    void loop() {
        (void)hipSetDevice(ctx_->device);
        std::vector<PackReq> act;
        while (true) {
            std::vector<std::pair<PackReq, std::pair<int, std::string>>> failed;
            bool stopping = false;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !queue_.empty() || !act.empty(); });
                stopping = stop_;
                // admit while slots and arena space last (GPU work is done unlocked)
                while (!stopping && !queue_.empty() && !free_slots_.empty()) {
                    PackJob* job = queue_.front();
                    queue_.pop_front();
                    lk.unlock();
                    PackReq q;
                    const int rc = admit(job, q);
                    lk.lock();
                    if (rc == 1) {
                        if (act.empty()) {   // larger than the whole arena: can never run
                            q.job = job;
                            failed.push_back({q, {SED_ENOMEM, "request larger than the packed engine's arena"}});
                            continue;
                        }
                        queue_.push_front(job);   // wait for a retirement
                        break;
                    }
                    if (rc < 0) {
                        failed.push_back({q, {rc, sed_last_error()}});
                        continue;
                    }
                    act.push_back(q);
                }
                if (stopping) {
                    for (PackJob* job : queue_) {
                        PackReq q;
                        q.job = job;
                        failed.push_back({q, {SED_EINVAL, "server stopped"}});
                    }
                    queue_.clear();
                }
            }
            for (auto& f : failed) finish(f.first, f.second.first, f.second.second);
            if (stopping) {
                for (PackReq& q : act) finish(q, SED_EINVAL, "server stopped");
                act.clear();
                break;
            }
            if (act.empty()) continue;
            const auto tc0 = std::chrono::steady_clock::now();
            const int rc = cycle(act);
            const auto tc1 = std::chrono::steady_clock::now();
            if (rc != SED_OK) {
                const std::string err = sed_last_error();
                for (PackReq& q : act) finish(q, rc, err);
                act.clear();
                continue;
            }
            // host decisions in parallel (pure computation), then the GPU actions in order
            std::vector<Decision> dec(act.size());
            {
                const int nthr = (int)std::min<size_t>(act.size(), (size_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
                std::vector<std::thread> pool;
                for (int t = 0; t < nthr; ++t)
                    pool.emplace_back([&, t] {
                        for (size_t i = t; i < act.size(); i += nthr) decide(act[i], dec[i]);
                    });
                for (auto& th : pool) th.join();
            }
            const auto tc2 = std::chrono::steady_clock::now();
            std::vector<PackReq> keep;
            std::vector<int> done_idx;
            for (size_t i = 0; i < act.size(); ++i) {
                PackReq& q = act[i];
                int r;
                if (q.phase == 0) r = apply<float2>(q, reinterpret_cast<float2*>(q.V32), reinterpret_cast<float2*>(q.xh32), dec[i]);
                else r = apply<double2>(q, reinterpret_cast<double2*>(q.V64), reinterpret_cast<double2*>(q.xh64), dec[i]);
                if (r < 0) finish(q, r, sed_last_error());
                else if (r == 0) keep.push_back(q);
            }
            // the Ritz coefficients in `dec` must outlive the copies queued above
            if (hipStreamSynchronize(st_) != hipSuccess) {
                for (PackReq& q : keep) finish(q, SED_EDEVICE, "packed engine stream failed");
                keep.clear();
            }
            act.swap(keep);
            if (trace_) {
                const auto tc3 = std::chrono::steady_clock::now();
                auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
                fprintf(stderr, "[pack] active=%zu gpu=%.2fms decide=%.2fms apply=%.2fms\n", dec.size(), ms(tc0, tc1),
                        ms(tc1, tc2), ms(tc2, tc3));
            }
        }
    }

    bool trace_ = std::getenv("SED_PACK_TRACE") != nullptr;
    sed_context* ctx_;
    int max_active_;
    void (*cb_)(void*, PackJob*);
    void* user_;
    hipStream_t st_ = nullptr;
    Arena arena_;
    DevBuf<double> slots_;
    DevBuf<uint8_t> items_;
    double* h_slots_ = nullptr;
    PackItem* h_items_ = nullptr;
    int slot_doubles_ = 0;
    std::vector<int> free_slots_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<PackJob*> queue_;
    bool stop_ = false;
    std::thread th_;
};

PackEngine* pack_create(sed_context* ctx, int64_t arena_bytes, int max_active, void (*on_done)(void*, PackJob*),
                        void* user) {
    auto* e = new PackEngine(ctx, max_active, on_done, user);
    if (e->start(arena_bytes) != SED_OK) {
        e->stop();
        delete e;
        return nullptr;
    }
    return e;
}
void pack_submit(PackEngine* e, PackJob* job) { e->submit(job); }
void pack_destroy(PackEngine* e) {
    if (!e) return;
    e->stop();
    delete e;
}

}  // namespace sed
