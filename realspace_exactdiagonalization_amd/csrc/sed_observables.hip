// Sector observables from an eigenvector, and vector precision conversions.
//
// Site densities <n_i> of a sector state psi = sum_r c_r |r~> with orthonormal orbit
// states |r~>: each of the |G| images g(rep_r) carries weight |c_r|^2 / |G|
// (observables/static_structure_factor.jl:62-83), so
//   <n_i> = sum_r |c_r|^2 / |G| * #{g : site i occupied in g(rep_r)}.
// Bit i of g(rep) is set iff site perm_g^-1(i) is occupied in rep, which needs only
// the inverse permutations (LDS) and no image construction.
#include <algorithm>

#include "sed_device.hpp"
#include "sed_internal.hpp"

namespace sed {

constexpr int kObsBlock = 256;

__global__ void __launch_bounds__(kObsBlock) k_density(const uint64_t* __restrict__ rep,
                                                       const double2* __restrict__ c, int64_t dim, int n, int nG,
                                                       const uint8_t* __restrict__ pinv,
                                                       double* __restrict__ partial) {
    __shared__ uint8_t lp[64 * 64];
    __shared__ double red[kObsBlock / kWave];
    for (int i = threadIdx.x; i < n * nG; i += blockDim.x) lp[i] = pinv[i];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * kObsBlock + threadIdx.x;
    uint64_t m = 0;
    double w = 0.0;
    if (r < dim) {
        m = rep[r];
        const double2 v = c[r];
        w = (v.x * v.x + v.y * v.y) / nG;
    }
    for (int i = 0; i < n; ++i) {
        int cnt = 0;
        for (int g = 0; g < nG; ++g) cnt += (int)((m >> lp[g * n + i]) & 1);
        double val = w * cnt;
        for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = val;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int k = 0; k < kObsBlock / kWave; ++k) s += red[k];
            partial[(int64_t)blockIdx.x * n + i] = s;
        }
        __syncthreads();
    }
}

int launch_density(sed_sector* sec, const double2* vec, double* density) {
    const sed_catalog* cat = sec->cat;
    hipStream_t s = cat->ctx->stream;
    const int n = cat->n_site, nG = cat->n_group;
    std::vector<uint8_t> pinv((size_t)n * nG);
    for (int g = 0; g < nG; ++g)
        for (int i = 0; i < n; ++i) pinv[(size_t)g * n + (cat->perm[(size_t)g * n + i] - 1)] = (uint8_t)i;
    DevBuf<uint8_t> dp;
    SED_TRY(dp.alloc(pinv.size()));
    SED_HIP(hipMemcpyAsync(dp.p, pinv.data(), pinv.size(), hipMemcpyHostToDevice, s));
    const int64_t nb = std::max<int64_t>(1, (sec->dim + kObsBlock - 1) / kObsBlock);
    DevBuf<double> part;
    SED_TRY(part.alloc((size_t)nb * n));
    hipLaunchKernelGGL(k_density, dim3((unsigned)nb), dim3(kObsBlock), 0, s, sec->d_rep.p, vec, sec->dim, n, nG,
                       dp.p, part.p);
    SED_HIP(hipGetLastError());
    std::vector<double> h((size_t)nb * n);
    SED_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) density[i] = 0.0;
    for (int64_t b = 0; b < nb; ++b)
        for (int i = 0; i < n; ++i) density[i] += h[(size_t)b * n + i];
    return SED_OK;
}

__global__ void __launch_bounds__(256) k_d2f(const double2* __restrict__ a, float2* __restrict__ b, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        b[r] = make_float2((float)a[r].x, (float)a[r].y);
}
__global__ void __launch_bounds__(256) k_f2d(const float2* __restrict__ a, double2* __restrict__ b, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        b[r] = make_double2(a[r].x, a[r].y);
}

int launch_convert(const double* a, float* b, int64_t n, hipStream_t s) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(k_d2f, dim3(g), dim3(256), 0, s, reinterpret_cast<const double2*>(a),
                       reinterpret_cast<float2*>(b), n);
    SED_HIP(hipGetLastError());
    return SED_OK;
}
int launch_convert(const float* a, double* b, int64_t n, hipStream_t s) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(k_f2d, dim3(g), dim3(256), 0, s, reinterpret_cast<const float2*>(a),
                       reinterpret_cast<double2*>(b), n);
    SED_HIP(hipGetLastError());
    return SED_OK;
}

}  // namespace sed
