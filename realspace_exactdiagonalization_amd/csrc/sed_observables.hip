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

// Pair occupations of the representatives: M[a][b] = sum_r |c_r|^2 n_a(rep_r) n_b(rep_r).
// Each block takes a contiguous range of rows, staged 256 at a time in LDS; thread t owns
// the entries t + 256 q of the n x n matrix and sums them in row order (no atomics), so
// the per-block partials and their fixed-order sum are bitwise reproducible.
constexpr int kPairOwn = 16;   // 16 x 256 entries = a 64 x 64 matrix
__global__ void __launch_bounds__(kObsBlock) k_pair_density(const uint64_t* __restrict__ rep,
                                                            const double2* __restrict__ c, int64_t dim, int n,
                                                            double* __restrict__ partial) {
    __shared__ uint64_t lm[kObsBlock];
    __shared__ double lw[kObsBlock];
    const int nn = n * n;
    const int64_t chunk = (dim + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = std::min<int64_t>(dim, r0 + chunk);
    int ea[kPairOwn], eb[kPairOwn];
    double acc[kPairOwn];
#pragma unroll
    for (int q = 0; q < kPairOwn; ++q) {
        const int e = threadIdx.x + q * kObsBlock;
        ea[q] = e < nn ? e / n : 0;
        eb[q] = e < nn ? e % n : 0;
        acc[q] = 0.0;
    }
    for (int64_t t0 = r0; t0 < r1; t0 += kObsBlock) {
        const int64_t r = t0 + threadIdx.x;
        uint64_t m = 0;
        double w = 0.0;
        if (r < r1) {
            m = rep[r];
            const double2 v = c[r];
            w = v.x * v.x + v.y * v.y;
        }
        __syncthreads();
        lm[threadIdx.x] = m;
        lw[threadIdx.x] = w;
        __syncthreads();
        const int cnt = (int)std::min<int64_t>(kObsBlock, r1 - t0);
        for (int q2 = 0; q2 < cnt; ++q2) {
            const uint64_t ms = lm[q2];
            const double ws = lw[q2];
#pragma unroll
            for (int q = 0; q < kPairOwn; ++q) acc[q] += ((ms >> ea[q]) & (ms >> eb[q]) & 1) ? ws : 0.0;
        }
    }
#pragma unroll
    for (int q = 0; q < kPairOwn; ++q) {
        const int e = threadIdx.x + q * kObsBlock;
        if (e < nn) partial[(int64_t)blockIdx.x * nn + e] = acc[q];
    }
}

// out[e] = sum_b partial[b * nn + e], blocks in order
__global__ void __launch_bounds__(kObsBlock) k_sum_partials(const double* __restrict__ partial, int nb, int nn,
                                                            double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn) return;
    double sm = 0.0;
    for (int b = 0; b < nb; ++b) sm += partial[(int64_t)b * nn + e];
    out[e] = sm;
}

int launch_pair_density(sed_sector* sec, const double2* vec, std::vector<double>& M) {
    const sed_catalog* cat = sec->cat;
    hipStream_t s = cat->ctx->stream;
    const int n = cat->n_site, nn = n * n;
    M.assign((size_t)nn, 0.0);
    if (sec->dim == 0) return SED_OK;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (sec->dim + kObsBlock - 1) / kObsBlock));
    DevBuf<double> part, out;
    SED_TRY(part.alloc((size_t)nb * nn));
    SED_TRY(out.alloc((size_t)nn));
    hipLaunchKernelGGL(k_pair_density, dim3((unsigned)nb), dim3(kObsBlock), 0, s, sec->d_rep.p, vec, sec->dim, n,
                       part.p);
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((nn + kObsBlock - 1) / kObsBlock)), dim3(kObsBlock), 0, s,
                       part.p, nb, nn, out.p);
    SED_HIP(hipGetLastError());
    SED_HIP(hipMemcpyAsync(M.data(), out.p, (size_t)nn * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    part.release();
    out.release();
    return SED_OK;
}

// partial[b] = sum over block b's rows of conj(x[r]) y[r] (fixed grid and order)
__global__ void __launch_bounds__(kObsBlock) k_cdot(const double2* __restrict__ x, const double2* __restrict__ y,
                                                    int64_t n, double2* __restrict__ partial) {
    __shared__ double2 red[kObsBlock / kWave];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = std::min<int64_t>(n, r0 + chunk);
    double2 a = make_double2(0.0, 0.0);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kObsBlock) {
        const double2 u = x[r], v = y[r];
        a.x += u.x * v.x + u.y * v.y;
        a.y += u.x * v.y - u.y * v.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        a.x += __shfl_down(a.x, off, kWave);
        a.y += __shfl_down(a.y, off, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 sm = make_double2(0.0, 0.0);
        for (int k = 0; k < kObsBlock / kWave; ++k) {
            sm.x += red[k].x;
            sm.y += red[k].y;
        }
        partial[blockIdx.x] = sm;
    }
}

// <x|y> for complex double vectors on the device (part: at least kDotBlocks entries)
int launch_cdot(const double2* x, const double2* y, int64_t n, double2* part, std::vector<double2>& host,
                hipStream_t s, cd* out) {
    constexpr int kDotBlocks = 256;
    host.resize(kDotBlocks);
    hipLaunchKernelGGL(k_cdot, dim3(kDotBlocks), dim3(kObsBlock), 0, s, x, y, n, part);
    SED_HIP(hipGetLastError());
    SED_HIP(hipMemcpyAsync(host.data(), part, kDotBlocks * sizeof(double2), hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    cd sm = 0.0;
    for (const double2& v : host) sm += cd(v.x, v.y);
    *out = sm;
    return SED_OK;
}

// Sector state -> Fock amplitudes (observables/entanglement_spectrum.jl:46-79): every
// representative r with |c_r| > atol adds, for each group element g,
//   c_r / sqrt(|G| stab_r) * conj(chi(g)) * alpha_g(rep_r)
// to the amplitude of the image g(rep_r), alpha = product of the site phases of the
// occupied sites times the reordering sign.  One thread owns one representative and
// its whole orbit (orbits are disjoint), so it adds its images in order into the
// rank-indexed output without races.
template <typename M, bool FERMI>
__global__ void __launch_bounds__(kObsBlock) k_fock_expand(const uint64_t* __restrict__ rep,
                                                           const uint8_t* __restrict__ stab,
                                                           const double2* __restrict__ c, int64_t dim, double atol,
                                                           int n, int nG, const uint8_t* __restrict__ perm,
                                                           const double2* __restrict__ phases,
                                                           const double2* __restrict__ chi,
                                                           const uint32_t* __restrict__ bin, int bs,
                                                           double2* __restrict__ out) {
    __shared__ uint8_t lp[64 * 64];
    __shared__ uint32_t lb[65 * 66];
    for (int i = threadIdx.x; i < n * nG; i += blockDim.x) lp[i] = perm[i];
    for (int i = threadIdx.x; i < (n + 1) * bs; i += blockDim.x) lb[i] = bin[i];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dim) return;
    const double2 cr = c[r];
    if (!(sqrt(cr.x * cr.x + cr.y * cr.y) > atol)) return;
    const double pre = 1.0 / sqrt((double)nG * (double)stab[r]);
    const M m = (M)rep[r];
    for (int g = 0; g < nG; ++g) {
        int par = 0;
        const M img = apply_perm<M, FERMI>(m, lp + g * n, par);
        double2 ph = make_double2(1.0, 0.0);
        for (M t = m; t; t &= t - 1) ph = cmul(ph, phases[g * n + ctz_m<M>(t)]);
        const double2 xc = make_double2(chi[g].x, -chi[g].y);
        double2 a = cmul(cmul(make_double2(cr.x * pre, cr.y * pre), xc), ph);
        if (FERMI && par) a = make_double2(-a.x, -a.y);
        const uint32_t k = colex_rank<M>(img, lb, bs);
        out[k].x += a.x;
        out[k].y += a.y;
    }
}

int launch_fock_expand(sed_sector* sec, const double2* c, double atol, double2* out, hipStream_t s) {
    const sed_catalog* cat = sec->cat;
    const int n = cat->n_site, nG = cat->n_group;
    std::vector<double2> ph((size_t)nG * n);   // the site phases the sector was built with
    for (size_t i = 0; i < ph.size(); ++i) ph[i] = make_double2(sec->phases[i].real(), sec->phases[i].imag());
    DevBuf<double2> dph;
    SED_TRY(dph.alloc(ph.size()));
    SED_HIP(hipMemcpyAsync(dph.p, ph.data(), ph.size() * 16, hipMemcpyHostToDevice, s));
    SED_HIP(hipMemsetAsync(out, 0, (size_t)cat->n_total * 16, s));
    const unsigned nb = (unsigned)std::max<int64_t>(1, (sec->dim + kObsBlock - 1) / kObsBlock);
    const auto* chi = reinterpret_cast<const double2*>(sec->d_chi64.p);
    const bool F = cat->stats == SED_FERMIONIC;
    if (n <= 32) {
        if (F) hipLaunchKernelGGL((k_fock_expand<uint32_t, true>), dim3(nb), dim3(kObsBlock), 0, s, sec->d_rep.p, sec->d_stab.p, c, sec->dim, atol, n, nG, cat->d_perm.p, dph.p, chi, cat->d_binom.p, cat->bstride, out);
        else hipLaunchKernelGGL((k_fock_expand<uint32_t, false>), dim3(nb), dim3(kObsBlock), 0, s, sec->d_rep.p, sec->d_stab.p, c, sec->dim, atol, n, nG, cat->d_perm.p, dph.p, chi, cat->d_binom.p, cat->bstride, out);
    } else {
        if (F) hipLaunchKernelGGL((k_fock_expand<uint64_t, true>), dim3(nb), dim3(kObsBlock), 0, s, sec->d_rep.p, sec->d_stab.p, c, sec->dim, atol, n, nG, cat->d_perm.p, dph.p, chi, cat->d_binom.p, cat->bstride, out);
        else hipLaunchKernelGGL((k_fock_expand<uint64_t, false>), dim3(nb), dim3(kObsBlock), 0, s, sec->d_rep.p, sec->d_stab.p, c, sec->dim, atol, n, nG, cat->d_perm.p, dph.p, chi, cat->d_binom.p, cat->bstride, out);
    }
    SED_HIP(hipGetLastError());
    SED_HIP(hipStreamSynchronize(s));
    dph.release();
    return SED_OK;
}

// Projected position operator <chi_to| U_d |chi_from> (observables/charge_pump.jl:47-91):
// column col of the from-sector has at most one entry, in the to-sector row of the same
// representative, (1/|G|) sum_g chi_to(g) conj(chi_from(g)) z(g(rep)) with
// z(mask) = prod_{i in mask} phi_i; row = -1 when the orbit is not in the to-sector.
template <typename M>
__global__ void __launch_bounds__(kObsBlock) k_position(const uint64_t* __restrict__ rep, int64_t dim, int n, int nG,
                                                        const uint8_t* __restrict__ perm,
                                                        const double2* __restrict__ phi,
                                                        const double2* __restrict__ chi_to,
                                                        const double2* __restrict__ chi_from,
                                                        const uint32_t* __restrict__ table_to, int shift_to,
                                                        const uint32_t* __restrict__ bin, int bs,
                                                        int64_t* __restrict__ rows, double2* __restrict__ vals) {
    __shared__ uint8_t lp[64 * 64];
    __shared__ uint32_t lb[65 * 66];
    for (int i = threadIdx.x; i < n * nG; i += blockDim.x) lp[i] = perm[i];
    for (int i = threadIdx.x; i < (n + 1) * bs; i += blockDim.x) lb[i] = bin[i];
    __syncthreads();
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= dim) return;
    const M m = (M)rep[col];
    const uint32_t e = table_to[colex_rank<M>(m, lb, bs)];
    if (e == kNoRow) {
        rows[col] = -1;
        vals[col] = make_double2(0.0, 0.0);
        return;
    }
    double2 el = make_double2(0.0, 0.0);
    for (int g = 0; g < nG; ++g) {
        int par = 0;
        const M img = apply_perm<M, false>(m, lp + g * n, par);
        double2 z = make_double2(1.0, 0.0);
        for (M t = img; t; t &= t - 1) z = cmul(z, phi[ctz_m<M>(t)]);
        const double2 w = cmul(chi_to[g], make_double2(chi_from[g].x, -chi_from[g].y));
        const double2 v = cmul(w, z);
        el.x += v.x;
        el.y += v.y;
    }
    rows[col] = (int64_t)(e >> shift_to);
    vals[col] = make_double2(el.x / nG, el.y / nG);
}

int launch_position(sed_sector* to, sed_sector* from, const double2* phi, int64_t* rows, double2* vals,
                    hipStream_t s) {
    const sed_catalog* cat = from->cat;
    const int n = cat->n_site, nG = cat->n_group;
    const unsigned nb = (unsigned)std::max<int64_t>(1, (from->dim + kObsBlock - 1) / kObsBlock);
    const auto* cto = reinterpret_cast<const double2*>(to->d_chi64.p);
    const auto* cfr = reinterpret_cast<const double2*>(from->d_chi64.p);
    if (n <= 32)
        hipLaunchKernelGGL((k_position<uint32_t>), dim3(nb), dim3(kObsBlock), 0, s, from->d_rep.p, from->dim, n, nG,
                           cat->d_perm.p, phi, cto, cfr, to->d_table.p, to->shift, cat->d_binom.p, cat->bstride, rows,
                           vals);
    else
        hipLaunchKernelGGL((k_position<uint64_t>), dim3(nb), dim3(kObsBlock), 0, s, from->d_rep.p, from->dim, n, nG,
                           cat->d_perm.p, phi, cto, cfr, to->d_table.p, to->shift, cat->d_binom.p, cat->bstride, rows,
                           vals);
    SED_HIP(hipGetLastError());
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
