// Representative tables, orbit catalog and sector compaction on gfx950.
//
// Replaces build_symmetry_orbit_catalog (src/symmetry_resolved_ed.jl:371-487) and
// build_symmetry_sector_basis (:534-556).  The reference walks all C(N,k) masks
// serially with Gosper's hack and, for each orbit minimum, writes the tables of every
// orbit image (the last group element g with g(rep) = m wins).  Here every
// combinadic rank is independent:
//   1. k_cat_images : one thread per rank r -> mask s = unrank(r); the |G| images
//      h(s) give rep = min_h h(s).  Among the h with h(s) = rep the one whose inverse
//      index is largest is kept: it is g^-1 for the largest g with g(rep) = s, i.e.
//      exactly the element the reference's last writer leaves in rep_sym_table.
//      The fermionic reordering sign of that h is stored with it.
//   2. prefix scan over the flags "r is its own representative" numbers the orbits
//      in ascending mask order (= colex rank order = Gosper order).
//   3. k_cat_final  : every rank takes the orbit number of its representative.
//   4. k_cat_orbits : one thread per orbit -> stabilizer set and stabilizer signs.
// Sector compaction is a per-orbit compatibility flag, a prefix scan and one
// per-rank pass writing the packed 4-byte sector table consumed by H|psi>.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "sed_device.hpp"
#include "sed_internal.hpp"

namespace sed {

constexpr int kCatBlock = 256;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = (int64_t)kScanBlock * kScanItems;
constexpr int kScanTopThreads = 1024;

// ------------------------------------------------------------------------
// block-level reductions and scans (wave64)
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// inclusive scan within a wave
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & (kWave - 1);
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive scan across the block; returns prefix, writes block total to *total.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    uint32_t inc = wave_incl_scan(v);
    if (lane == kWave - 1) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < NW; ++i) {
            uint32_t t = lds[i];
            lds[i] = acc;
            acc += t;
        }
        lds[NW] = acc;
    }
    __syncthreads();
    uint32_t r = lds[w] + inc - v;
    *total = lds[NW];
    __syncthreads();
    return r;
}

template <typename Flag>
__global__ void __launch_bounds__(kScanBlock) k_scan_count(Flag f, int64_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[kScanBlock / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
#pragma unroll 4
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kScanBlock + threadIdx.x;
        if (idx < n) c += f(idx);
    }
    c = wave_sum_u32(c);
    const int lane = threadIdx.x & (kWave - 1);
    if (lane == 0) lds[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < kScanBlock / kWave; ++i) s += lds[i];
        sums[blockIdx.x] = s;
    }
}

// single-block exclusive scan of nb block sums; offsets[nb] = total
__global__ void __launch_bounds__(kScanTopThreads) k_scan_top(const uint32_t* __restrict__ sums, int64_t nb,
                                                              uint32_t* __restrict__ offsets) {
    __shared__ uint32_t lds[kScanTopThreads / kWave + 1];
    const int64_t per = (nb + kScanTopThreads - 1) / kScanTopThreads;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    const int64_t b1 = std::min<int64_t>(b0 + per, nb);
    uint32_t local = 0;
    for (int64_t b = b0; b < b1; ++b) local += sums[b];
    uint32_t total;
    uint32_t run = block_excl_scan<kScanTopThreads>(local, lds, &total);
    for (int64_t b = b0; b < b1; ++b) {
        uint32_t t = sums[b];
        offsets[b] = run;
        run += t;
    }
    if (threadIdx.x == 0) offsets[nb] = total;
}

template <typename Flag, typename Writer>
__global__ void __launch_bounds__(kScanBlock) k_scan_write(Flag f, Writer w, int64_t n,
                                                           const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t lds[kScanBlock / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t run = offsets[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kScanBlock + threadIdx.x;
        uint32_t fl = idx < n ? f(idx) : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan<kScanBlock>(fl, lds, &tot);
        if (idx < n) w(idx, fl, run + ex);
        run += tot;
    }
}

// Exclusive scan driver: counts, scans the block sums, returns the total on the
// host, then (after the caller allocated what the writer needs) writes.
struct ScanPlan {
    DevBuf<uint32_t> sums, offsets;
    int64_t nb = 0;
    uint32_t total = 0;
};

template <typename Flag>
static int scan_prepare(Flag f, int64_t n, hipStream_t s, ScanPlan& plan) {
    plan.nb = (n + kScanTile - 1) / kScanTile;
    if (plan.nb == 0) plan.nb = 1;
    SED_TRY(plan.sums.alloc(plan.nb));
    SED_TRY(plan.offsets.alloc(plan.nb + 1));
    hipLaunchKernelGGL(k_scan_count<Flag>, dim3((unsigned)plan.nb), dim3(kScanBlock), 0, s, f, n, plan.sums.p);
    SED_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, s, plan.sums.p, plan.nb, plan.offsets.p);
    SED_HIP(hipGetLastError());
    SED_HIP(hipMemcpyAsync(&plan.total, plan.offsets.p + plan.nb, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

template <typename Flag, typename Writer>
static int scan_write(Flag f, Writer w, int64_t n, hipStream_t s, ScanPlan& plan) {
    hipLaunchKernelGGL((k_scan_write<Flag, Writer>), dim3((unsigned)plan.nb), dim3(kScanBlock), 0, s, f, w, n,
                       plan.offsets.p);
    SED_HIP(hipGetLastError());
    return SED_OK;
}

// ------------------------------------------------------------------------
// LDS staging of the catalog constants
// ------------------------------------------------------------------------
struct CatConst {
    const uint32_t* bin;
    const uint8_t* perm;
    const uint8_t* inv;
    int n, k, nG, bs;
};

__device__ __forceinline__ void stage_cat(const CatConst& c, uint32_t*& bin, uint8_t*& perm, uint8_t*& inv) {
    extern __shared__ __align__(16) uint8_t smem[];
    bin = reinterpret_cast<uint32_t*>(smem);
    const int nb = (c.n + 1) * c.bs;
    perm = reinterpret_cast<uint8_t*>(bin + nb);
    inv = perm + c.nG * c.n;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bin[i] = c.bin[i];
    for (int i = threadIdx.x; i < c.nG * c.n; i += blockDim.x) perm[i] = c.perm[i];
    for (int i = threadIdx.x; i < c.nG; i += blockDim.x) inv[i] = c.inv[i];
    __syncthreads();
}

static size_t cat_lds_bytes(const sed_catalog* cat) {
    return (size_t)(cat->n_site + 1) * cat->bstride * 4 + (size_t)cat->n_group * cat->n_site + cat->n_group + 16;
}

// ------------------------------------------------------------------------
// 1. per-rank representative, group element and sign
// ------------------------------------------------------------------------
// Catalogs of 2^31 <= C(N,k) < 2^32 ranks (BIG) cannot keep a representative's rank in
// the int32 rank table: this pass then stores only whether rank r is its own
// representative (1) or not (0), and k_cat_final_big fills in the orbits after the scan.
template <typename M, bool FERMI, bool BIG>
__global__ void __launch_bounds__(kCatBlock) k_cat_images(CatConst c, int64_t n_total,
                                                          int32_t* __restrict__ rep_rank,
                                                          uint8_t* __restrict__ sym_par) {
    uint32_t* bin;
    uint8_t *perm, *inv;
    stage_cat(c, bin, perm, inv);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        const M s = colex_unrank<M>((uint32_t)r, c.n, c.k, bin, c.bs);
        M best = ~(M)0;
        int best_inv = -1, best_h = 0, best_par = 0;
        for (int h = 0; h < c.nG; ++h) {
            int par;
            const M img = apply_perm<M, FERMI>(s, perm + h * c.n, par);
            const int ih = inv[h];
            if (img < best || (img == best && ih > best_inv)) {
                best = img;
                best_inv = ih;
                best_h = h;
                best_par = par;
            }
        }
        if (BIG) rep_rank[r] = best == s ? 1 : 0;
        else rep_rank[r] = (int32_t)colex_rank<M>(best, bin, c.bs);
        sym_par[r] = (uint8_t)(best_h | (best_par << 7));
    }
}

struct RepFlag {
    const int32_t* rr;
    __device__ uint32_t operator()(int64_t i) const { return rr[i] == (int32_t)i ? 1u : 0u; }
};
struct RepFlagBig {
    const int32_t* rr;
    __device__ uint32_t operator()(int64_t i) const { return rr[i] == 1 ? 1u : 0u; }
};
struct RepWriter {
    int32_t* rr;
    int64_t* orbit_rank;
    __device__ void operator()(int64_t i, uint32_t fl, uint32_t ex) const {
        if (fl) {
            rr[i] = -(int32_t)ex - 1;
            orbit_rank[ex] = i;
        }
    }
};

// 3. every rank takes its representative's orbit number (stored as -(orbit+1))
__global__ void __launch_bounds__(256) k_cat_final(int64_t n_total, int32_t* __restrict__ rr) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = rr[r];
        if (v >= 0) rr[r] = rr[v];
    }
}

// 3 (BIG). every rank that is not a representative takes its representative's orbit
// number: the representative (minimal image) is found again and its rank, where the scan
// wrote -(orbit+1), read.  Representatives are not rewritten.
template <typename M, bool FERMI>
__global__ void __launch_bounds__(kCatBlock) k_cat_final_big(CatConst c, int64_t n_total, int32_t* __restrict__ rr) {
    uint32_t* bin;
    uint8_t *perm, *inv;
    stage_cat(c, bin, perm, inv);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (rr[r] < 0) continue;
        const M s = colex_unrank<M>((uint32_t)r, c.n, c.k, bin, c.bs);
        M best = ~(M)0;
        for (int h = 0; h < c.nG; ++h) {
            int par;
            const M img = apply_perm<M, FERMI>(s, perm + h * c.n, par);
            best = img < best ? img : best;
        }
        rr[r] = rr[colex_rank<M>(best, bin, c.bs)];
    }
}

// 4. per-orbit representative mask, stabilizer set and stabilizer signs
template <typename M, bool FERMI>
__global__ void __launch_bounds__(kCatBlock) k_cat_orbits(CatConst c, int64_t n_orbits,
                                                          const int64_t* __restrict__ orbit_rank,
                                                          uint64_t* __restrict__ orep,
                                                          uint64_t* __restrict__ ostab,
                                                          uint64_t* __restrict__ ostab_par) {
    uint32_t* bin;
    uint8_t *perm, *inv;
    stage_cat(c, bin, perm, inv);
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_orbits;
         o += (int64_t)gridDim.x * blockDim.x) {
        const M s = colex_unrank<M>((uint32_t)orbit_rank[o], c.n, c.k, bin, c.bs);
        uint64_t st = 0, sp = 0;
        for (int h = 0; h < c.nG; ++h) {
            int par;
            const M img = apply_perm<M, FERMI>(s, perm + h * c.n, par);
            if (img == s) {
                st |= 1ull << h;
                if (par) sp |= 1ull << h;
            }
        }
        orep[o] = (uint64_t)s;
        ostab[o] = st;
        ostab_par[o] = sp;
    }
}

// Reference rep_amp_table entry conj(alpha_g(rep)) for g = inverse of the stored
// element, with the reference's left-to-right phase product (no FMA contraction).
template <typename M, bool FERMI>
__global__ void __launch_bounds__(kCatBlock) k_cat_amp(CatConst c, const double2* __restrict__ phases,
                                                       int64_t n_total, const int32_t* __restrict__ rr,
                                                       const uint8_t* __restrict__ sym_par,
                                                       const uint64_t* __restrict__ orep,
                                                       double2* __restrict__ amp) {
    uint32_t* bin;
    uint8_t *perm, *inv;
    stage_cat(c, bin, perm, inv);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = -(int64_t)rr[r] - 1;
        M rep = (M)orep[o];
        const int g = inv[sym_par[r] & 0x7f];
        const double2* ph = phases + (size_t)g * c.n;
        const uint8_t* p = perm + g * c.n;
        double2 a = make_double2(1.0, 0.0);
        M img = 0;
        int par = 0;
        while (rep) {
            const int i = ctz_m<M>(rep);
            rep &= rep - 1;
            const int q = p[i];
            if (FERMI) par ^= popc_m<M>((img >> q) >> 1) & 1;
            img |= (M)1 << q;
            const double2 b = ph[i];
            a = make_double2(__dsub_rn(__dmul_rn(a.x, b.x), __dmul_rn(a.y, b.y)),
                             __dadd_rn(__dmul_rn(a.x, b.y), __dmul_rn(a.y, b.x)));
        }
        if (par) a = make_double2(-a.x, -a.y);
        amp[r] = make_double2(a.x, -a.y);
    }
}

template <typename M, bool F, bool BIG>
static int run_catalog(sed_catalog* cat) {
    hipStream_t s = cat->ctx->stream;
    CatConst c{cat->d_binom.p, cat->d_perm.p, cat->d_inv.p, cat->n_site, cat->n_filled, cat->n_group,
               cat->bstride};
    const size_t lds = cat_lds_bytes(cat);
    const int64_t n = cat->n_total;
    SED_TRY(cat->d_rep_rank.alloc(n));
    SED_TRY(cat->d_sym_par.alloc(n));
    int64_t nblk = std::min<int64_t>((n + kCatBlock - 1) / kCatBlock, 1 << 22);
    hipLaunchKernelGGL((k_cat_images<M, F, BIG>), dim3((unsigned)nblk), dim3(kCatBlock), lds, s, c, n,
                       cat->d_rep_rank.p, cat->d_sym_par.p);
    SED_HIP(hipGetLastError());
    ScanPlan plan;
    using Flag = typename std::conditional<BIG, RepFlagBig, RepFlag>::type;
    Flag flag{cat->d_rep_rank.p};
    SED_TRY(scan_prepare(flag, n, s, plan));
    cat->n_orbits = plan.total;
    if (cat->n_orbits > ((int64_t)1 << 31))   // rank entries hold -(orbit + 1) in int32
        return fail(SED_EUNSUPPORTED, std::to_string(cat->n_orbits) + " orbits exceed 2^31");
    DevBuf<int64_t> orbit_rank;
    SED_TRY(orbit_rank.alloc(cat->n_orbits));
    SED_TRY(scan_write(flag, RepWriter{cat->d_rep_rank.p, orbit_rank.p}, n, s, plan));
    if (BIG) {
        hipLaunchKernelGGL((k_cat_final_big<M, F>), dim3((unsigned)nblk), dim3(kCatBlock), lds, s, c, n,
                           cat->d_rep_rank.p);
    } else {
        nblk = std::min<int64_t>((n + 255) / 256, 1 << 22);
        hipLaunchKernelGGL(k_cat_final, dim3((unsigned)nblk), dim3(256), 0, s, n, cat->d_rep_rank.p);
    }
    SED_HIP(hipGetLastError());
    SED_TRY(cat->d_orbit_rep.alloc(cat->n_orbits));
    SED_TRY(cat->d_orbit_stab.alloc(cat->n_orbits));
    SED_TRY(cat->d_orbit_stab_par.alloc(cat->n_orbits));
    nblk = std::max<int64_t>(1, std::min<int64_t>((cat->n_orbits + kCatBlock - 1) / kCatBlock, 1 << 22));
    hipLaunchKernelGGL((k_cat_orbits<M, F>), dim3((unsigned)nblk), dim3(kCatBlock), lds, s, c, cat->n_orbits,
                       orbit_rank.p, cat->d_orbit_rep.p, cat->d_orbit_stab.p, cat->d_orbit_stab_par.p);
    SED_HIP(hipGetLastError());
    SED_HIP(hipStreamSynchronize(s));
    orbit_rank.release();
    return SED_OK;
}

int launch_catalog(sed_catalog* cat) {
    const bool f = cat->stats == SED_FERMIONIC;
    // C(32, k) < 2^31: only catalogs of more than 32 sites can be BIG
    if (cat->n_site <= 32) return f ? run_catalog<uint32_t, true, false>(cat) : run_catalog<uint32_t, false, false>(cat);
    if (cat->n_total >= ((int64_t)1 << 31))
        return f ? run_catalog<uint64_t, true, true>(cat) : run_catalog<uint64_t, false, true>(cat);
    return f ? run_catalog<uint64_t, true, false>(cat) : run_catalog<uint64_t, false, false>(cat);
}

template <typename M, bool F>
static int run_amp(const sed_catalog* cat, double* d_amp) {
    hipStream_t s = cat->ctx->stream;
    CatConst c{cat->d_binom.p, cat->d_perm.p, cat->d_inv.p, cat->n_site, cat->n_filled, cat->n_group,
               cat->bstride};
    DevBuf<double> ph;
    SED_TRY(ph.alloc((size_t)2 * cat->n_group * cat->n_site));
    SED_HIP(hipMemcpyAsync(ph.p, cat->phases.data(), ph.bytes(), hipMemcpyHostToDevice, s));
    const int64_t n = cat->n_total;
    int64_t nblk = std::min<int64_t>((n + kCatBlock - 1) / kCatBlock, 1 << 22);
    hipLaunchKernelGGL((k_cat_amp<M, F>), dim3((unsigned)nblk), dim3(kCatBlock), cat_lds_bytes(cat), s, c,
                       reinterpret_cast<const double2*>(ph.p), n, cat->d_rep_rank.p, cat->d_sym_par.p,
                       cat->d_orbit_rep.p, reinterpret_cast<double2*>(d_amp));
    SED_HIP(hipGetLastError());
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

// Per-rank hop amplitude of a group with general site phases: for rank r (mask s in the
// orbit of rep, stored element h with h(s) = rep, g = h^-1), prod_{i in rep} eta_g(i) in
// ascending site order (the reordering sign lives in the sector table's parity bit): the
// conjugate of the reference's rep_amp_table entry (src/symmetry_resolved_ed.jl:466)
// without its sign.
template <typename M>
__global__ void __launch_bounds__(kCatBlock) k_cat_hop_amp(CatConst c, const double2* __restrict__ phases,
                                                           int64_t n_total, const int32_t* __restrict__ rr,
                                                           const uint8_t* __restrict__ sym_par,
                                                           const uint64_t* __restrict__ orep,
                                                           double2* __restrict__ a64, float2* __restrict__ a32) {
    uint32_t* bin;
    uint8_t *perm, *inv;
    stage_cat(c, bin, perm, inv);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = -(int64_t)rr[r] - 1;
        M rep = (M)orep[o];
        const double2* ph = phases + (size_t)inv[sym_par[r] & 0x7f] * c.n;
        double2 a = make_double2(1.0, 0.0);
        while (rep) {
            const int i = ctz_m<M>(rep);
            rep &= rep - 1;
            const double2 b = ph[i];
            a = make_double2(__dsub_rn(__dmul_rn(a.x, b.x), __dmul_rn(a.y, b.y)),
                             __dadd_rn(__dmul_rn(a.x, b.y), __dmul_rn(a.y, b.x)));
        }
        a64[r] = a;
        a32[r] = make_float2((float)a.x, (float)a.y);
    }
}

int launch_hop_amp(const sed_catalog* cat, HopAmp* out) {
    hipStream_t s = cat->ctx->stream;
    CatConst c{cat->d_binom.p, cat->d_perm.p, cat->d_inv.p, cat->n_site, cat->n_filled, cat->n_group,
               cat->bstride};
    DevBuf<double> ph;
    SED_TRY(ph.alloc((size_t)2 * cat->n_group * cat->n_site));
    SED_HIP(hipMemcpyAsync(ph.p, cat->phases.data(), ph.bytes(), hipMemcpyHostToDevice, s));
    const int64_t n = cat->n_total;
    SED_TRY(out->a64.alloc((size_t)2 * n));
    SED_TRY(out->a32.alloc((size_t)2 * n));
    const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>((n + kCatBlock - 1) / kCatBlock, 1 << 22));
    auto* a64 = reinterpret_cast<double2*>(out->a64.p);
    auto* a32 = reinterpret_cast<float2*>(out->a32.p);
    const auto* phd = reinterpret_cast<const double2*>(ph.p);
    if (cat->n_site <= 32)
        hipLaunchKernelGGL(k_cat_hop_amp<uint32_t>, dim3((unsigned)nblk), dim3(kCatBlock), cat_lds_bytes(cat), s, c,
                           phd, n, cat->d_rep_rank.p, cat->d_sym_par.p, cat->d_orbit_rep.p, a64, a32);
    else
        hipLaunchKernelGGL(k_cat_hop_amp<uint64_t>, dim3((unsigned)nblk), dim3(kCatBlock), cat_lds_bytes(cat), s, c,
                           phd, n, cat->d_rep_rank.p, cat->d_sym_par.p, cat->d_orbit_rep.p, a64, a32);
    SED_HIP(hipGetLastError());
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

int launch_catalog_amp(const sed_catalog* cat, double* d_amp) {
    const bool f = cat->stats == SED_FERMIONIC;
    if (cat->n_site <= 32) return f ? run_amp<uint32_t, true>(cat, d_amp) : run_amp<uint32_t, false>(cat, d_amp);
    return f ? run_amp<uint64_t, true>(cat, d_amp) : run_amp<uint64_t, false>(cat, d_amp);
}

// ------------------------------------------------------------------------
// sector compaction
// ------------------------------------------------------------------------
// _is_orbit_compatible (:534-542): every stabilizer element h of the orbit must
// satisfy |chi(h) - alpha_h(rep)| <= atol.  With gauge-form site phases the
// stabilizer phase alpha_h(rep) is the reordering sign alone.
__global__ void __launch_bounds__(256) k_sec_flags(int64_t n_orbits, int nG, const double2* __restrict__ chi,
                                                   double atol, const uint64_t* __restrict__ ostab,
                                                   const uint64_t* __restrict__ ostab_par,
                                                   uint8_t* __restrict__ flags) {
    __shared__ double2 lchi[64];
    for (int i = threadIdx.x; i < nG; i += blockDim.x) lchi[i] = chi[i];
    __syncthreads();
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_orbits;
         o += (int64_t)gridDim.x * blockDim.x) {
        uint64_t st = ostab[o];
        const uint64_t sp = ostab_par[o];
        uint8_t ok = 1;
        while (st) {
            const int h = __builtin_ctzll(st);
            st &= st - 1;
            const double target = ((sp >> h) & 1) ? -1.0 : 1.0;
            const double dx = lchi[h].x - target, dy = lchi[h].y;
            if (!(sqrt(dx * dx + dy * dy) <= atol)) ok = 0;
        }
        flags[o] = ok;
    }
}

// The same test with general site phases: alpha_h(rep) = sign * prod_{i in rep} eta_h(i)
// (ascending site order, as apply_operation_to_mask multiplies them).
__global__ void __launch_bounds__(256) k_sec_flags_general(int64_t n_orbits, int n, int nG,
                                                           const double2* __restrict__ chi, double atol,
                                                           const double2* __restrict__ phases,
                                                           const uint64_t* __restrict__ orep,
                                                           const uint64_t* __restrict__ ostab,
                                                           const uint64_t* __restrict__ ostab_par,
                                                           uint8_t* __restrict__ flags) {
    __shared__ double2 lchi[64];
    for (int i = threadIdx.x; i < nG; i += blockDim.x) lchi[i] = chi[i];
    __syncthreads();
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_orbits;
         o += (int64_t)gridDim.x * blockDim.x) {
        uint64_t st = ostab[o];
        const uint64_t sp = ostab_par[o];
        const uint64_t rep = orep[o];
        uint8_t ok = 1;
        while (st) {
            const int h = __builtin_ctzll(st);
            st &= st - 1;
            double2 a = make_double2(1.0, 0.0);
            for (uint64_t t = rep; t; t &= t - 1) a = cmul(a, phases[(size_t)h * n + __builtin_ctzll(t)]);
            if ((sp >> h) & 1) a = make_double2(-a.x, -a.y);
            const double dx = lchi[h].x - a.x, dy = lchi[h].y - a.y;
            if (!(sqrt(dx * dx + dy * dy) <= atol)) ok = 0;
        }
        flags[o] = ok;
    }
}

struct U8Flag {
    const uint8_t* f;
    __device__ uint32_t operator()(int64_t i) const { return f[i]; }
};
struct SectorWriter {
    int32_t* row_of;
    const uint64_t* orep;
    const uint64_t* ostab;
    uint64_t* srep;
    uint8_t* sstab;
    __device__ void operator()(int64_t o, uint32_t fl, uint32_t ex) const {
        row_of[o] = fl ? (int32_t)ex : -1;
        if (fl) {
            srep[ex] = orep[o];
            sstab[ex] = (uint8_t)__builtin_popcountll(ostab[o]);
        }
    }
};

template <typename TE>
__global__ void __launch_bounds__(256) k_sec_table(int64_t n_total, int shift, const int32_t* __restrict__ rr,
                                                   const uint8_t* __restrict__ sym_par,
                                                   const int32_t* __restrict__ row_of,
                                                   TE* __restrict__ table) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_total;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = -(int64_t)rr[r] - 1;
        const int32_t row = row_of[o];
        const uint32_t sp = sym_par[r];
        table[r] = row < 0 ? ~(TE)0 : (((TE)row << shift) | (TE)(((sp & 0x7fu) << 1) | (sp >> 7)));
    }
}

int launch_sector(sed_sector* sec, const double* chi_dev) {
    sed_catalog* cat = sec->cat;
    hipStream_t s = cat->ctx->stream;
    const int64_t no = cat->n_orbits;
    DevBuf<uint8_t> flags;
    SED_TRY(flags.alloc(no));
    int64_t nblk = std::max<int64_t>(1, std::min<int64_t>((no + 255) / 256, 1 << 22));
    DevBuf<double> ph;
    if (cat->general) {
        SED_TRY(ph.alloc((size_t)2 * cat->n_group * cat->n_site));
        SED_HIP(hipMemcpyAsync(ph.p, cat->phases.data(), ph.bytes(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_sec_flags_general, dim3((unsigned)nblk), dim3(256), 0, s, no, cat->n_site,
                           cat->n_group, reinterpret_cast<const double2*>(chi_dev), sec->atol,
                           reinterpret_cast<const double2*>(ph.p), cat->d_orbit_rep.p, cat->d_orbit_stab.p,
                           cat->d_orbit_stab_par.p, flags.p);
        sec->amp = cat->amp;
    } else {
        hipLaunchKernelGGL(k_sec_flags, dim3((unsigned)nblk), dim3(256), 0, s, no, cat->n_group,
                           reinterpret_cast<const double2*>(chi_dev), sec->atol, cat->d_orbit_stab.p,
                           cat->d_orbit_stab_par.p, flags.p);
    }
    SED_HIP(hipGetLastError());
    ScanPlan plan;
    U8Flag fl{flags.p};
    SED_TRY(scan_prepare(fl, no, s, plan));
    sec->dim = plan.total;
    int gbits = 0;
    while ((1 << gbits) < cat->n_group) ++gbits;
    sec->shift = gbits + 1;
    // rows are int32 (2^31 complex doubles are 32 GiB per Krylov vector); a row that does
    // not fit the 4-byte entry beside the group bits takes the 8-byte table
    if (sec->dim >= ((int64_t)1 << 31))
        return fail(SED_EUNSUPPORTED, "sector dimension " + std::to_string(sec->dim) + " exceeds 2^31 - 1");
    sec->wide = sec->dim > 0 && (uint64_t)sec->dim > ((1ull << (32 - sec->shift)) - 1);
    if (sec->wide && cat->n_site <= 32)   // C(32, k) / |G| always fits the 4-byte entry
        return fail(SED_EUNSUPPORTED, "8-byte sector table on a catalog of at most 32 sites");
    if (sec->wide && cat->general)
        return fail(SED_EUNSUPPORTED, "site phases not of gauge form on a sector of " + std::to_string(sec->dim) +
                                          " rows (8-byte table)");
    DevBuf<int32_t> row_of;
    SED_TRY(row_of.alloc(no));
    SED_TRY(sec->d_rep.alloc(sec->dim));
    SED_TRY(sec->d_stab.alloc(sec->dim));
    SED_TRY(scan_write(fl, SectorWriter{row_of.p, cat->d_orbit_rep.p, cat->d_orbit_stab.p, sec->d_rep.p,
                                        sec->d_stab.p},
                       no, s, plan));
    nblk = std::min<int64_t>((cat->n_total + 255) / 256, 1 << 22);
    if (sec->wide) {
        SED_TRY(sec->d_table64.alloc(cat->n_total));
        hipLaunchKernelGGL(k_sec_table<uint64_t>, dim3((unsigned)nblk), dim3(256), 0, s, cat->n_total, sec->shift,
                           cat->d_rep_rank.p, cat->d_sym_par.p, row_of.p, sec->d_table64.p);
    } else {
        SED_TRY(sec->d_table.alloc(cat->n_total));
        hipLaunchKernelGGL(k_sec_table<uint32_t>, dim3((unsigned)nblk), dim3(256), 0, s, cat->n_total, sec->shift,
                           cat->d_rep_rank.p, cat->d_sym_par.p, row_of.p, sec->d_table.p);
    }
    SED_HIP(hipGetLastError());
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

}  // namespace sed
