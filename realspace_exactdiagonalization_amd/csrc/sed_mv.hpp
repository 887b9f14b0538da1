// Shared pieces of the row-gather H|psi> (device constants in LDS and the per-row
// hop loop), used by the stand-alone matvec and by the packed Lanczos kernel.
#pragma once

#include "sed_device.hpp"

namespace sed {

constexpr int kMvBlock = 256;
// Each source site's term list in the padded copy is a multiple of kTermPad long; padding
// terms hop from the site to itself (always gated out: the site is occupied) with t' = 0.
constexpr int kTermPad = 4;

struct MvConst {
    const uint32_t* bin;
    const int32_t* tstart;
    const uint8_t* tto;
    const void* tcoef;
    const int32_t* dstart;
    const uint8_t* dj;
    const void* dv;
    const void* chi;
    // padded term lists: per entry b | (b * bs) << 8, conj(t'); per-site starts; and the
    // characters with the sign folded in: chi2[g << 1 | p] = (-1)^p chi(g) for index
    // g << 1 | p < 1 << shift, chi2[1 << shift] = 0 (dead hops)
    const int32_t* pstart;
    const uint32_t* pinfo;
    const void* pcoef;
    const void* chi2;
    int n, k, nG, bs, shift, n_terms, n_dens, n_pterms;
    uint32_t gmask;
    int64_t dim;
};

// Constants of the hop loops in LDS.  With the padded term lists (`padded` layouts),
// tstart / tcoef / chi hold the padded starts, coefficients and sign-folded characters,
// pinfo the packed targets, and tto is unused.
template <typename R>
struct MvLds {
    using T = typename C2<R>::T;
    uint32_t* bin;
    int32_t* tstart;
    int32_t* dstart;
    T* tcoef;
    T* dv;
    T* chi;
    uint8_t* tto;
    uint32_t* pinfo;
    uint8_t* dj;
    uint32_t* S;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Rank scratch entries per thread: k (per-site S of the site-by-site loops) or 2 (k + 1) (the row's
// prefix arrays of row_sums_flat).
__host__ __device__ inline int mv_scratch(const MvConst& c, bool flat) {
    return flat ? 2 * (c.k + 1) : (c.k > 0 ? c.k : 1);
}

__host__ __device__ inline int mv_chi2_count(const MvConst& c) { return (1 << c.shift) + 1; }

template <typename R>
__host__ __device__ inline size_t mv_lds_layout(const MvConst& c, size_t* off, int block, bool flat = false,
                                                bool padded = false) {
    using T = typename C2<R>::T;
    const size_t nt = padded ? (size_t)c.n_pterms : (size_t)c.n_terms;
    const size_t nchi = padded ? (size_t)mv_chi2_count(c) : (size_t)c.nG;
    size_t o = 0;
    off[0] = o; o = align16(o + (size_t)(c.n + 1) * c.bs * 4);       // bin
    off[1] = o; o = align16(o + (size_t)(c.n + 1) * 4);               // tstart
    off[2] = o; o = align16(o + (size_t)(c.n + 1) * 4);               // dstart
    off[3] = o; o = align16(o + nt * sizeof(T));                      // tcoef
    off[4] = o; o = align16(o + (size_t)c.n_dens * sizeof(T));        // dv
    off[5] = o; o = align16(o + nchi * sizeof(T));                    // chi
    off[6] = o; o = align16(o + nt * (padded ? 4 : 1));               // tto / pinfo
    off[7] = o; o = align16(o + (size_t)c.n_dens);                    // dj
    off[8] = o; o = align16(o + (size_t)mv_scratch(c, flat) * block * 4);   // S
    return o;
}

template <typename R>
__device__ __forceinline__ MvLds<R> stage_mv(const MvConst& c, int block, bool flat = false, bool padded = false) {
    using T = typename C2<R>::T;
    extern __shared__ __align__(16) uint8_t smem[];
    size_t off[9];
    mv_lds_layout<R>(c, off, block, flat, padded);
    MvLds<R> L;
    L.bin = reinterpret_cast<uint32_t*>(smem + off[0]);
    L.tstart = reinterpret_cast<int32_t*>(smem + off[1]);
    L.dstart = reinterpret_cast<int32_t*>(smem + off[2]);
    L.tcoef = reinterpret_cast<T*>(smem + off[3]);
    L.dv = reinterpret_cast<T*>(smem + off[4]);
    L.chi = reinterpret_cast<T*>(smem + off[5]);
    L.tto = smem + off[6];
    L.pinfo = reinterpret_cast<uint32_t*>(smem + off[6]);
    L.dj = smem + off[7];
    L.S = reinterpret_cast<uint32_t*>(smem + off[8]);
    const int nt = blockDim.x;
    for (int i = threadIdx.x; i < (c.n + 1) * c.bs; i += nt) L.bin[i] = c.bin[i];
    const int32_t* ts = padded ? c.pstart : c.tstart;
    for (int i = threadIdx.x; i <= c.n; i += nt) {
        L.tstart[i] = ts[i];
        L.dstart[i] = c.dstart[i];
    }
    const T* dv = reinterpret_cast<const T*>(c.dv);
    if (padded) {
        const T* tc = reinterpret_cast<const T*>(c.pcoef);
        const T* ch = reinterpret_cast<const T*>(c.chi2);
        for (int i = threadIdx.x; i < c.n_pterms; i += nt) {
            L.tcoef[i] = tc[i];
            L.pinfo[i] = c.pinfo[i];
        }
        for (int i = threadIdx.x; i < mv_chi2_count(c); i += nt) L.chi[i] = ch[i];
    } else {
        const T* tc = reinterpret_cast<const T*>(c.tcoef);
        const T* ch = reinterpret_cast<const T*>(c.chi);
        for (int i = threadIdx.x; i < c.n_terms; i += nt) {
            L.tcoef[i] = tc[i];
            L.tto[i] = c.tto[i];
        }
        for (int i = threadIdx.x; i < c.nG; i += nt) L.chi[i] = ch[i];
    }
    for (int i = threadIdx.x; i < c.n_dens; i += nt) {
        L.dv[i] = dv[i];
        L.dj[i] = c.dj[i];
    }
    __syncthreads();
    return L;
}

// Hop sums of one row m for the lane with index `sub` among the LPR lanes of the row:
// it takes occupied sites sub, sub+LPR, ... in ascending order.  S points at this
// thread's rank scratch (k entries, stride SST).  Returns the off-diagonal sum
// (ax, ay) = sum conj(t') s_JW sign chi(g) xhat[c] and the density sum (dx, dy).
//
// Per occupied site a, with mp = m without a (bits q_0 < ... < q_{k-2}), the rank of
// mp + b for an empty site b that lands at index ib among mp's bits is
//   sum_{i<ib} C(q_i, i+1) + sum_{i>=ib} C(q_i, i+2) + C(b, ib+1)
//   = base + S[ib] + C(b, ib+1),  base = sum_i C(q_i, i+2),
//     S[j] = sum_{i<j} (C(q_i, i+1) - C(q_i, i+2))    (uint32, modulo 2^32),
// one pass over mp's bits per site.  Offsets into the binomial table are 24-bit
// products (site < 64, stride < 70).
//
// The term lists are the padded ones (stage_mv with padded = true): every site's list is
// a multiple of kTermPad long, so a batch needs no bounds test, each entry carries its
// target's binomial row offset, and the hop's sign (table parity xor Jordan-Wigner parity)
// selects the signed character chi2[g << 1 | sign]; hops that are gated out or leave the
// sector select chi2[1 << shift] = 0.  Each hop then adds conj(t') chi2 xhat[c] with four
// fused multiply-adds.  With AMP (site phases not of gauge form) each hop also reads the
// target rank's hop amplitude amp[rank], issued beside its table read, and adds
// conj(t') chi2 amp xhat[c].
template <typename M, typename R, bool FERMI, int LPR, int BATCH, int SST, typename TE, bool AMP = false>
__device__ __forceinline__ void row_sums_padded(const MvLds<R>& L, const MvConst& c, const M m, const int sub,
                                                uint32_t* S, const TE* __restrict__ table,
                                                const typename C2<R>::T* __restrict__ xhat, R& ax, R& ay, R& dx,
                                                R& dy, const typename C2<R>::T* __restrict__ amp = nullptr) {
    using T = typename C2<R>::T;
    static_assert(kTermPad % BATCH == 0, "batch must divide the term padding");
    const uint32_t mask2 = (1u << c.shift) - 1u;
    const uint32_t dead = mask2 + 1u;
    M rest = m;
    for (int i = 0; i < sub && rest; ++i) rest &= rest - 1;
    int ia = sub;
    while (rest) {
        const int a = ctz_m<M>(rest);
        for (int i = 0; i < LPR && rest; ++i) rest &= rest - 1;
        const M mp = m ^ ((M)1 << a);
        uint32_t base = 0;
        {
            uint32_t cur = 0;
            S[0] = 0;
            M t = mp;
            int j = 0;
            while (t) {
                const int q = ctz_m<M>(t);
                t &= t - 1;
                const uint32_t* bq = L.bin + __umul24(q, c.bs) + j;
                const uint32_t c1 = bq[1], c2 = bq[2];
                base += c2;
                cur += c1 - c2;
                ++j;
                S[j * SST] = cur;
            }
        }
        for (int d = L.dstart[a]; d < L.dstart[a + 1]; ++d) {
            if ((m >> L.dj[d]) & 1) {
                dx += L.dv[d].x;
                dy += L.dv[d].y;
            }
        }
        const int t1 = L.tstart[a + 1];
        for (int t = L.tstart[a]; t < t1; t += BATCH) {
            uint32_t idx[BATCH];
            bool ok[BATCH];
            uint32_t pj[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const uint32_t info = L.pinfo[t + u];
                const int b = (int)(info & 63u);
                const int ib = popc_m<M>(mp & (((M)1 << b) - 1));
                ok[u] = !((m >> b) & 1);
                idx[u] = base + S[ib * SST] + L.bin[(info >> 8) + ib + 1];
                pj[u] = FERMI ? (uint32_t)((ib ^ ia) & 1) : 0u;
            }
            constexpr TE none = ~(TE)0;
            TE e[BATCH];
            T am[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const TE v = table[ok[u] ? idx[u] : 0u];
                if constexpr (AMP) am[u] = amp[ok[u] ? idx[u] : 0u];
                e[u] = ok[u] ? v : none;
            }
            T v[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) v[u] = xhat[e[u] != none ? (e[u] >> c.shift) : (TE)0];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const uint32_t ci = e[u] != none ? ((uint32_t)e[u] ^ pj[u]) & mask2 : dead;
                T cf = cmul(L.tcoef[t + u], L.chi[ci]);
                if constexpr (AMP) cf = cmul(cf, am[u]);
                ax += cf.x * v[u].x;
                ax -= cf.y * v[u].y;
                ay += cf.x * v[u].y;
                ay += cf.y * v[u].x;
            }
        }
        ia += LPR;
    }
}

// The same sums with every hop of the lane independent.  For m with bits
// q_0 < ... < q_{k-1}, R(m) = sum_j C(q_j, j+1) and the prefix sums (per row, in the
// scratch: PD at P[0..k], PU at P[k+1..2k+1], stride SST)
//   PD[j] = sum_{i<j} (C(q_i, i) - C(q_i, i+1)),   PU[j] = sum_{i<j} (C(q_i, i+2) - C(q_i, i+1)),
// the hop a = q_ia -> b (b empty; i_b = number of bits of m below b other than a) lands on
//   rank = R - C(a, i_a+1) + C(b, i_b+1) + (a < b ? PD[i_b+1] - PD[i_a+1] : PU[i_a] - PU[i_b])
// (the bits between a and b move down or up one index), in uint32 arithmetic modulo 2^32
// (the host enables this loop only when C(n-1, k+1) < 2^32).  The lane's candidates
// (sites sub, sub+LPR, ... and their term lists, in order) are walked in batches of
// BATCH that run across site boundaries: BATCH table reads, then BATCH gathers.
template <typename M, typename R, bool FERMI, int LPR, int BATCH, int SST, typename TE>
__device__ __forceinline__ void row_sums_flat(const MvLds<R>& L, const MvConst& c, const M m, const int sub,
                                              uint32_t* P, const TE* __restrict__ table,
                                              const typename C2<R>::T* __restrict__ xhat, R& ax, R& ay, R& dx,
                                              R& dy) {
    using T = typename C2<R>::T;
    const int K1 = c.k + 1;
    uint32_t Rm = 0;
    {
        uint32_t pd = 0, pu = 0;
        M t = m;
        int j = 0;
        while (t) {
            const int q = ctz_m<M>(t);
            t &= t - 1;
            const uint32_t* bq = L.bin + __umul24(q, c.bs) + j;
            const uint32_t c0 = bq[0], c1 = bq[1], c2 = bq[2];
            P[j * SST] = pd;
            P[(K1 + j) * SST] = pu;
            Rm += c1;
            pd += c0 - c1;
            pu += c2 - c1;
            ++j;
        }
        P[j * SST] = pd;
        P[(K1 + j) * SST] = pu;
    }
    // this lane's first site
    M rest = m;
    for (int i = 0; i < sub && rest; ++i) rest &= rest - 1;
    int ia = sub;
    int a = -1, t = 0, t1 = 0;
    uint32_t ca = 0;
    auto next_site = [&]() {
        while (rest) {
            a = ctz_m<M>(rest);
            for (int i = 0; i < LPR && rest; ++i) rest &= rest - 1;
            for (int d = L.dstart[a]; d < L.dstart[a + 1]; ++d) {
                if ((m >> L.dj[d]) & 1) {
                    dx += L.dv[d].x;
                    dy += L.dv[d].y;
                }
            }
            t = L.tstart[a];
            t1 = L.tstart[a + 1];
            ca = L.bin[__umul24(a, c.bs) + ia + 1];
            if (t < t1) return;
            ia += LPR;
        }
        a = -1;
    };
    next_site();
    while (a >= 0) {
        uint32_t idx[BATCH];
        bool ok[BATCH];
        int tt[BATCH], pj[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            if (a >= 0 && t >= t1) {
                ia += LPR;
                next_site();
            }
            const bool live = a >= 0;
            const int tc = live ? t : 0;
            const int b = L.tto[tc];
            const int below = popc_m<M>(m & (((M)1 << b) - 1));
            const int ib = below - ((live && a < b) ? 1 : 0);
            const uint32_t mid = (live && a < b) ? P[(ib + 1) * SST] - P[(ia + 1) * SST]
                                                 : P[(K1 + (live ? ia : 0)) * SST] - P[(K1 + ib) * SST];
            ok[u] = live && !((m >> b) & 1);
            idx[u] = Rm - ca + L.bin[__umul24(b, c.bs) + ib + 1] + mid;
            tt[u] = tc;
            pj[u] = FERMI ? ((ib - ia) & 1) : 0;
            if (live) ++t;
        }
        constexpr TE none = ~(TE)0;
        TE e[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const TE v = table[ok[u] ? idx[u] : 0u];
            e[u] = ok[u] ? v : none;
        }
        T v[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) v[u] = xhat[e[u] != none ? (e[u] >> c.shift) : (TE)0];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const bool live = e[u] != none;
            const int g = live ? (int)(((uint32_t)e[u] >> 1) & c.gmask) : 0;
            const T cf = cmul(L.tcoef[tt[u]], L.chi[g]);
            const T pr = cmul(cf, v[u]);
            const R sg = live ? (((((uint32_t)e[u] & 1) ^ (uint32_t)pj[u]) != 0) ? (R)-1 : (R)1) : (R)0;
            ax += sg * pr.x;
            ay += sg * pr.y;
        }
        if (a >= 0 && t >= t1) {
            ia += LPR;
            next_site();
        }
    }
}

}  // namespace sed
