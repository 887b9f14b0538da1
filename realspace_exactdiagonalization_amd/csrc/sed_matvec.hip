// Row-gather H|psi> for one symmetry sector on gfx950.
//
// Replaces MatrixFreeHamiltonian / _apply_table! (src/symmetry_resolved_ed.jl:731-791,
// :880-884) and the on-the-fly apply_hamiltonian! (:813-878).  The reference scatters
// column contributions H[row, col] x[col] into per-thread buffers; here one thread owns
// one output row r and gathers, using Hermiticity H[r, c] = conj(H[c, r]):
//
//   H[c, r] = t_k * s_JW(m_r) * alpha * conj(chi(g)) * sqrt(stab_c / stab_r)
//
// for every term k = (a -> b) with site a occupied and b empty in the representative
// m_r (the reference's gate (m & hop) == from, :709), where (c, g, alpha) is the
// canonical image of m_r ^ hop (project_to_sector, :627-632).  No atomics, and every
// row is summed in a fixed order, so results are bitwise reproducible.
//
// Per hop the thread needs the combinadic rank of the new mask.  With m' = m_r minus
// bit a (bits q_0 < ... < q_{k-2}), inserting b at position i_b = popcount(m' below b)
// gives  rank = S[i_b] + C(b, i_b + 1)  where
//   S[i] = sum_{j<i} C(q_j, j+1) + sum_{j>=i} C(q_j, j+2),
// built once per occupied site a (k values, kept in LDS).  The Jordan-Wigner sign of
// hopping_phase_for_stats (:646-654) is (-1)^|i_b - i_a| with i_a the index of a in m_r.
//
// Table entry (4 bytes per combinadic rank): row << shift | g << 1 | parity, or
// 0xffffffff when the orbit is not in the sector.  The vector gathered is
// xhat = sqrt(stab) * x, so the stabilizer norm of the target row costs no load.
//
// Twisted (flux) groups: with site phases eta_g(i) = f(i)/f(perm_g(i)), the reference's
// matrix equals D^-1 H' D with D = diag(prod_{i in rep} f(i)) and H' built from the
// sign-only tables and terms t'_k = t_k f(b)/f(a).  This kernel applies H'; the host
// applies D where vectors cross the boundary.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "sed_device.hpp"
#include "sed_internal.hpp"
#include "sed_mv.hpp"

namespace sed {

// LPR lanes share one row (rows per block = kMvBlock / LPR); the row sum is closed by a
// fixed xor butterfly over the row's lanes.  Hops are processed BATCH at a time: BATCH
// independent table reads are issued before the BATCH dependent gathers, so each lane
// keeps several memory requests in flight.
template <typename M, typename R, bool FERMI, int LPR, int BATCH, int BLOCK, bool FLAT, typename TE, bool AMP>
__global__ void __launch_bounds__(BLOCK) k_matvec(MvConst c, const uint64_t* __restrict__ rep,
                                                  const uint8_t* __restrict__ stab,
                                                  const TE* __restrict__ table,
                                                  const typename C2<R>::T* __restrict__ xhat,
                                                  typename C2<R>::T* __restrict__ y, int64_t nblocks,
                                                  int64_t chunk, const typename C2<R>::T* __restrict__ amp) {
    using T = typename C2<R>::T;
    static_assert(!(FLAT && AMP), "hop amplitudes run on the site-by-site loop only");
    constexpr int RPB = BLOCK / LPR;
    const MvLds<R> L = stage_mv<R>(c, BLOCK, FLAT, !FLAT);
    const int64_t lb = xcd_block(nblocks, chunk);
    if (lb < 0) return;
    const int sub = threadIdx.x % LPR;
    const int64_t r = lb * RPB + threadIdx.x / LPR;
    const bool live = r < c.dim;
    const M m = live ? (M)rep[r] : (M)0;
    R ax = 0, ay = 0, dx = 0, dy = 0;
    if constexpr (FLAT)
        row_sums_flat<M, R, FERMI, LPR, BATCH, BLOCK, TE>(L, c, m, sub, L.S + threadIdx.x, table, xhat, ax, ay, dx,
                                                          dy);
    else
        row_sums_padded<M, R, FERMI, LPR, BATCH, BLOCK, TE, AMP>(L, c, m, sub, L.S + threadIdx.x, table, xhat, ax,
                                                                 ay, dx, dy, amp);
    if (LPR > 1) {
#pragma unroll
        for (int off = LPR / 2; off >= 1; off >>= 1) {
            ax += __shfl_xor(ax, off, kWave);
            ay += __shfl_xor(ay, off, kWave);
            dx += __shfl_xor(dx, off, kWave);
            dy += __shfl_xor(dy, off, kWave);
        }
    }
    if (!live || sub != 0) return;
    // diag(m_r) x[r] = diag(m_r) xhat[r] / sqrt(stab_r): one scale for both parts
    const R inv = (R)1 / sqrt((R)stab[r]);
    const T xr = xhat[r];
    T out;
    out.x = (ax + (dx * xr.x - dy * xr.y)) * inv;
    out.y = (ay + (dx * xr.y + dy * xr.x)) * inv;
    y[r] = out;
}

// Launch shapes the autotuner chooses from: lanes per row (rows per wavefront =
// 64 / lanes), hops per batch and workgroup size.  A shape whose LDS (rank scratch of
// k entries per thread) does not fit a workgroup is skipped.
// Shapes with `flat` walk each lane's hops in batches across site boundaries
// (row_sums_flat), the others site by site (row_sums_padded).
struct MvShape {
    int lanes, batch, block;
    bool flat;
};
static const MvShape kShapes[] = {{1, 1, 256, false}, {1, 4, 256, false}, {2, 4, 256, false},
                                  {4, 4, 256, false}, {8, 4, 256, false}, {4, 2, 256, false},
                                  {2, 4, 128, false}, {1, 4, 512, false}, {2, 4, 512, false},
                                  {1, 4, 256, true},  {1, 8, 256, true},  {2, 4, 256, true},
                                  {2, 8, 256, true}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
// the shape of sectors too small to tune and of launches before any tuning
constexpr int kDefaultShape = 1;
// XCD run lengths the autotuner tries for the chosen shape (0 = one range per XCD)
static const int kChunks[] = {1, 16, 64, 256, 1024, 0};
constexpr int kNumChunks = sizeof(kChunks) / sizeof(kChunks[0]);
static_assert(2 * kNumShapes + kNumChunks <= 64, "autotuner timings must fit sed_sector::tune_ms");

// Dense H' (row-major, dim x dim, complex double) for the small-sector path.
template <typename M, bool FERMI>
__global__ void __launch_bounds__(kMvBlock) k_dense(MvConst c, const uint64_t* __restrict__ rep,
                                                    const uint8_t* __restrict__ stab,
                                                    const uint32_t* __restrict__ table,
                                                    const double2* __restrict__ amp, double2* __restrict__ H) {
    const MvLds<double> L = stage_mv<double>(c, kMvBlock);
    const int64_t r = (int64_t)blockIdx.x * kMvBlock + threadIdx.x;
    if (r >= c.dim) return;
    const M m = (M)rep[r];
    const double inv = 1.0 / sqrt((double)stab[r]);
    double2* row = H + r * c.dim;
    M rest = m;
    int ia = 0;
    double dx = 0, dy = 0;
    while (rest) {
        const int a = ctz_m<M>(rest);
        rest &= rest - 1;
        const M mp = m ^ ((M)1 << a);
        for (int d = L.dstart[a]; d < L.dstart[a + 1]; ++d) {
            if ((m >> L.dj[d]) & 1) {
                dx += L.dv[d].x;
                dy += L.dv[d].y;
            }
        }
        for (int t = L.tstart[a]; t < L.tstart[a + 1]; ++t) {
            const int b = L.tto[t];
            if ((m >> b) & 1) continue;
            const M nm = mp | ((M)1 << b);
            const uint32_t rank = colex_rank<M>(nm, L.bin, c.bs);
            const uint32_t e = table[rank];
            if (e == kNoRow) continue;
            const uint32_t col = e >> c.shift;
            const int g = (e >> 1) & c.gmask;
            int p = e & 1;
            const int ib = popc_m<M>(mp & (((M)1 << b) - 1));
            if (FERMI) p ^= (ib - ia) & 1;
            double2 cf = cmul(L.tcoef[t], L.chi[g]);
            if (amp) cf = cmul(cf, amp[rank]);
            const double sc = sqrt((double)stab[col]) * inv * (p ? -1.0 : 1.0);
            row[col].x += cf.x * sc;
            row[col].y += cf.y * sc;
        }
        ++ia;
    }
    row[r].x += dx;
    row[r].y += dy;
}

// Exact counts of gated hops (one table read each) and in-sector hops (one gather).
template <typename M, typename TE>
__global__ void __launch_bounds__(kMvBlock) k_hop_counts(MvConst c, const uint64_t* __restrict__ rep,
                                                         const TE* __restrict__ table,
                                                         unsigned long long* __restrict__ partial) {
    const MvLds<double> L = stage_mv<double>(c, kMvBlock);
    __shared__ unsigned long long red[2][kMvBlock / kWave];
    const int64_t r = (int64_t)blockIdx.x * kMvBlock + threadIdx.x;
    unsigned long long ng = 0, ns = 0;
    if (r < c.dim) {
        const M m = (M)rep[r];
        M rest = m;
        while (rest) {
            const int a = ctz_m<M>(rest);
            rest &= rest - 1;
            const M mp = m ^ ((M)1 << a);
            for (int t = L.tstart[a]; t < L.tstart[a + 1]; ++t) {
                const int b = L.tto[t];
                if ((m >> b) & 1) continue;
                ++ng;
                const TE e = table[colex_rank<M>(mp | ((M)1 << b), L.bin, c.bs)];
                if (e != ~(TE)0) ++ns;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ng += __shfl_down(ng, off, kWave);
        ns += __shfl_down(ns, off, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][threadIdx.x / kWave] = ng;
        red[1][threadIdx.x / kWave] = ns;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int i = 0; i < kMvBlock / kWave; ++i) {
            a += red[0][i];
            b += red[1][i];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prescale(const T* __restrict__ x, const uint8_t* __restrict__ stab,
                                                  T* __restrict__ xh, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const T v = x[r];
        const auto s = sqrt((decltype(v.x))stab[r]);
        T o;
        o.x = v.x * s;
        o.y = v.y * s;
        xh[r] = o;
    }
}

__global__ void __launch_bounds__(256) k_gauge(double2* __restrict__ v, const double2* __restrict__ F,
                                               int64_t n, int forward) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 f = F[r];
        if (!forward) f.y = -f.y;
        v[r] = cmul(f, v[r]);
    }
}

// the same on FP32 vectors: out = F v (forward) or conj(F) v, F in FP64
__global__ void __launch_bounds__(256) k_gauge32(const float2* v, float2* out,
                                                 const double2* __restrict__ F, int64_t n, int forward) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 f = F[r];
        if (!forward) f.y = -f.y;
        const double2 a = cmul(f, make_double2(v[r].x, v[r].y));
        out[r] = make_float2((float)a.x, (float)a.y);
    }
}

int launch_gauge32(sed_sector* sec, const float* v, float* out, bool forward, hipStream_t s) {
    if (sec->dim == 0) return SED_OK;
    const int64_t g = std::min<int64_t>((sec->dim + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_gauge32, dim3((unsigned)g), dim3(256), 0, s, reinterpret_cast<const float2*>(v),
                       reinterpret_cast<float2*>(out), reinterpret_cast<const double2*>(sec->d_gauge.p), sec->dim,
                       forward ? 1 : 0);
    SED_HIP(hipGetLastError());
    return SED_OK;
}

// F(rep_r) = prod_{i in rep_r} f(i)
__global__ void __launch_bounds__(256) k_gauge_rows(const uint64_t* __restrict__ rep, const double2* __restrict__ f,
                                                    int n_site, int64_t dim, double2* __restrict__ F) {
    __shared__ double2 lf[64];
    for (int i = threadIdx.x; i < n_site; i += blockDim.x) lf[i] = f[i];
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t m = rep[r];
        double2 a = make_double2(1.0, 0.0);
        while (m) {
            const int i = __builtin_ctzll(m);
            m &= m - 1;
            a = cmul(a, lf[i]);
        }
        F[r] = a;
    }
}

// ------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------
template <typename R>
static MvConst make_const(const sed_sector* sec) {
    const sed_catalog* cat = sec->cat;
    MvConst c;
    c.bin = cat->d_binom.p;
    c.tstart = sec->d_term_start.p;
    c.tto = sec->d_term_to.p;
    c.dstart = sec->d_dens_start.p;
    c.dj = sec->d_dens_j.p;
    if (sizeof(R) == 8) {
        c.tcoef = sec->d_term_coef64.p;
        c.dv = sec->d_dens_v64.p;
        c.chi = sec->d_chi64.p;
    } else {
        c.tcoef = sec->d_term_coef32.p;
        c.dv = sec->d_dens_v32.p;
        c.chi = sec->d_chi32.p;
    }
    c.n = cat->n_site;
    c.k = cat->n_filled;
    c.nG = cat->n_group;
    c.bs = cat->bstride;
    c.shift = sec->shift;
    c.n_terms = sec->n_terms;
    c.n_dens = sec->n_dens;
    c.n_pterms = sec->n_pterms;
    c.pstart = sec->d_pterm_start.p;
    c.pinfo = sec->d_pterm_info.p;
    if (sizeof(R) == 8) {
        c.pcoef = sec->d_pterm_coef64.p;
        c.chi2 = sec->d_chi2_64.p;
    } else {
        c.pcoef = sec->d_pterm_coef32.p;
        c.chi2 = sec->d_chi2_32.p;
    }
    c.gmask = (1u << (sec->shift - 1)) - 1u;
    c.dim = sec->dim;
    return c;
}

MvConst mv_const(const sed_sector* sec, bool fp64) { return fp64 ? make_const<double>(sec) : make_const<float>(sec); }

template <typename R>
size_t matvec_lds_bytes(const sed_sector* sec) {
    MvConst c = make_const<R>(sec);
    size_t off[9];
    return std::max(mv_lds_layout<R>(c, off, kMvBlock), mv_lds_layout<R>(c, off, kMvBlock, false, true));
}
template size_t matvec_lds_bytes<float>(const sed_sector*);
template size_t matvec_lds_bytes<double>(const sed_sector*);

// The flat hop loop's prefix sums need C(n-1, k+1) below 2^32 (the binomial table
// saturates there).
static bool flat_ranks_exact(const sed_sector* sec) {
    const sed_catalog* cat = sec->cat;
    if (cat->n_site < 1) return false;
    return cat->binom[(size_t)(cat->n_site - 1) * cat->bstride + cat->n_filled + 1] != 0xffffffffu;
}

template <typename M, typename R, bool F, typename TE, int LPR, int B, int BLOCK, bool FLAT = false, bool AMP = false>
static int launch_shape(const MvConst& c, sed_sector* sec, const typename C2<R>::T* xh, typename C2<R>::T* y,
                        int chunk, hipStream_t s) {
    using T = typename C2<R>::T;
    size_t off[9];
    const size_t lds = mv_lds_layout<R>(c, off, BLOCK, FLAT, !FLAT);
    if ((int)lds > sec->cat->ctx->max_lds) return fail(SED_EUNSUPPORTED, "launch shape needs too much LDS");
    if (FLAT && !flat_ranks_exact(sec)) return fail(SED_EUNSUPPORTED, "flat hop loop: C(n-1, k+1) >= 2^32");
    const int64_t rpb = BLOCK / LPR;
    const int64_t nb = (sec->dim + rpb - 1) / rpb;
    const int64_t ch = chunk > 0 ? chunk : std::max<int64_t>(1, (nb + 7) / 8);
    const int64_t grid = ((nb + 8 * ch - 1) / (8 * ch)) * (8 * ch);
    const TE* table;
    if constexpr (sizeof(TE) == 8) table = sec->d_table64.p;
    else table = sec->d_table.p;
    const T* amp = nullptr;
    if constexpr (AMP) {
        if (!sec->amp) return fail(SED_ESTATE, "hop amplitudes missing");
        if constexpr (sizeof(R) == 8) amp = reinterpret_cast<const T*>(sec->amp->a64.p);
        else amp = reinterpret_cast<const T*>(sec->amp->a32.p);
    }
    hipLaunchKernelGGL((k_matvec<M, R, F, LPR, B, BLOCK, FLAT, TE, AMP>), dim3((unsigned)grid), dim3(BLOCK), lds, s,
                       c, sec->d_rep.p, sec->d_stab.p, table, xh, y, nb, ch, amp);
    SED_HIP(hipGetLastError());
    return SED_OK;
}

// Sectors with per-rank hop amplitudes (site phases not of gauge form, 4-byte tables) run
// shape 1 only.
constexpr int kAmpShape = 1;
template <typename M, typename R, bool F, typename TE>
static int run_shape(sed_sector* sec, int shape, int chunk, const typename C2<R>::T* xh, typename C2<R>::T* y,
                     hipStream_t s) {
    MvConst c = make_const<R>(sec);
    if (sec->amp) {
        if constexpr (sizeof(TE) == 4) {
            if (shape == kAmpShape) return launch_shape<M, R, F, TE, 1, 4, 256, false, true>(c, sec, xh, y, chunk, s);
        }
        return fail(SED_EUNSUPPORTED, "launch shape " + std::to_string(shape) +
                                          " with site phases not of gauge form (shape 1 only)");
    }
    switch (shape) {
        case 0: return launch_shape<M, R, F, TE, 1, 1, 256>(c, sec, xh, y, chunk, s);
        case 1: return launch_shape<M, R, F, TE, 1, 4, 256>(c, sec, xh, y, chunk, s);
        case 2: return launch_shape<M, R, F, TE, 2, 4, 256>(c, sec, xh, y, chunk, s);
        case 3: return launch_shape<M, R, F, TE, 4, 4, 256>(c, sec, xh, y, chunk, s);
        case 4: return launch_shape<M, R, F, TE, 8, 4, 256>(c, sec, xh, y, chunk, s);
        case 5: return launch_shape<M, R, F, TE, 4, 2, 256>(c, sec, xh, y, chunk, s);
        case 6: return launch_shape<M, R, F, TE, 2, 4, 128>(c, sec, xh, y, chunk, s);
        case 7: return launch_shape<M, R, F, TE, 1, 4, 512>(c, sec, xh, y, chunk, s);
        case 8: return launch_shape<M, R, F, TE, 2, 4, 512>(c, sec, xh, y, chunk, s);
        case 9: return launch_shape<M, R, F, TE, 1, 4, 256, true>(c, sec, xh, y, chunk, s);
        case 10: return launch_shape<M, R, F, TE, 1, 8, 256, true>(c, sec, xh, y, chunk, s);
        case 11: return launch_shape<M, R, F, TE, 2, 4, 256, true>(c, sec, xh, y, chunk, s);
        default: return launch_shape<M, R, F, TE, 2, 8, 256, true>(c, sec, xh, y, chunk, s);
    }
}

// A shape's workgroup LDS fits the device (and the flat loop's ranks are exact)
template <typename R>
static bool shape_fits(const sed_sector* sec, int shape) {
    MvConst c = make_const<R>(sec);
    size_t off[9];
    if (kShapes[shape].flat && !flat_ranks_exact(sec)) return false;
    if (sec->amp && shape != kAmpShape) return false;
    return (int)mv_lds_layout<R>(c, off, kShapes[shape].block, kShapes[shape].flat, !kShapes[shape].flat) <=
           sec->cat->ctx->max_lds;
}

template <typename R>
static int run_any_shape(sed_sector* sec, int shape, int chunk, const typename C2<R>::T* xh, typename C2<R>::T* y,
                         hipStream_t s) {
    const bool F = sec->cat->stats == SED_FERMIONIC;
    if (sec->cat->n_site <= 32)
        return F ? run_shape<uint32_t, R, true, uint32_t>(sec, shape, chunk, xh, y, s)
                 : run_shape<uint32_t, R, false, uint32_t>(sec, shape, chunk, xh, y, s);
    if (sec->wide)
        return F ? run_shape<uint64_t, R, true, uint64_t>(sec, shape, chunk, xh, y, s)
                 : run_shape<uint64_t, R, false, uint64_t>(sec, shape, chunk, xh, y, s);
    return F ? run_shape<uint64_t, R, true, uint32_t>(sec, shape, chunk, xh, y, s)
             : run_shape<uint64_t, R, false, uint32_t>(sec, shape, chunk, xh, y, s);
}

// Autotuner: on the first matvec of a sector of at least kTuneSmallDim rows, time every
// shape (XCD run length 1) on the call's own buffers, then (from kTuneMinDim rows) every XCD
// run length for the fastest shape, then the shapes that came within 20 % of the fastest
// at the fastest run length (deterministic kernels, so the repeated launches leave y as
// one launch would), and keep the fastest pair.  A candidate whose first timed launch is
// within 10 % of the best so far is timed twice more and keeps the minimum, so work of
// other streams running meanwhile inflates it less; a slower one keeps its single time.
// Each kernel instantiation gets one untimed warm-up launch.  A process-wide cache keyed
// by the model's shape parameters and the sector dimension rounded to a quarter octave
// skips re-tuning (all momentum sectors of a model share it).
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int, int, int, int>, std::pair<int, int>> g_tune_cache;
// Sectors from kTuneMinDim rows are tuned over shapes and XCD run lengths with single
// launches; smaller ones (down to kTuneSmallDim rows) over the shapes only, at run length
// 1 (their grids are a few workgroups per XCD), each candidate timed over back-to-back
// launches (their kernels take microseconds, so one launch between two events would time
// the launch rather than the kernel).  Below kTuneSmallDim rows the default shape runs.
constexpr int64_t kTuneMinDim = 1 << 17;
constexpr int64_t kTuneSmallDim = 512;

// launches per timed sample: about kTuneMinDim rows' worth of work, at most 32
static int tune_reps(int64_t dim) {
    if (dim >= kTuneMinDim) return 1;
    return (int)std::min<int64_t>(32, std::max<int64_t>(1, kTuneMinDim / std::max<int64_t>(dim, 1)));
}

// Time one (shape, run length): an optional warm-up launch, one timed sample, and two more
// when the first is within 10 % of `best` (the best time so far); the minimum.  A sample
// is `reps` back-to-back launches, reported per launch.
template <typename R>
static int time_shape(sed_sector* sec, int shape, int chunk, const typename C2<R>::T* xh, typename C2<R>::T* y,
                      hipStream_t s, hipEvent_t e0, hipEvent_t e1, bool warm, float best, float* ms_out, int reps = 1) {
    if (warm) SED_TRY(run_any_shape<R>(sec, shape, chunk, xh, y, s));
    float mine = 1e30f;
    for (int k = 0; k < 3; ++k) {
        SED_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) SED_TRY(run_any_shape<R>(sec, shape, chunk, xh, y, s));
        SED_HIP(hipEventRecord(e1, s));
        SED_HIP(hipEventSynchronize(e1));
        float ms = 0;
        SED_HIP(hipEventElapsedTime(&ms, e0, e1));
        mine = std::min(mine, ms / (float)reps);
        if (k == 0 && mine > 1.1f * best) break;
    }
    *ms_out = mine;
    return SED_OK;
}

template <typename R>
static int pick_shape(sed_sector* sec, const typename C2<R>::T* xh, typename C2<R>::T* y, hipStream_t s) {
    const int pi = sizeof(R) == 8 ? 0 : 1;
    if (sec->mv_shape[pi] >= 0) return sec->mv_shape[pi];
    if (const char* env = std::getenv("SED_MATVEC_SHAPE")) {
        // fixed launch (profiling): shape from the environment, run length SED_MATVEC_CHUNK
        const int v = std::atoi(env);
        const char* ce = std::getenv("SED_MATVEC_CHUNK");
        if (v >= 0 && v < kNumShapes) {
            sec->mv_chunk[pi] = ce ? std::max(0, std::atoi(ce)) : 1;
            return sec->mv_shape[pi] = v;
        }
    }
    if (sec->dim < kTuneSmallDim) {
        sec->mv_chunk[pi] = 1;
        return sec->mv_shape[pi] = kDefaultShape;
    }
    const bool small = sec->dim < kTuneMinDim;
    const int reps = tune_reps(sec->dim);
    const auto key = std::make_tuple(sec->cat->n_site, sec->cat->n_filled, sec->cat->n_group, sec->cat->stats,
                                     sec->n_terms, (int)std::lround(4.0 * std::log2((double)sec->dim)), pi);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            sec->mv_chunk[pi] = it->second.second;
            return sec->mv_shape[pi] = it->second.first;
        }
    }
    hipEvent_t e0, e1;
    SED_HIP(hipEventCreate(&e0));
    SED_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_i = kDefaultShape, best_c = 1;
    for (int i = 0; i < kNumShapes; ++i) {
        float ms = 0;
        if (!shape_fits<R>(sec, i)) {
            sec->tune_ms[pi][i] = -1.0f;
            continue;
        }
        SED_TRY(time_shape<R>(sec, i, 1, xh, y, s, e0, e1, true, best, &ms, reps));
        sec->tune_ms[pi][i] = ms;
        if (ms < best) {
            best = ms;
            best_i = i;
        }
    }
    for (int c = 0; c < kNumChunks && small; ++c) sec->tune_ms[pi][kNumShapes + c] = -1.0f;
    for (int i = 0; i < kNumShapes && small; ++i) sec->tune_ms[pi][kNumShapes + kNumChunks + i] = -1.0f;
    for (int c = 0; c < kNumChunks && !small; ++c) {
        float ms = 0;
        if (kChunks[c] == 1) ms = sec->tune_ms[pi][best_i];
        else SED_TRY(time_shape<R>(sec, best_i, kChunks[c], xh, y, s, e0, e1, false, best, &ms));
        sec->tune_ms[pi][kNumShapes + c] = ms;
        if (ms < best) {
            best = ms;
            best_c = kChunks[c];
        }
    }
    // the close shapes once more at the chosen run length (-1: not fitting or not close)
    const int shape_c = best_i;
    const float best1 = sec->tune_ms[pi][shape_c];
    for (int i = 0; i < kNumShapes && !small; ++i) {
        float ms = 0;
        if (!shape_fits<R>(sec, i) || !(sec->tune_ms[pi][i] <= 1.2f * best1)) {
            sec->tune_ms[pi][kNumShapes + kNumChunks + i] = -1.0f;
            continue;
        }
        if (i == shape_c) ms = sec->tune_ms[pi][kNumShapes + (int)(std::find(kChunks, kChunks + kNumChunks, best_c) - kChunks)];
        else SED_TRY(time_shape<R>(sec, i, best_c, xh, y, s, e0, e1, false, best, &ms));
        sec->tune_ms[pi][kNumShapes + kNumChunks + i] = ms;
        if (ms < best) {
            best = ms;
            best_i = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {best_i, best_c};
    }
    sec->mv_chunk[pi] = best_c;
    return sec->mv_shape[pi] = best_i;
}

template <typename R>
static int run_mv(sed_sector* sec, const typename C2<R>::T* xh, typename C2<R>::T* y, hipStream_t s) {
    const int shape = pick_shape<R>(sec, xh, y, s);
    if (shape < 0) return shape;
    return run_any_shape<R>(sec, shape, sec->mv_chunk[sizeof(R) == 8 ? 0 : 1], xh, y, s);
}

int matvec_shape_info(const sed_sector* sec, int precision, int* lanes, int* batch, int* chunk, float* tune_ms,
                      int* n_chunks, int* block) {
    const int pi = precision == SED_PREC_FP64 ? 0 : 1;
    const int sh = sec->mv_shape[pi];
    if (lanes) *lanes = sh >= 0 ? kShapes[sh].lanes : 0;
    if (batch) *batch = sh >= 0 ? kShapes[sh].batch : 0;
    if (block) *block = sh >= 0 ? kShapes[sh].block : 0;
    if (chunk) *chunk = sec->mv_chunk[pi];
    if (n_chunks) *n_chunks = kNumChunks;
    if (tune_ms)
        for (int i = 0; i < 2 * kNumShapes + kNumChunks; ++i) tune_ms[i] = sec->tune_ms[pi][i];
    return kNumShapes;
}

template <typename R>
static int matvec_any(sed_sector* sec, const typename C2<R>::T* x, typename C2<R>::T* y, typename C2<R>::T* xh,
                      hipStream_t s) {
    using T = typename C2<R>::T;
    if (!sec->has_terms) return fail(SED_ESTATE, "matvec before sed_sector_set_terms");
    if (sec->dim == 0) return SED_OK;
    const int64_t g = std::min<int64_t>((sec->dim + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_prescale<T>, dim3((unsigned)g), dim3(256), 0, s, x, sec->d_stab.p, xh, sec->dim);
    SED_HIP(hipGetLastError());
    return run_mv<R>(sec, xh, y, s);
}

int launch_matvec64(sed_sector* sec, const double* x, double* y, double* xhat, hipStream_t s) {
    return matvec_any<double>(sec, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y),
                              reinterpret_cast<double2*>(xhat), s);
}
int launch_matvec32(sed_sector* sec, const float* x, float* y, float* xhat, hipStream_t s) {
    return matvec_any<float>(sec, reinterpret_cast<const float2*>(x), reinterpret_cast<float2*>(y),
                             reinterpret_cast<float2*>(xhat), s);
}

// H|psi> kernel alone on an already prescaled xhat = sqrt(stab) x (Lanczos steps, timing).
int launch_matvec_xhat(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s) {
    if (sec->dim == 0) return SED_OK;
    if (fp64) return run_mv<double>(sec, reinterpret_cast<const double2*>(xhat), reinterpret_cast<double2*>(y), s);
    return run_mv<float>(sec, reinterpret_cast<const float2*>(xhat), reinterpret_cast<float2*>(y), s);
}

// The same without the autotuner (no events, no host synchronisation): the shape an
// earlier matvec or solve chose for this precision, else the default shape at run length 1.
int launch_matvec_xhat_fixed(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s) {
    if (sec->dim == 0) return SED_OK;
    const int pi = fp64 ? 0 : 1;
    const int shape = sec->mv_shape[pi] >= 0 ? sec->mv_shape[pi] : kDefaultShape;
    const int chunk = sec->mv_shape[pi] >= 0 ? sec->mv_chunk[pi] : 1;
    if (fp64)
        return run_any_shape<double>(sec, shape, chunk, reinterpret_cast<const double2*>(xhat),
                                     reinterpret_cast<double2*>(y), s);
    return run_any_shape<float>(sec, shape, chunk, reinterpret_cast<const float2*>(xhat),
                                reinterpret_cast<float2*>(y), s);
}

int launch_gauge64(sed_sector* sec, double* v, bool forward, hipStream_t s) {
    if (!sec->has_gauge || sec->dim == 0) return SED_OK;
    const int64_t g = std::min<int64_t>((sec->dim + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_gauge, dim3((unsigned)g), dim3(256), 0, s, reinterpret_cast<double2*>(v),
                       reinterpret_cast<const double2*>(sec->d_gauge.p), sec->dim, forward ? 1 : 0);
    SED_HIP(hipGetLastError());
    return SED_OK;
}

int launch_gauge_rows(sed_sector* sec, const double* f_dev) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((sec->dim + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(k_gauge_rows, dim3((unsigned)g), dim3(256), 0, sec->cat->ctx->stream, sec->d_rep.p,
                       reinterpret_cast<const double2*>(f_dev), sec->cat->n_site, sec->dim,
                       reinterpret_cast<double2*>(sec->d_gauge.p));
    SED_HIP(hipGetLastError());
    return SED_OK;
}

template <typename M, typename TE>
static int run_counts(sed_sector* sec, const TE* table, int64_t* gated, int64_t* ins) {
    hipStream_t s = sec->cat->ctx->stream;
    MvConst c = make_const<double>(sec);
    size_t off[9];
    const size_t lds = mv_lds_layout<double>(c, off, kMvBlock);
    const int64_t nb = std::max<int64_t>(1, (sec->dim + kMvBlock - 1) / kMvBlock);
    DevBuf<unsigned long long> part;
    SED_TRY(part.alloc(2 * nb));
    hipLaunchKernelGGL((k_hop_counts<M, TE>), dim3((unsigned)nb), dim3(kMvBlock), lds, s, c, sec->d_rep.p, table,
                       part.p);
    SED_HIP(hipGetLastError());
    std::vector<unsigned long long> h(2 * nb);
    SED_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes(), hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    unsigned long long a = 0, b = 0;
    for (int64_t i = 0; i < nb; ++i) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    *gated = (int64_t)a;
    *ins = (int64_t)b;
    return SED_OK;
}

int launch_hop_counts(sed_sector* sec, int64_t* gated, int64_t* in_sector) {
    if (sec->dim == 0) {
        *gated = *in_sector = 0;
        return SED_OK;
    }
    if (sec->cat->n_site <= 32) return run_counts<uint32_t>(sec, sec->d_table.p, gated, in_sector);
    if (sec->wide) return run_counts<uint64_t>(sec, sec->d_table64.p, gated, in_sector);
    return run_counts<uint64_t>(sec, sec->d_table.p, gated, in_sector);
}

int launch_dense(sed_sector* sec, double* H_dev, hipStream_t s) {
    if (sec->wide) return fail(SED_EUNSUPPORTED, "dense H' of an 8-byte-table sector");
    MvConst c = make_const<double>(sec);
    size_t off[9];
    const size_t lds = mv_lds_layout<double>(c, off, kMvBlock);
    SED_HIP(hipMemsetAsync(H_dev, 0, (size_t)sec->dim * sec->dim * 16, s));
    const int64_t nb = (sec->dim + kMvBlock - 1) / kMvBlock;
    const bool F = sec->cat->stats == SED_FERMIONIC;
    auto H = reinterpret_cast<double2*>(H_dev);
    const auto* A = sec->amp ? reinterpret_cast<const double2*>(sec->amp->a64.p) : nullptr;
    if (sec->cat->n_site <= 32) {
        if (F) hipLaunchKernelGGL((k_dense<uint32_t, true>), dim3((unsigned)nb), dim3(kMvBlock), lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, A, H);
        else hipLaunchKernelGGL((k_dense<uint32_t, false>), dim3((unsigned)nb), dim3(kMvBlock), lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, A, H);
    } else {
        if (F) hipLaunchKernelGGL((k_dense<uint64_t, true>), dim3((unsigned)nb), dim3(kMvBlock), lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, A, H);
        else hipLaunchKernelGGL((k_dense<uint64_t, false>), dim3((unsigned)nb), dim3(kMvBlock), lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, A, H);
    }
    SED_HIP(hipGetLastError());
    return SED_OK;
}

}  // namespace sed
