// Matrix mode: the sector block as a sparse matrix on gfx950.
//
// Replaces build_ed_Hamiltonian_symmetry_block (src/symmetry_resolved_ed.jl:903-963),
// which loops over columns (representatives) and pushes, for every term k = (a -> b)
// gated by the column's representative, the entry
//   H[row, col] = t_k * s_JW * alpha * conj(chi(g)) * sqrt(stab_row / stab_col)
// plus the diagonal, then sums duplicates (sparse) and drops explicit zeros.  Here one
// thread owns one column r and walks its representative's hops as the H|psi> kernel
// does; a count pass sizes every column, a fill pass writes (row, value) pairs at the
// column's offset.  The values come out of the same sign-only tables as the H|psi>
// kernel: H' (sign tables, gauge-transformed terms) has H'[c, r] = conj(H'[r, c]) =
// t'_k conj(chi(g)) s sqrt(stab_c / stab_r), and with twisted boundaries
// H[c, r] = conj(F(c)) H'[c, r] F(r), F(rep) = prod_{i in rep} f(i).  With site phases
// not of gauge form (no F) each entry carries the conjugate of the target rank's hop
// amplitude, H[c, r] = t_k conj(A) conj(chi(g)) s sqrt(stab_c / stab_r).
#include <algorithm>
#include <vector>

#include "sed_device.hpp"
#include "sed_internal.hpp"
#include "sed_mv.hpp"

namespace sed {

MvConst mv_const(const sed_sector* sec, bool fp64);

template <typename M, bool FERMI, bool FILL>
__global__ void __launch_bounds__(kMvBlock) k_csc_entries(MvConst c, const uint64_t* __restrict__ rep,
                                                          const uint8_t* __restrict__ stab,
                                                          const uint32_t* __restrict__ table,
                                                          const double2* __restrict__ F,
                                                          const double2* __restrict__ amp,
                                                          const int64_t* __restrict__ off, int32_t* __restrict__ cnt,
                                                          int64_t* __restrict__ erow, double2* __restrict__ eval) {
    const MvLds<double> L = stage_mv<double>(c, kMvBlock);
    const int64_t r = (int64_t)blockIdx.x * kMvBlock + threadIdx.x;
    if (r >= c.dim) return;
    const M m = (M)rep[r];
    const double inv = 1.0 / sqrt((double)stab[r]);
    const double2 Fr = F ? F[r] : make_double2(1.0, 0.0);
    int64_t o = FILL ? off[r] + 1 : 0;   // slot 0 of the column: the diagonal
    int n = 1;
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
            ++n;
            if (!FILL) continue;
            const uint32_t col = e >> c.shift;
            const int g = (e >> 1) & c.gmask;
            int p = e & 1;
            const int ib = popc_m<M>(mp & (((M)1 << b) - 1));
            if (FERMI) p ^= (ib - ia) & 1;
            double2 cf = cmul(L.tcoef[t], L.chi[g]);                // conj(t') chi(g)
            if (amp) cf = cmul(cf, amp[rank]);                      // times the hop amplitude
            const double sc = sqrt((double)stab[col]) * inv * (p ? -1.0 : 1.0);
            double2 v = make_double2(cf.x * sc, -cf.y * sc);         // H'[col, r]
            if (F) {
                const double2 Fc = F[col];
                v = cmul(cmul(make_double2(Fc.x, -Fc.y), v), Fr);
            }
            erow[o] = col;
            eval[o] = v;
            ++o;
        }
        ++ia;
    }
    if (FILL) {
        erow[off[r]] = r;
        eval[off[r]] = make_double2(dx, dy);
    } else {
        cnt[r] = n;
    }
}

template <bool FILL>
static int launch_entries(sed_sector* sec, const int64_t* off, int32_t* cnt, int64_t* erow, double2* eval,
                          hipStream_t s) {
    MvConst c = mv_const(sec, true);
    size_t offs[9];
    const size_t lds = mv_lds_layout<double>(c, offs, kMvBlock);
    const int64_t nb = (sec->dim + kMvBlock - 1) / kMvBlock;
    const bool F = sec->cat->stats == SED_FERMIONIC;
    const double2* G = sec->has_gauge ? reinterpret_cast<const double2*>(sec->d_gauge.p) : nullptr;
    const double2* A = sec->amp ? reinterpret_cast<const double2*>(sec->amp->a64.p) : nullptr;
    const dim3 grid((unsigned)nb), blk(kMvBlock);
    if (sec->cat->n_site <= 32) {
        if (F) hipLaunchKernelGGL((k_csc_entries<uint32_t, true, FILL>), grid, blk, lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, G, A, off, cnt, erow, eval);
        else hipLaunchKernelGGL((k_csc_entries<uint32_t, false, FILL>), grid, blk, lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, G, A, off, cnt, erow, eval);
    } else {
        if (F) hipLaunchKernelGGL((k_csc_entries<uint64_t, true, FILL>), grid, blk, lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, G, A, off, cnt, erow, eval);
        else hipLaunchKernelGGL((k_csc_entries<uint64_t, false, FILL>), grid, blk, lds, s, c, sec->d_rep.p, sec->d_stab.p, sec->d_table.p, G, A, off, cnt, erow, eval);
    }
    SED_HIP(hipGetLastError());
    return SED_OK;
}

// The merged CSC (0-based here; the ABI adds the Julia base): per column, entries
// sorted by row, duplicates summed in emission order, exact zeros dropped.
int build_sector_csc(sed_sector* sec, std::vector<int64_t>& colptr, std::vector<int64_t>& rowval,
                     std::vector<double>& nzval) {
    hipStream_t s = sec->cat->ctx->stream;
    const int64_t n = sec->dim;
    colptr.assign((size_t)n + 1, 0);
    rowval.clear();
    nzval.clear();
    if (n == 0) return SED_OK;
    DevBuf<int32_t> dcnt;
    SED_TRY(dcnt.alloc((size_t)n));
    SED_TRY(launch_entries<false>(sec, nullptr, dcnt.p, nullptr, nullptr, s));
    std::vector<int32_t> cnt((size_t)n);
    SED_HIP(hipMemcpyAsync(cnt.data(), dcnt.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    dcnt.release();
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t r = 0; r < n; ++r) off[r + 1] = off[r] + cnt[r];
    const int64_t total = off[n];
    DevBuf<int64_t> doff, drow;
    DevBuf<double2> dval;
    SED_TRY(doff.alloc((size_t)n + 1));
    SED_TRY(drow.alloc((size_t)total));
    SED_TRY(dval.alloc((size_t)total));
    SED_HIP(hipMemcpyAsync(doff.p, off.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, s));
    SED_TRY(launch_entries<true>(sec, doff.p, nullptr, drow.p, dval.p, s));
    std::vector<int64_t> er((size_t)total);
    std::vector<double2> ev((size_t)total);
    SED_HIP(hipMemcpyAsync(er.data(), drow.p, (size_t)total * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipMemcpyAsync(ev.data(), dval.p, (size_t)total * 16, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    doff.release();
    drow.release();
    dval.release();
    rowval.reserve((size_t)total);
    nzval.reserve((size_t)2 * total);
    std::vector<int64_t> idx;
    for (int64_t col = 0; col < n; ++col) {
        const int64_t a = off[col], b = off[col + 1];
        idx.resize((size_t)(b - a));
        for (int64_t i = a; i < b; ++i) idx[(size_t)(i - a)] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return er[x] < er[y]; });
        for (size_t q = 0; q < idx.size();) {
            const int64_t row = er[idx[q]];
            double re = 0.0, im = 0.0;
            for (; q < idx.size() && er[idx[q]] == row; ++q) {
                re += ev[idx[q]].x;
                im += ev[idx[q]].y;
            }
            if (re == 0.0 && im == 0.0) continue;   // dropzeros!
            rowval.push_back(row);
            nzval.push_back(re);
            nzval.push_back(im);
        }
        colptr[col + 1] = (int64_t)rowval.size();
    }
    return SED_OK;
}

}  // namespace sed
