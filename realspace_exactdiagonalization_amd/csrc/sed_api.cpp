// C ABI of libsed: handles, validation, host-side preparation of the device tables.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <utility>

#include "sed_internal.hpp"
#include "sed_mv.hpp"

namespace sed {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int launch_gauge_rows(sed_sector* sec, const double* f_dev);
template <typename R>
size_t matvec_lds_bytes(const sed_sector* sec);
int launch_density(sed_sector* sec, const double2* vec, double* density);
int launch_pair_density(sed_sector* sec, const double2* vec, std::vector<double>& M);
int build_sector_csc(sed_sector* sec, std::vector<int64_t>& colptr, std::vector<int64_t>& rowval,
                     std::vector<double>& nzval);
int launch_fock_expand(sed_sector* sec, const double2* c, double atol, double2* out, hipStream_t s);
int launch_position(sed_sector* to, sed_sector* from, const double2* phi, int64_t* rows, double2* vals,
                    hipStream_t s);
int launch_cdot(const double2* x, const double2* y, int64_t n, double2* part, std::vector<double2>& host,
                hipStream_t s, cd* out);
int launch_convert(const double* a, float* b, int64_t n, hipStream_t s);
int launch_convert(const float* a, double* b, int64_t n, hipStream_t s);
int launch_matvec_xhat(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s);
int matvec_shape_info(const sed_sector* sec, int precision, int* lanes, int* batch, int* chunk, float* tune_ms,
                      int* n_chunks, int* block);

template <typename T>
static int upload(DevBuf<T>& b, const T* h, size_t n, hipStream_t s) {
    SED_TRY(b.alloc(n));
    if (n) SED_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return SED_OK;
}

static int max_lds(const sed_context* ctx) { return ctx->max_lds; }

static std::atomic<int64_t> g_term_version{0};

static void free_context(sed_context* ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->ws) {
        ctx->ws->release();
        delete ctx->ws;
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void context_child_added(sed_context* ctx) { ++ctx->n_children; }
void context_child_removed(sed_context* ctx) {
    if (--ctx->n_children == 0 && ctx->released) free_context(ctx);
}

static void free_catalog(sed_catalog* cat) {
    sed_context* ctx = cat->ctx;
    (void)hipSetDevice(ctx->device);
    cat->d_binom.release();
    cat->d_perm.release();
    cat->d_inv.release();
    cat->d_rep_rank.release();
    cat->d_sym_par.release();
    cat->d_orbit_rep.release();
    cat->d_orbit_stab.release();
    cat->d_orbit_stab_par.release();
    delete cat;
    context_child_removed(ctx);
}

}  // namespace sed

using namespace sed;

extern "C" {

const char* sed_last_error(void) { return g_last_error.c_str(); }

int sed_version(void) { return 200; }

int sed_create(int device, sed_context** out) {
    if (!out) return fail(SED_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    SED_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SED_EINVAL, "device " + std::to_string(device) + " out of range");
    SED_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    SED_HIP(hipGetDeviceProperties(&p, device));
    auto* ctx = new sed_context();
    ctx->device = device;
    ctx->n_cu = p.multiProcessorCount;
    ctx->max_lds = (int)p.sharedMemPerBlock;
    ctx->arch = p.gcnArchName;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(SED_EDEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    ctx->ws = new Workspace();
    ctx->ws->stream = ctx->stream;
    // the no-allocation matvec's scratch array on this device (its own error is kept for
    // that entry point; every other entry point works without it)
    (void)noalloc_scratch_init(device);
    *out = ctx;
    return SED_OK;
}

int sed_destroy(sed_context* ctx) {
    if (!ctx) return SED_OK;
    if (ctx->n_children > 0) {
        ctx->released = true;   // freed with its last catalog / server
        return SED_OK;
    }
    free_context(ctx);
    return SED_OK;
}

// Site phases of the group (null = all 1): the inverse indices of
// _inverse_operation_indices (:109-117) and the gauge form f of the phases.  The orbit
// tables depend on the permutations only.
static int set_group_phases(sed_catalog* cat, const double* phases) {
    const int n_site = cat->n_site, n_group = cat->n_group;
    cat->phases.resize((size_t)n_group * n_site);
    for (size_t i = 0; i < cat->phases.size(); ++i)
        cat->phases[i] = phases ? cd(phases[2 * i], phases[2 * i + 1]) : cd(1.0, 0.0);
    // _inverse_operation_indices (:109-117): first h with h o g = id and matching phases
    cat->inv.assign(n_group, -1);
    for (int g = 0; g < n_group; ++g) {
        for (int h = 0; h < n_group && cat->inv[g] < 0; ++h) {
            bool ok = true;
            for (int i = 0; i < n_site && ok; ++i) {
                const int gi = cat->perm[(size_t)g * n_site + i];
                if (cat->perm[(size_t)h * n_site + gi - 1] != i + 1) ok = false;
            }
            for (int i = 0; i < n_site && ok; ++i) {
                const int gi = cat->perm[(size_t)g * n_site + i];
                if (std::abs(cat->phases[(size_t)g * n_site + i] * cat->phases[(size_t)h * n_site + gi - 1] - 1.0) > 1e-12)
                    ok = false;
            }
            if (ok) cat->inv[g] = h;
        }
        if (cat->inv[g] < 0) return fail(SED_EINVAL, "Symmetry operations are not closed under inversion.");
    }
    // gauge form of the site phases: eta_g(i) = f(i) / f(perm_g(i))
    cat->gauge_f.assign(n_site, cd(0.0));
    std::vector<char> have(n_site, 0);
    cat->has_phases = false;
    for (const cd& e : cat->phases)
        if (e != cd(1.0, 0.0)) cat->has_phases = true;
    for (int s0 = 0; s0 < n_site; ++s0) {
        if (have[s0]) continue;
        have[s0] = 1;
        cat->gauge_f[s0] = 1.0;
        std::vector<int> stack{s0};
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            for (int g = 0; g < n_group; ++g) {
                const int t = cat->perm[(size_t)g * n_site + s] - 1;
                if (!have[t]) {
                    have[t] = 1;
                    cat->gauge_f[t] = cat->gauge_f[s] / cat->phases[(size_t)g * n_site + s];
                    stack.push_back(t);
                }
            }
        }
    }
    cat->general = false;
    for (int g = 0; g < n_group && !cat->general; ++g)
        for (int i = 0; i < n_site; ++i) {
            const int t = cat->perm[(size_t)g * n_site + i] - 1;
            if (std::abs(cat->phases[(size_t)g * n_site + i] * cat->gauge_f[t] - cat->gauge_f[i]) > 1e-10) {
                cat->general = true;
                break;
            }
        }
    if (!cat->general) return SED_OK;
    // Not of gauge form: accepted when the operators U_g|s> = alpha_g(s)|g(s)> form a
    // representation, U_g U_h = U_{g o h}, i.e. for every pair the composed permutation is
    // an element k of the group and eta_k(i) = eta_h(i) eta_g(perm_h(i)) (a cocycle);
    // only then are the projected sector states and the reference's sector matrix what
    // they are meant to be.
    cat->gauge_f.assign(n_site, cd(1.0, 0.0));
    std::vector<int> comp(n_site);
    for (int g = 0; g < n_group; ++g)
        for (int h = 0; h < n_group; ++h) {
            for (int i = 0; i < n_site; ++i)
                comp[i] = cat->perm[(size_t)g * n_site + cat->perm[(size_t)h * n_site + i] - 1];
            int k = -1;
            for (int c = 0; c < n_group && k < 0; ++c)
                if (std::equal(comp.begin(), comp.end(), cat->perm.begin() + (size_t)c * n_site)) k = c;
            if (k < 0)
                return fail(SED_EUNSUPPORTED, "site phases not of gauge form f(i)/f(perm(i)) on permutations "
                                              "that are not closed under composition");
            for (int i = 0; i < n_site; ++i) {
                const int hi = cat->perm[(size_t)h * n_site + i] - 1;
                const cd want = cat->phases[(size_t)h * n_site + i] * cat->phases[(size_t)g * n_site + hi];
                if (std::abs(cat->phases[(size_t)k * n_site + i] - want) > 1e-10)
                    return fail(SED_EUNSUPPORTED, "site phases are neither of gauge form f(i)/f(perm(i)) nor a "
                                                  "representation of the group (eta_gh(i) != eta_h(i) eta_g(h(i)))");
            }
        }
    return SED_OK;
}

int sed_catalog_build(sed_context* ctx, int n_site, int n_filled, int statistics, int n_group, int identity_idx,
                      const int32_t* perm, const double* phases, sed_catalog** out) {
    if (!ctx || !perm || !out) return fail(SED_EINVAL, "NULL argument");
    *out = nullptr;
    if (n_site < 1 || n_site > 64) return fail(SED_EUNSUPPORTED, "n_site must be in 1..64");
    if (n_filled < 0 || n_filled > n_site) return fail(SED_EINVAL, "n_filled out of range");
    if (statistics != SED_BOSONIC && statistics != SED_FERMIONIC) return fail(SED_EINVAL, "bad statistics");
    if (n_group < 1 || n_group > 64) return fail(SED_EUNSUPPORTED, "group order must be in 1..64");
    if (identity_idx < 1 || identity_idx > n_group) return fail(SED_EINVAL, "identity_idx out of range");
    SED_HIP(hipSetDevice(ctx->device));
    auto* cat = new sed_catalog();
    std::unique_ptr<sed_catalog> guard(cat);
    cat->ctx = ctx;
    cat->n_site = n_site;
    cat->n_filled = n_filled;
    cat->stats = statistics;
    cat->n_group = n_group;
    cat->identity_idx = identity_idx;
    cat->perm.assign(perm, perm + (size_t)n_group * n_site);
    // permutations: 1-based, bijective
    for (int g = 0; g < n_group; ++g) {
        std::vector<char> seen(n_site, 0);
        for (int i = 0; i < n_site; ++i) {
            const int t = cat->perm[(size_t)g * n_site + i];
            if (t < 1 || t > n_site || seen[t - 1])
                return fail(SED_EINVAL, "perm of group element " + std::to_string(g + 1) + " is not a permutation");
            seen[t - 1] = 1;
        }
    }
    for (int i = 0; i < n_site; ++i)
        if (cat->perm[(size_t)(identity_idx - 1) * n_site + i] != i + 1)
            return fail(SED_EINVAL, "Identity must have trivial permutation");
    SED_TRY(set_group_phases(cat, phases));
    // binomials C(n, j), n <= n_site, j <= n_filled + 1 (saturating at 2^32-1)
    uint64_t tot = 1;
    {
        std::vector<std::vector<uint64_t>> C(n_site + 1, std::vector<uint64_t>(n_site + 2, 0));
        for (int a = 0; a <= n_site; ++a) {
            C[a][0] = 1;
            for (int b = 1; b <= a; ++b) {
                const uint64_t v = C[a - 1][b - 1] + (b <= a - 1 ? C[a - 1][b] : 0);
                C[a][b] = std::min<uint64_t>(v, (uint64_t)1 << 62);
            }
        }
        tot = C[n_site][n_filled];
        // ranks are uint32 (the binomial table saturates at 2^32 - 1, above every rank)
        if (tot >= 0xffffffffull)
            return fail(SED_EUNSUPPORTED, "C(n_site, n_filled) = " + std::to_string(tot) + " exceeds 2^32 - 2");
        cat->bstride = n_filled + 2;
        cat->binom.assign((size_t)(n_site + 1) * cat->bstride, 0);
        for (int a = 0; a <= n_site; ++a)
            for (int b = 0; b < cat->bstride && b <= n_site + 1; ++b)
                cat->binom[(size_t)a * cat->bstride + b] = (uint32_t)std::min<uint64_t>(C[a][b], 0xffffffffull);
    }
    cat->n_total = (int64_t)tot;
    // per-rank hop amplitudes (24 B per rank) are refused where every sector of typical size
    // would need the 8-byte table, which they do not run on
    if (cat->general && cat->n_total >= ((int64_t)1 << 31))
        return fail(SED_EUNSUPPORTED, "site phases not of gauge form on a catalog of " + std::to_string(tot) +
                                          " ranks (at most 2^31 - 1)");
    hipStream_t s = ctx->stream;
    std::vector<uint8_t> p8((size_t)n_group * n_site), inv8(n_group);
    for (size_t i = 0; i < p8.size(); ++i) p8[i] = (uint8_t)(cat->perm[i] - 1);
    for (int g = 0; g < n_group; ++g) inv8[g] = (uint8_t)cat->inv[g];
    SED_TRY(upload(cat->d_binom, cat->binom.data(), cat->binom.size(), s));
    SED_TRY(upload(cat->d_perm, p8.data(), p8.size(), s));
    SED_TRY(upload(cat->d_inv, inv8.data(), inv8.size(), s));
    SED_TRY(launch_catalog(cat));
    if (cat->general) {
        cat->amp = std::make_shared<HopAmp>();
        SED_TRY(launch_hop_amp(cat, cat->amp.get()));
    }
    *out = guard.release();
    context_child_added(ctx);
    return SED_OK;
}

int sed_catalog_info(const sed_catalog* cat, int64_t* n_total, int64_t* n_orbits) {
    if (!cat) return fail(SED_EINVAL, "NULL catalog");
    if (n_total) *n_total = cat->n_total;
    if (n_orbits) *n_orbits = cat->n_orbits;
    return SED_OK;
}

int sed_catalog_copy(const sed_catalog* cat, int32_t* rep_rank, int32_t* rep_sym, double* rep_amp, uint64_t* reps,
                     int64_t* stab_order) {
    if (!cat) return fail(SED_EINVAL, "NULL catalog");
    SED_HIP(hipSetDevice(cat->ctx->device));
    hipStream_t s = cat->ctx->stream;
    const int64_t n = cat->n_total, no = cat->n_orbits;
    if (rep_rank) {
        SED_HIP(hipMemcpyAsync(rep_rank, cat->d_rep_rank.p, n * 4, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) rep_rank[i] = -rep_rank[i];  // -(orbit+1) -> orbit+1
    }
    if (rep_sym) {
        std::vector<uint8_t> sp(n);
        SED_HIP(hipMemcpyAsync(sp.data(), cat->d_sym_par.p, n, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) rep_sym[i] = (sp[i] & 0x7f) + 1;
    }
    if (rep_amp) {
        DevBuf<double> amp;
        SED_TRY(amp.alloc((size_t)2 * n));
        SED_TRY(launch_catalog_amp(cat, amp.p));
        SED_HIP(hipMemcpy(rep_amp, amp.p, (size_t)16 * n, hipMemcpyDeviceToHost));
    }
    if (reps) {
        SED_HIP(hipMemcpyAsync(reps, cat->d_orbit_rep.p, no * 8, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
    }
    if (stab_order) {
        std::vector<uint64_t> st(no);
        SED_HIP(hipMemcpyAsync(st.data(), cat->d_orbit_stab.p, no * 8, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < no; ++i) stab_order[i] = __builtin_popcountll(st[i]);
    }
    return SED_OK;
}

int sed_catalog_stabilizers(const sed_catalog* cat, int32_t* stab_gidx, double* stab_phase) {
    if (!cat) return fail(SED_EINVAL, "NULL catalog");
    SED_HIP(hipSetDevice(cat->ctx->device));
    hipStream_t s = cat->ctx->stream;
    const int64_t no = cat->n_orbits;
    const int nG = cat->n_group, N = cat->n_site;
    std::vector<uint64_t> st(no), reps(no);
    SED_HIP(hipMemcpyAsync(st.data(), cat->d_orbit_stab.p, no * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipMemcpyAsync(reps.data(), cat->d_orbit_rep.p, no * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    for (int64_t o = 0; o < no; ++o) {
        int j = 0;
        for (int h = 0; h < nG; ++h) {
            if (!((st[o] >> h) & 1)) continue;
            // alpha_h(rep): reference product of site phases and reordering sign
            uint64_t m = reps[o], img = 0;
            cd a(1.0, 0.0);
            int par = 1;
            while (m) {
                const int i = __builtin_ctzll(m);
                m &= m - 1;
                const int p = cat->perm[(size_t)h * N + i];  // 1-based
                if (cat->stats == SED_FERMIONIC) {
                    const uint64_t above = p >= 64 ? 0 : (img >> p);
                    if (__builtin_popcountll(above) & 1) par = -par;
                }
                img |= 1ull << (p - 1);
                const cd b = cat->phases[(size_t)h * N + i];
                a = cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
            }
            a *= (double)par;
            if (stab_gidx) stab_gidx[o * nG + j] = h + 1;
            if (stab_phase) {
                stab_phase[2 * (o * nG + j)] = a.real();
                stab_phase[2 * (o * nG + j) + 1] = a.imag();
            }
            ++j;
        }
        for (; j < nG; ++j) {
            if (stab_gidx) stab_gidx[o * nG + j] = 0;
            if (stab_phase) stab_phase[2 * (o * nG + j)] = stab_phase[2 * (o * nG + j) + 1] = 0.0;
        }
    }
    return SED_OK;
}

int sed_catalog_update_phases(sed_catalog* cat, const double* phases) {
    if (!cat) return fail(SED_EINVAL, "NULL catalog");
    SED_HIP(hipSetDevice(cat->ctx->device));
    const std::vector<cd> old_phases = cat->phases, old_f = cat->gauge_f;
    const std::vector<int32_t> old_inv = cat->inv;
    const bool old_has = cat->has_phases, old_general = cat->general;
    int rc = set_group_phases(cat, phases);
    if (rc == SED_OK && cat->inv != old_inv)
        rc = fail(SED_EUNSUPPORTED, "the new phases change the inverse group elements (rebuild the catalog)");
    if (rc == SED_OK && cat->general && cat->n_total >= ((int64_t)1 << 31))
        rc = fail(SED_EUNSUPPORTED, "site phases not of gauge form on a catalog of more than 2^31 - 1 ranks");
    std::shared_ptr<HopAmp> amp;   // sectors built before keep the table of their phases
    if (rc == SED_OK && cat->general) {
        amp = std::make_shared<HopAmp>();
        rc = launch_hop_amp(cat, amp.get());
    }
    if (rc != SED_OK) {
        cat->phases = old_phases;
        cat->gauge_f = old_f;
        cat->inv = old_inv;
        cat->has_phases = old_has;
        cat->general = old_general;
        return rc;
    }
    cat->amp = amp;
    return SED_OK;
}

int sed_catalog_destroy(sed_catalog* cat) {
    if (!cat) return SED_OK;
    if (cat->n_sectors_alive > 0) {
        cat->released = true;   // freed with its last sector
        return SED_OK;
    }
    free_catalog(cat);
    return SED_OK;
}

int sed_sector_build(sed_catalog* cat, const double* chi, double atol, sed_sector** out, int64_t* dim) {
    if (!cat || !chi || !out) return fail(SED_EINVAL, "NULL argument");
    *out = nullptr;
    SED_HIP(hipSetDevice(cat->ctx->device));
    auto* sec = new sed_sector();
    std::unique_ptr<sed_sector> guard(sec);
    sec->cat = cat;
    sec->atol = atol;
    sec->chi.resize(cat->n_group);
    for (int g = 0; g < cat->n_group; ++g) sec->chi[g] = cd(chi[2 * g], chi[2 * g + 1]);
    hipStream_t s = cat->ctx->stream;
    SED_TRY(upload(sec->d_chi64, chi, (size_t)2 * cat->n_group, s));
    std::vector<float> c32(2 * cat->n_group);
    for (int i = 0; i < 2 * cat->n_group; ++i) c32[i] = (float)chi[i];
    SED_TRY(upload(sec->d_chi32, c32.data(), c32.size(), s));
    SED_TRY(launch_sector(sec, sec->d_chi64.p));
    {
        // characters with the hop sign folded in, indexed like a table entry's low bits
        // (g << 1 | sign), and a zero entry for hops that do not land in the sector
        const int nc = (1 << sec->shift) + 1;
        std::vector<double> c2(2 * (size_t)nc, 0.0);
        for (int g = 0; g < cat->n_group && (2 * g + 1) < nc - 1; ++g)
            for (int sg = 0; sg < 2; ++sg) {
                c2[2 * (2 * g + sg)] = sg ? -chi[2 * g] : chi[2 * g];
                c2[2 * (2 * g + sg) + 1] = sg ? -chi[2 * g + 1] : chi[2 * g + 1];
            }
        std::vector<float> c2f(c2.begin(), c2.end());
        SED_TRY(upload(sec->d_chi2_64, c2.data(), c2.size(), s));
        SED_TRY(upload(sec->d_chi2_32, c2f.data(), c2f.size(), s));
        SED_HIP(hipStreamSynchronize(s));   // the host arrays end with this scope
    }
    sec->has_gauge = cat->has_phases && !cat->general;
    sec->gauge_f = cat->gauge_f;
    sec->phases = cat->phases;
    if (sec->has_gauge && sec->dim > 0) {
        std::vector<double> f(2 * cat->n_site);
        for (int i = 0; i < cat->n_site; ++i) {
            f[2 * i] = cat->gauge_f[i].real();
            f[2 * i + 1] = cat->gauge_f[i].imag();
        }
        DevBuf<double> fd;
        SED_TRY(upload(fd, f.data(), f.size(), s));
        SED_TRY(sec->d_gauge.alloc((size_t)2 * sec->dim));
        SED_TRY(launch_gauge_rows(sec, fd.p));
        SED_HIP(hipStreamSynchronize(s));
    }
    if (dim) *dim = sec->dim;
    ++cat->n_sectors_alive;
    *out = guard.release();
    return SED_OK;
}

int sed_sector_basis(const sed_sector* sec, uint64_t* reps, int64_t* stab_order) {
    if (!sec) return fail(SED_EINVAL, "NULL sector");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    hipStream_t s = sec->cat->ctx->stream;
    if (reps) SED_HIP(hipMemcpyAsync(reps, sec->d_rep.p, sec->dim * 8, hipMemcpyDeviceToHost, s));
    if (stab_order) {
        std::vector<uint8_t> st(sec->dim);
        SED_HIP(hipMemcpyAsync(st.data(), sec->d_stab.p, sec->dim, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < sec->dim; ++i) stab_order[i] = st[i];
    }
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

int sed_sector_set_terms(sed_sector* sec, int n_bil, const int32_t* i_from, const int32_t* i_to, const double* t,
                         int n_den, const int32_t* di, const int32_t* dj, const double* v) {
    if (!sec) return fail(SED_EINVAL, "NULL sector");
    if (n_bil < 0 || n_den < 0) return fail(SED_EINVAL, "negative term count");
    if ((n_bil && (!i_from || !i_to || !t)) || (n_den && (!di || !dj || !v)))
        return fail(SED_EINVAL, "NULL term array");
    const sed_catalog* cat = sec->cat;
    const int N = cat->n_site;
    SED_HIP(hipSetDevice(cat->ctx->device));
    const int old_terms = sec->has_terms ? sec->n_terms : -1, old_dens = sec->has_terms ? sec->n_dens : -1;
    // merge hops with equal (from, to): their matrix contributions add
    std::map<std::pair<int, int>, cd> hops;
    for (int k = 0; k < n_bil; ++k) {
        const int a = i_from[k], b = i_to[k];
        if (a < 1 || a > N || b < 1 || b > N) return fail(SED_EINVAL, "bilinear term site out of range");
        if (a == b) return fail(SED_EINVAL, "bilinear term with i_from == i_to changes the particle number");
        // gauge-transformed amplitude t' = t f(to)/f(from); the kernel uses conj(t')
        const cd tp = cd(t[2 * k], t[2 * k + 1]) * sec->gauge_f[b - 1] / sec->gauge_f[a - 1];
        hops[{a - 1, b - 1}] += tp;
    }
    std::map<std::pair<int, int>, cd> dens;  // ordered by first site
    std::vector<std::vector<std::pair<int, cd>>> dl(N);
    for (int k = 0; k < n_den; ++k) {
        const int a = di[k], b = dj[k];
        if (a < 1 || a > N || b < 1 || b > N) return fail(SED_EINVAL, "density term site out of range");
        dl[a - 1].push_back({b - 1, cd(v[2 * k], v[2 * k + 1])});
    }
    sec->term_start.assign(N + 1, 0);
    sec->term_to.clear();
    sec->term_coef.clear();
    for (int a = 0; a < N; ++a) {
        sec->term_start[a] = (int)sec->term_to.size();
        for (auto it = hops.lower_bound({a, -1}); it != hops.end() && it->first.first == a; ++it) {
            sec->term_to.push_back((uint8_t)it->first.second);
            sec->term_coef.push_back(std::conj(it->second));
        }
    }
    sec->term_start[N] = (int)sec->term_to.size();
    sec->dens_start.assign(N + 1, 0);
    sec->dens_j.clear();
    sec->dens_v.clear();
    for (int a = 0; a < N; ++a) {
        sec->dens_start[a] = (int)sec->dens_j.size();
        for (auto& pr : dl[a]) {
            sec->dens_j.push_back((uint8_t)pr.first);
            sec->dens_v.push_back(pr.second);
        }
    }
    sec->dens_start[N] = (int)sec->dens_j.size();
    sec->n_terms = (int)sec->term_to.size();
    sec->n_dens = (int)sec->dens_j.size();
    hipStream_t s = cat->ctx->stream;
    std::vector<double> tc64(2 * std::max(1, sec->n_terms)), dv64(2 * std::max(1, sec->n_dens));
    std::vector<float> tc32(tc64.size()), dv32(dv64.size());
    for (int i = 0; i < sec->n_terms; ++i) {
        tc64[2 * i] = sec->term_coef[i].real();
        tc64[2 * i + 1] = sec->term_coef[i].imag();
    }
    for (int i = 0; i < sec->n_dens; ++i) {
        dv64[2 * i] = sec->dens_v[i].real();
        dv64[2 * i + 1] = sec->dens_v[i].imag();
    }
    for (size_t i = 0; i < tc64.size(); ++i) tc32[i] = (float)tc64[i];
    for (size_t i = 0; i < dv64.size(); ++i) dv32[i] = (float)dv64[i];
    std::vector<uint8_t> tto = sec->term_to, djv = sec->dens_j;
    if (tto.empty()) tto.push_back(0);
    if (djv.empty()) djv.push_back(0);
    SED_TRY(upload(sec->d_term_start, sec->term_start.data(), sec->term_start.size(), s));
    SED_TRY(upload(sec->d_dens_start, sec->dens_start.data(), sec->dens_start.size(), s));
    SED_TRY(upload(sec->d_term_to, tto.data(), tto.size(), s));
    SED_TRY(upload(sec->d_dens_j, djv.data(), djv.size(), s));
    SED_TRY(upload(sec->d_term_coef64, tc64.data(), tc64.size(), s));
    SED_TRY(upload(sec->d_dens_v64, dv64.data(), dv64.size(), s));
    SED_TRY(upload(sec->d_term_coef32, tc32.data(), tc32.size(), s));
    SED_TRY(upload(sec->d_dens_v32, dv32.data(), dv32.size(), s));
    // padded copy: each site's list to a multiple of kTermPad, padding hops a -> a with 0
    {
        std::vector<int32_t> ps(N + 1, 0);
        std::vector<uint32_t> info;
        std::vector<double> pc64;
        for (int a = 0; a < N; ++a) {
            ps[a] = (int32_t)info.size();
            const int t0 = sec->term_start[a], t1 = sec->term_start[a + 1];
            const int cnt = t1 - t0, padded = (cnt + kTermPad - 1) / kTermPad * kTermPad;
            for (int i = 0; i < padded; ++i) {
                const int b = i < cnt ? sec->term_to[t0 + i] : a;
                info.push_back((uint32_t)b | ((uint32_t)(b * cat->bstride) << 8));
                pc64.push_back(i < cnt ? tc64[2 * (t0 + i)] : 0.0);
                pc64.push_back(i < cnt ? tc64[2 * (t0 + i) + 1] : 0.0);
            }
        }
        ps[N] = (int32_t)info.size();
        sec->n_pterms = (int)info.size();
        if (info.empty()) {
            info.push_back(0);
            pc64.assign(2, 0.0);
        }
        std::vector<float> pc32(pc64.begin(), pc64.end());
        SED_TRY(upload(sec->d_pterm_start, ps.data(), ps.size(), s));
        SED_TRY(upload(sec->d_pterm_info, info.data(), info.size(), s));
        SED_TRY(upload(sec->d_pterm_coef64, pc64.data(), pc64.size(), s));
        SED_TRY(upload(sec->d_pterm_coef32, pc32.data(), pc32.size(), s));
        SED_HIP(hipStreamSynchronize(s));   // the host arrays end with this scope
    }
    SED_HIP(hipStreamSynchronize(s));
    const size_t need = std::max(matvec_lds_bytes<double>(sec), matvec_lds_bytes<float>(sec));
    if ((int)need > max_lds(cat->ctx))
        return fail(SED_EUNSUPPORTED, "term lists need " + std::to_string(need) + " bytes of LDS per workgroup");
    sec->has_terms = true;
    sec->hops_gated = sec->hops_in_sector = -1;
    if (sec->n_terms != old_terms || sec->n_dens != old_dens) {
        // a launch shape chosen (and LDS-checked) for other term lists: choose again
        for (int p = 0; p < 2; ++p) {
            sec->mv_shape[p] = -1;
            sec->mv_chunk[p] = 1;
        }
    }
    sec->csc_ready = false;   // a kept matrix belongs to the previous terms
    std::vector<int64_t>().swap(sec->csc_colptr);
    std::vector<int64_t>().swap(sec->csc_rowval);
    std::vector<double>().swap(sec->csc_nzval);
    // process-wide unique: a captured Lanczos step keyed on (sector address, version)
    // can never match a later sector that reuses a destroyed sector's address
    sec->version = ++g_term_version;
    return SED_OK;
}

int sed_sector_hop_counts(sed_sector* sec, int64_t* hops_gated, int64_t* hops_in_sector) {
    if (!sec) return fail(SED_EINVAL, "NULL sector");
    if (!sec->has_terms) return fail(SED_ESTATE, "hop counts before sed_sector_set_terms");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    if (sec->hops_gated < 0) SED_TRY(launch_hop_counts(sec, &sec->hops_gated, &sec->hops_in_sector));
    if (hops_gated) *hops_gated = sec->hops_gated;
    if (hops_in_sector) *hops_in_sector = sec->hops_in_sector;
    return SED_OK;
}

// y = H x with x, y complex double vectors on the device, reference basis
static int matvec_dev64(sed_sector* sec, const double* x, double* y, int precision) {
    hipStream_t s = sec->cat->ctx->stream;
    const int64_t n = sec->dim;
    DevBuf<double> xp;
    SED_TRY(xp.alloc((size_t)2 * n));
    SED_HIP(hipMemcpyAsync(xp.p, x, (size_t)16 * n, hipMemcpyDeviceToDevice, s));
    SED_TRY(launch_gauge64(sec, xp.p, true, s));
    Workspace& ws = *sec->cat->ctx->ws;
    if (precision == SED_PREC_FP64) {
        SED_TRY(grow(ws.xh64, (size_t)2 * n));
        SED_TRY(launch_matvec64(sec, xp.p, y, ws.xh64.p, s));
    } else {
        DevBuf<float> x32, y32;
        SED_TRY(x32.alloc((size_t)2 * n));
        SED_TRY(y32.alloc((size_t)2 * n));
        SED_TRY(grow(ws.xh32, (size_t)2 * n));
        SED_TRY(launch_convert(xp.p, x32.p, n, s));
        SED_TRY(launch_matvec32(sec, x32.p, y32.p, ws.xh32.p, s));
        SED_TRY(launch_convert(y32.p, y, n, s));
    }
    SED_TRY(launch_gauge64(sec, y, false, s));
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

int sed_matvec(sed_sector* sec, const double* x, double* y, int precision) {
    if (!sec || !x || !y) return fail(SED_EINVAL, "NULL argument");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32) return fail(SED_EINVAL, "bad precision");
    if (!sec->has_terms) return fail(SED_ESTATE, "matvec before sed_sector_set_terms");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    const int64_t n = sec->dim;
    if (n == 0) return SED_OK;
    hipStream_t s = sec->cat->ctx->stream;
    DevBuf<double> dx, dy;
    SED_TRY(dx.alloc((size_t)2 * n));
    SED_TRY(dy.alloc((size_t)2 * n));
    SED_HIP(hipMemcpyAsync(dx.p, x, (size_t)16 * n, hipMemcpyHostToDevice, s));
    SED_TRY(matvec_dev64(sec, dx.p, dy.p, precision));
    SED_HIP(hipMemcpyAsync(y, dy.p, (size_t)16 * n, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    return SED_OK;
}

int sed_matvec_device(sed_sector* sec, const void* x_dev, void* y_dev, int precision) {
    if (!sec || !x_dev || !y_dev) return fail(SED_EINVAL, "NULL argument");
    if (!sec->has_terms) return fail(SED_ESTATE, "matvec before sed_sector_set_terms");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    hipStream_t s = sec->cat->ctx->stream;
    if (sec->has_gauge && precision == SED_PREC_FP64)
        return matvec_dev64(sec, (const double*)x_dev, (double*)y_dev, precision);
    Workspace& ws = *sec->cat->ctx->ws;
    if (sec->has_gauge && precision == SED_PREC_FP32) {
        // x' = D x (sector basis -> gauge-free basis), y = D^-1 H' x'
        DevBuf<float> xp;
        SED_TRY(xp.alloc((size_t)2 * sec->dim));
        SED_TRY(launch_gauge32(sec, (const float*)x_dev, xp.p, true, s));
        SED_TRY(grow(ws.xh32, (size_t)2 * sec->dim));
        SED_TRY(launch_matvec32(sec, xp.p, (float*)y_dev, ws.xh32.p, s));
        SED_TRY(launch_gauge32(sec, (const float*)y_dev, (float*)y_dev, false, s));
        SED_HIP(hipStreamSynchronize(s));   // xp is released on return
        return SED_OK;
    }
    if (precision == SED_PREC_FP64) {
        SED_TRY(grow(ws.xh64, (size_t)2 * sec->dim));
        return launch_matvec64(sec, (const double*)x_dev, (double*)y_dev, ws.xh64.p, s);
    }
    if (precision == SED_PREC_FP32) {
        SED_TRY(grow(ws.xh32, (size_t)2 * sec->dim));
        return launch_matvec32(sec, (const float*)x_dev, (float*)y_dev, ws.xh32.p, s);
    }
    return fail(SED_EINVAL, "bad precision");
}

int sed_matvec_benchmark(sed_sector* sec, int precision, int n_calls, double* ms_per_call) {
    if (!sec || !ms_per_call || n_calls < 1) return fail(SED_EINVAL, "bad argument");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32) return fail(SED_EINVAL, "bad precision");
    if (!sec->has_terms) return fail(SED_ESTATE, "matvec before sed_sector_set_terms");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    hipStream_t s = sec->cat->ctx->stream;
    const int64_t n = sec->dim;
    const bool fp64 = precision == SED_PREC_FP64;
    const size_t es = fp64 ? 16 : 8;
    DevBuf<uint8_t> x, y, xh;
    SED_TRY(x.alloc(es * n));
    SED_TRY(y.alloc(es * n));
    SED_TRY(xh.alloc(es * n));
    // x = 1 + i (any dense vector: every hop gathers a live value)
    {
        std::vector<double> h((size_t)2 * n, 1.0);
        DevBuf<double> d;
        SED_TRY(d.alloc((size_t)2 * n));
        SED_HIP(hipMemcpyAsync(d.p, h.data(), (size_t)16 * n, hipMemcpyHostToDevice, s));
        if (fp64) SED_HIP(hipMemcpyAsync(x.p, d.p, (size_t)16 * n, hipMemcpyDeviceToDevice, s));
        else SED_TRY(launch_convert(d.p, (float*)x.p, n, s));
        SED_HIP(hipStreamSynchronize(s));
    }
    // one full call (prescale + kernel) fills xhat; timed calls run the H|psi> kernel alone
    if (fp64) SED_TRY(launch_matvec64(sec, (const double*)x.p, (double*)y.p, (double*)xh.p, s));
    else SED_TRY(launch_matvec32(sec, (const float*)x.p, (float*)y.p, (float*)xh.p, s));
    for (int i = 0; i < 2; ++i) SED_TRY(launch_matvec_xhat(sec, xh.p, y.p, fp64, s));
    hipEvent_t ev[2] = {nullptr, nullptr};
    float ms = 0;
    auto timed = [&]() -> int {
        SED_HIP(hipEventCreate(&ev[0]));
        SED_HIP(hipEventCreate(&ev[1]));
        SED_HIP(hipEventRecord(ev[0], s));
        for (int i = 0; i < n_calls; ++i) SED_TRY(launch_matvec_xhat(sec, xh.p, y.p, fp64, s));
        SED_HIP(hipEventRecord(ev[1], s));
        SED_HIP(hipEventSynchronize(ev[1]));
        SED_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        return SED_OK;
    };
    const int rc = timed();
    (void)hipStreamSynchronize(s);   // the buffers are freed on return
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == SED_OK) *ms_per_call = ms / n_calls;
    return rc;
}

int sed_sector_set_shape(sed_sector* sec, int precision, int shape, int xcd_chunk) {
    if (!sec) return fail(SED_EINVAL, "NULL sector");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32) return fail(SED_EINVAL, "bad precision");
    const int n = matvec_shape_info(sec, precision, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (shape < -1 || shape >= n) return fail(SED_EINVAL, "shape out of range");
    if (shape >= 0 && xcd_chunk < 0) return fail(SED_EINVAL, "xcd_chunk must be >= 0");
    const int pi = precision == SED_PREC_FP64 ? 0 : 1;
    sec->mv_shape[pi] = shape;
    sec->mv_chunk[pi] = shape >= 0 ? xcd_chunk : 1;
    return SED_OK;
}

int sed_sector_shape(const sed_sector* sec, int precision, int* shape, int* lanes, int* batch, int* xcd_chunk,
                     int* n_shapes, int* n_chunks, float* tune_ms, int* workgroup) {
    if (!sec) return fail(SED_EINVAL, "NULL sector");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32) return fail(SED_EINVAL, "bad precision");
    const int n = matvec_shape_info(sec, precision, lanes, batch, xcd_chunk, tune_ms, n_chunks, workgroup);
    if (shape) *shape = sec->mv_shape[precision == SED_PREC_FP64 ? 0 : 1];
    if (n_shapes) *n_shapes = n;
    return SED_OK;
}

int sed_solve(sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter, int precision, double* evals,
              double* evecs, sed_solve_stats* stats) {
    if (!sec || !evals) return fail(SED_EINVAL, "NULL argument");
    if (nev < 1) return fail(SED_EINVAL, "nev must be >= 1");
    if (!(tol > 0)) return fail(SED_EINVAL, "tol must be > 0");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32 && precision != SED_PREC_FP32_REFINE)
        return fail(SED_EINVAL, "bad precision");
    if (!sec->has_terms) return fail(SED_ESTATE, "solve before sed_sector_set_terms");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    std::vector<double> ev;
    DevBuf<double2> vecs;
    const int rc = solve_sector(sec, nev, tol, seed, maxiter < 1 ? 300 : maxiter, precision, ev, &vecs, stats,
                                *sec->cat->ctx->ws);
    const int64_t n = sec->dim;
    const int k = (int)ev.size();
    for (int i = 0; i < k; ++i) evals[i] = ev[i];
    for (int i = k; i < nev; ++i) evals[i] = NAN;
    sec->last_evals = ev;
    sec->last_nev = k;
    sec->last_evecs.clear();
    if (k > 0 && n > 0 && vecs.p) {
        hipStream_t s = sec->cat->ctx->stream;
        for (int i = 0; i < k; ++i) SED_TRY(launch_gauge64(sec, (double*)(vecs.p + (size_t)i * n), false, s));
        sec->last_evecs.resize((size_t)n * k);
        SED_HIP(hipMemcpyAsync(sec->last_evecs.data(), vecs.p, (size_t)16 * n * k, hipMemcpyDeviceToHost, s));
        SED_HIP(hipStreamSynchronize(s));
        if (evecs) std::memcpy(evecs, sec->last_evecs.data(), (size_t)16 * n * k);
    }
    return rc;
}

int sed_sector_density(sed_sector* sec, int which, double* density) {
    if (!sec || !density) return fail(SED_EINVAL, "NULL argument");
    if (which < 0 || which >= sec->last_nev) return fail(SED_ESTATE, "no eigenvector " + std::to_string(which));
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    DevBuf<double2> v;
    SED_TRY(v.alloc(sec->dim));
    SED_HIP(hipMemcpy(v.p, sec->last_evecs.data() + (size_t)which * sec->dim, (size_t)16 * sec->dim,
                      hipMemcpyHostToDevice));
    return launch_density(sec, v.p, density);
}

// the requested eigenvector of the last solve, uploaded (reference basis)
static int upload_eigvec(sed_sector* sec, int which, DevBuf<double2>& v) {
    if (which < 0 || which >= sec->last_nev) return fail(SED_ESTATE, "no eigenvector " + std::to_string(which));
    SED_TRY(v.alloc(sec->dim));
    SED_HIP(hipMemcpy(v.p, sec->last_evecs.data() + (size_t)which * sec->dim, (size_t)16 * sec->dim,
                      hipMemcpyHostToDevice));
    return SED_OK;
}

int sed_sector_density_correlation(sed_sector* sec, int which, double* nn, double* density) {
    if (!sec || !nn || !density) return fail(SED_EINVAL, "NULL argument");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    DevBuf<double2> v;
    SED_TRY(upload_eigvec(sec, which, v));
    std::vector<double> M;
    const int rc = launch_pair_density(sec, v.p, M);
    v.release();
    SED_TRY(rc);
    // each image g(rep) occupies site i iff rep occupies pinv_g(i), perm_g(pinv_g(i)) = i:
    // <n_i n_j> = (1/|G|) sum_g M[pinv_g(i)][pinv_g(j)]
    const sed_catalog* cat = sec->cat;
    const int n = cat->n_site, nG = cat->n_group;
    std::vector<int> pinv((size_t)nG * n);
    for (int g = 0; g < nG; ++g)
        for (int a = 0; a < n; ++a) pinv[(size_t)g * n + (cat->perm[(size_t)g * n + a] - 1)] = a;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double sm = 0.0;
            for (int g = 0; g < nG; ++g) sm += M[(size_t)pinv[(size_t)g * n + i] * n + pinv[(size_t)g * n + j]];
            nn[(size_t)i * n + j] = sm / nG;
        }
    for (int i = 0; i < n; ++i) density[i] = nn[(size_t)i * n + i];
    return SED_OK;
}

// rho_ij = <psi| a_i^dag a_j |psi>.  A sector state is an eigenvector of every U_g, so
// rho_ij = <psi| O_ij |psi> with O_ij = (1/|G|) sum_g U_g a_i^dag a_j U_g^dag
//        = (1/|G|) sum_g eta_g(i) conj(eta_g(j)) a_{g(i)}^dag a_{g(j)}
// (U_g a_i^dag U_g^dag = eta_g(i) a_{g(i)}^dag with the group's site phases), a
// translation-invariant hopping operator: the sector H|psi> kernel applies it.  The
// kernel gathers rows through Hermiticity, so it is fed the Hermitian parts
// A = O_ij + O_ji and B = i (O_ij - O_ji): rho_ij = (<A> - i <B>) / 2, rho_ji = conj.
int sed_sector_one_body(sed_sector* sec, int which, double* rho) {
    if (!sec || !rho) return fail(SED_EINVAL, "NULL argument");
    sed_catalog* cat = sec->cat;
    SED_HIP(hipSetDevice(cat->ctx->device));
    const int n = cat->n_site, nG = cat->n_group;
    const int64_t dim = sec->dim;
    hipStream_t s = cat->ctx->stream;
    DevBuf<double2> x;
    SED_TRY(upload_eigvec(sec, which, x));
    std::vector<double> dens(n);
    SED_TRY(launch_density(sec, x.p, dens.data()));
    for (int i = 0; i < 2 * n * n; ++i) rho[i] = 0.0;
    for (int i = 0; i < n; ++i) rho[2 * ((size_t)i * n + i)] = dens[i];
    SED_TRY(launch_gauge64(sec, reinterpret_cast<double*>(x.p), true, s));   // primed basis
    // a scratch sector on this sector's tables with the pair operators as its terms
    sed_sector tmp;
    tmp.cat = cat;
    tmp.chi = sec->chi;
    tmp.dim = dim;
    tmp.shift = sec->shift;
    tmp.d_rep.borrow(sec->d_rep);
    tmp.d_stab.borrow(sec->d_stab);
    tmp.d_table.borrow(sec->d_table);
    tmp.wide = sec->wide;
    tmp.d_table64.borrow(sec->d_table64);
    tmp.amp = sec->amp;
    tmp.d_chi64.borrow(sec->d_chi64);
    tmp.d_chi32.borrow(sec->d_chi32);
    tmp.d_chi2_64.borrow(sec->d_chi2_64);
    tmp.d_chi2_32.borrow(sec->d_chi2_32);
    tmp.has_gauge = sec->has_gauge;
    tmp.d_gauge.borrow(sec->d_gauge);
    tmp.gauge_f = sec->gauge_f;
    tmp.phases = sec->phases;
    tmp.mv_shape[0] = sec->mv_shape[0] >= 0 ? sec->mv_shape[0] : 1;
    tmp.mv_chunk[0] = sec->mv_shape[0] >= 0 ? sec->mv_chunk[0] : 1;
    DevBuf<double2> y, xh, part;
    int rc = y.alloc(dim);
    if (rc == SED_OK) rc = xh.alloc(dim);
    if (rc == SED_OK) rc = part.alloc(256);
    std::vector<int32_t> from(2 * nG), to(2 * nG);
    std::vector<double> t(4 * nG);
    std::vector<double2> host;
    for (int i = 0; i < n && rc == SED_OK; ++i)
        for (int j = i + 1; j < n && rc == SED_OK; ++j) {
            double ev[2] = {0.0, 0.0};
            for (int part_b = 0; part_b < 2 && rc == SED_OK; ++part_b) {
                for (int g = 0; g < nG; ++g) {
                    const cd tg = sec->phases[(size_t)g * n + i] * std::conj(sec->phases[(size_t)g * n + j]) /
                                  (double)nG;
                    const cd f = part_b ? cd(0.0, 1.0) * tg : tg;   // O_ij part; the O_ji part is its conjugate
                    from[2 * g] = cat->perm[(size_t)g * n + j];
                    to[2 * g] = cat->perm[(size_t)g * n + i];
                    from[2 * g + 1] = cat->perm[(size_t)g * n + i];
                    to[2 * g + 1] = cat->perm[(size_t)g * n + j];
                    t[4 * g] = f.real();
                    t[4 * g + 1] = f.imag();
                    t[4 * g + 2] = f.real();
                    t[4 * g + 3] = -f.imag();
                }
                rc = sed_sector_set_terms(&tmp, 2 * nG, from.data(), to.data(), t.data(), 0, nullptr, nullptr,
                                          nullptr);
                if (rc == SED_OK)
                    rc = launch_matvec64(&tmp, reinterpret_cast<const double*>(x.p), reinterpret_cast<double*>(y.p),
                                         reinterpret_cast<double*>(xh.p), s);
                cd dot = 0.0;
                if (rc == SED_OK) rc = launch_cdot(x.p, y.p, dim, part.p, host, s, &dot);
                ev[part_b] = dot.real();
            }
            const cd r(0.5 * ev[0], -0.5 * ev[1]);   // (<A> - i <B>) / 2
            rho[2 * ((size_t)i * n + j)] = r.real();
            rho[2 * ((size_t)i * n + j) + 1] = r.imag();
            rho[2 * ((size_t)j * n + i)] = r.real();
            rho[2 * ((size_t)j * n + i) + 1] = -r.imag();
        }
    // the scratch sector's own term lists and the work vectors are freed on return (the
    // borrowed tables are not); nothing is in flight after launch_cdot's synchronization
    SED_HIP(hipStreamSynchronize(s));
    return rc;
}

int sed_sector_matrix_csc(sed_sector* sec, int64_t* nnz, int64_t* colptr, int64_t* rowval, double* nzval) {
    if (!sec || !nnz) return fail(SED_EINVAL, "NULL argument");
    if (!sec->has_terms) return fail(SED_ESTATE, "matrix before sed_sector_set_terms");
    const bool query = !colptr && !rowval && !nzval;
    if (!query && (!colptr || !rowval || !nzval)) return fail(SED_EINVAL, "colptr, rowval and nzval: all or none");
    if (sec->wide) return fail(SED_EUNSUPPORTED, "sparse matrix of a sector with an 8-byte table (dimension " +
                                                     std::to_string(sec->dim) + ")");
    SED_HIP(hipSetDevice(sec->cat->ctx->device));
    if (!sec->csc_ready) {
        SED_TRY(build_sector_csc(sec, sec->csc_colptr, sec->csc_rowval, sec->csc_nzval));
        sec->csc_ready = true;
    }
    *nnz = (int64_t)sec->csc_rowval.size();
    if (query) return SED_OK;
    for (size_t i = 0; i < sec->csc_colptr.size(); ++i) colptr[i] = sec->csc_colptr[i] + 1;
    for (size_t i = 0; i < sec->csc_rowval.size(); ++i) rowval[i] = sec->csc_rowval[i] + 1;
    std::memcpy(nzval, sec->csc_nzval.data(), sec->csc_nzval.size() * sizeof(double));
    std::vector<int64_t>().swap(sec->csc_colptr);
    std::vector<int64_t>().swap(sec->csc_rowval);
    std::vector<double>().swap(sec->csc_nzval);
    sec->csc_ready = false;
    return SED_OK;
}

// mask of colex rank r (k-subsets of {0..n-1}), from the catalog's binomial table
static uint64_t unrank_mask(const sed_catalog* cat, uint64_t r) {
    uint64_t m = 0;
    int b = cat->n_site - 1;
    for (int j = cat->n_filled; j >= 1; --j) {
        while (cat->binom[(size_t)b * cat->bstride + j] > r) --b;
        m |= (uint64_t)1 << b;
        r -= cat->binom[(size_t)b * cat->bstride + j];
        --b;
    }
    return m;
}

int sed_sector_fock_expansion(sed_sector* sec, const double* c, double atol, int64_t capacity, int64_t* n_out,
                              uint64_t* masks, double* amps) {
    if (!sec || !c || !n_out) return fail(SED_EINVAL, "NULL argument");
    if (capacity < 0 || (capacity > 0 && (!masks || !amps))) return fail(SED_EINVAL, "bad output arrays");
    const sed_catalog* cat = sec->cat;
    SED_HIP(hipSetDevice(cat->ctx->device));
    hipStream_t s = cat->ctx->stream;
    const int64_t n = sec->dim, nt = cat->n_total;
    *n_out = 0;
    if (n == 0) return fail(SED_EINVAL, "Expanded state has near-zero norm.");
    DevBuf<double2> dc, out;
    SED_TRY(dc.alloc((size_t)n));
    SED_TRY(out.alloc((size_t)nt));
    SED_HIP(hipMemcpyAsync(dc.p, c, (size_t)n * 16, hipMemcpyHostToDevice, s));
    SED_TRY(launch_fock_expand(sec, dc.p, atol, out.p, s));
    std::vector<double2> h((size_t)nt);
    SED_HIP(hipMemcpy(h.data(), out.p, (size_t)nt * 16, hipMemcpyDeviceToHost));
    dc.release();
    out.release();
    double norm2 = 0.0;
    int64_t cnt = 0;
    for (int64_t k = 0; k < nt; ++k) {
        const double a = std::hypot(h[k].x, h[k].y);
        if (a > atol) {
            norm2 += h[k].x * h[k].x + h[k].y * h[k].y;
            ++cnt;
        }
    }
    if (!(norm2 > atol)) return fail(SED_EINVAL, "Expanded state has near-zero norm.");
    *n_out = cnt;
    if (capacity == 0) return SED_OK;
    if (capacity < cnt) return fail(SED_EINVAL, "capacity " + std::to_string(capacity) + " < " + std::to_string(cnt));
    const double inv = 1.0 / std::sqrt(norm2);
    int64_t q = 0;
    for (int64_t k = 0; k < nt; ++k) {
        if (!(std::hypot(h[k].x, h[k].y) > atol)) continue;
        masks[q] = unrank_mask(cat, (uint64_t)k);   // colex rank order = ascending masks
        amps[2 * q] = h[k].x * inv;
        amps[2 * q + 1] = h[k].y * inv;
        ++q;
    }
    return SED_OK;
}

int sed_sector_position_operator(sed_sector* to, sed_sector* from, const double* site_phases, int64_t* rows,
                                 double* vals) {
    if (!to || !from || !site_phases || !rows || !vals) return fail(SED_EINVAL, "NULL argument");
    if (to->cat != from->cat) return fail(SED_EINVAL, "sectors of different catalogs (symmetry groups)");
    if (to->wide) return fail(SED_EUNSUPPORTED, "position operator into a sector with an 8-byte table");
    const sed_catalog* cat = from->cat;
    SED_HIP(hipSetDevice(cat->ctx->device));
    hipStream_t s = cat->ctx->stream;
    const int64_t n = from->dim;
    if (n == 0) return SED_OK;
    DevBuf<double2> phi, dv;
    DevBuf<int64_t> dr;
    SED_TRY(phi.alloc((size_t)cat->n_site));
    SED_TRY(dv.alloc((size_t)n));
    SED_TRY(dr.alloc((size_t)n));
    SED_HIP(hipMemcpyAsync(phi.p, site_phases, (size_t)cat->n_site * 16, hipMemcpyHostToDevice, s));
    SED_TRY(launch_position(to, from, phi.p, dr.p, dv.p, s));
    SED_HIP(hipMemcpyAsync(rows, dr.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    SED_HIP(hipMemcpyAsync(vals, dv.p, (size_t)n * 16, hipMemcpyDeviceToHost, s));
    SED_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) rows[i] += 1;   // 1-based, 0 = no entry
    phi.release();
    dv.release();
    dr.release();
    return SED_OK;
}

int sed_sector_destroy(sed_sector* sec) {
    if (!sec) return SED_OK;
    (void)hipSetDevice(sec->cat->ctx->device);
    --sec->cat->n_sectors_alive;
    // the synchronous workspace's captured step references this sector's tables
    Workspace* ws = sec->cat->ctx->ws;
    if (ws && ws->graph_key.sec == sec) ws->drop_graph();
    (void)hipStreamSynchronize(sec->cat->ctx->stream);
    sed_catalog* cat = sec->cat;
    delete sec;   // frees the sector's device tables
    if (cat->n_sectors_alive == 0 && cat->released) free_catalog(cat);
    return SED_OK;
}

}  // extern "C"
