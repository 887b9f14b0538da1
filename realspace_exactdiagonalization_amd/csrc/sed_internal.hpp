// Internal host-side structures and helpers of libsed (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sed.h"

namespace sed {

using cd = std::complex<double>;

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SED_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::sed::fail(SED_EDEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define SED_TRY(call)            \
    do {                         \
        int rc_ = (call);        \
        if (rc_ != SED_OK) return rc_; \
    } while (0)

// Device buffer.  It owns its allocation (freed by release() or when it goes out of
// scope; move-only), or, after borrow(), points at another buffer's allocation without
// ever freeing it.  Code that frees a buffer a kernel may still be reading synchronizes
// its stream first.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) {
        o.p = nullptr;
        o.n = 0;
        o.owned = true;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owned = o.owned;
            o.p = nullptr;
            o.n = 0;
            o.owned = true;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(SED_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
        return SED_OK;
    }
    void borrow(const DevBuf& o) {
        release();
        p = o.p;
        n = o.n;
        owned = false;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct Workspace;

// Per-rank hop amplitudes of a group whose site phases are not of gauge form: for rank s
// in the orbit of rep with stored element h (h(s) = rep), prod_{i in rep} eta_g(i) with
// g = h^-1 (the phase part of the reference's conj(rep_amp_table[s])), FP64 and FP32.
// Sectors keep the table of the phases they were built with (shared with the catalog).
struct HopAmp {
    DevBuf<double> a64;   // complex interleaved, n_total
    DevBuf<float> a32;
};

}  // namespace sed

struct sed_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    int max_lds = 65536;              // LDS bytes per workgroup
    std::string arch;
    sed::Workspace* ws = nullptr;     // workspace of the synchronous API (on `stream`)
    int64_t n_children = 0;           // live catalogs and servers
    bool released = false;            // sed_destroy called while children were alive
};

struct sed_catalog {
    sed_context* ctx = nullptr;
    int n_site = 0, n_filled = 0, stats = 0, n_group = 0, identity_idx = 1;
    int64_t n_total = 0, n_orbits = 0;
    std::vector<int32_t> perm;        // 1-based, n_group x n_site
    std::vector<sed::cd> phases;      // n_group x n_site
    std::vector<int32_t> inv;         // 0-based inverse element index
    std::vector<sed::cd> gauge_f;     // f(i): eta_g(i) = f(i)/f(perm_g(i)) (all 1 if general)
    bool has_phases = false;          // any eta != 1
    // phases not of gauge form (a representation of the group with non-trivial phases on
    // sites that some element fixes): hops take their amplitude from the per-rank table
    bool general = false;
    std::shared_ptr<sed::HopAmp> amp;  // general phases only
    std::vector<uint32_t> binom;      // (n_site+1) x bstride, saturating
    int bstride = 0;
    // device state
    sed::DevBuf<uint32_t> d_binom;
    sed::DevBuf<uint8_t> d_perm;      // 0-based targets, n_group x n_site
    sed::DevBuf<uint8_t> d_inv;
    sed::DevBuf<int32_t> d_rep_rank;  // per rank: -(orbit+1)
    sed::DevBuf<uint8_t> d_sym_par;   // per rank: group index h (bits 0-6), parity (bit 7)
    sed::DevBuf<uint64_t> d_orbit_rep;
    sed::DevBuf<uint64_t> d_orbit_stab;      // bit h set if h(rep) == rep
    sed::DevBuf<uint64_t> d_orbit_stab_par;  // bit h set if that stabilizer element has sign -1
    int64_t n_sectors_alive = 0;
    bool released = false;            // sed_catalog_destroy called while sectors were alive
};

struct sed_sector {
    sed_catalog* cat = nullptr;
    std::vector<sed::cd> chi;
    double atol = 1e-12;
    int64_t dim = 0;
    int shift = 1;                    // table entry = row << shift | g << 1 | parity
    sed::DevBuf<uint64_t> d_rep;      // sector representatives (ascending)
    sed::DevBuf<uint8_t> d_stab;      // stabilizer orders
    sed::DevBuf<uint32_t> d_table;    // per rank: packed (row, g, parity) or 0xffffffff
    // sectors whose row << shift does not fit 32 bits (more than 32 sites only) keep the
    // same entry in 8 bytes: d_table64 (no row = all ones) and wide = true, d_table empty
    bool wide = false;
    sed::DevBuf<uint64_t> d_table64;
    std::shared_ptr<sed::HopAmp> amp;  // the catalog's per-rank hop amplitudes (general phases)
    // terms
    bool has_terms = false;
    int n_terms = 0, n_dens = 0;
    std::vector<int32_t> term_start, dens_start;  // CSR by from-site (0-based), n_site+1
    std::vector<uint8_t> term_to, dens_j;
    std::vector<sed::cd> term_coef;   // conj(t'_k), gauge-transformed
    std::vector<sed::cd> dens_v;
    bool has_gauge = false;           // D = diag(F(rep)) != 1
    std::vector<sed::cd> gauge_f;     // the catalog's f(i) and site phases when this sector
    std::vector<sed::cd> phases;      // was built (a later flux refresh leaves them alone)
    sed::DevBuf<int32_t> d_term_start, d_dens_start;
    sed::DevBuf<uint8_t> d_term_to, d_dens_j;
    sed::DevBuf<double> d_term_coef64, d_dens_v64, d_chi64;  // interleaved complex
    sed::DevBuf<float> d_term_coef32, d_dens_v32, d_chi32;
    sed::DevBuf<double> d_gauge;      // F(rep_r) per row, complex interleaved (only if has_gauge)
    // the term lists padded per source site to a multiple of kTermPad (site-by-site hop loop)
    int n_pterms = 0;
    sed::DevBuf<int32_t> d_pterm_start;
    sed::DevBuf<uint32_t> d_pterm_info;    // target b | (b * binomial stride) << 8
    sed::DevBuf<double> d_pterm_coef64;
    sed::DevBuf<float> d_pterm_coef32;
    sed::DevBuf<double> d_chi2_64;         // (-1)^p chi(g) at g << 1 | p, then one 0 entry
    sed::DevBuf<float> d_chi2_32;
    int64_t hops_gated = -1, hops_in_sector = -1;
    int mv_shape[2] = {-1, -1};       // H|psi> launch shape per precision (FP64, FP32), -1 = not tuned
    int mv_chunk[2] = {1, 1};         // XCD run length of the H|psi> blocks (0 = one range per XCD)
    int64_t version = 0;              // bumped whenever device term tables are (re)uploaded
    float tune_ms[2][64] = {};        // autotuner timings: shapes, XCD run lengths, shapes again
    int busy = 0;                     // a server request is running on this sector
    // last solve
    std::vector<double> last_evals;
    std::vector<sed::cd> last_evecs;  // dim x nev, reference basis
    int last_nev = 0;
    // matrix mode: merged CSC kept between the size query and the copy (0-based)
    std::vector<int64_t> csc_colptr, csc_rowval;
    std::vector<double> csc_nzval;
    bool csc_ready = false;
};

namespace sed {

// Reusable device/pinned buffers of one solver lane (one stream).  Grown on demand,
// never shrunk, so repeated solves do not allocate in the hot loop.
// Everything a captured Lanczos step depends on: the sector and its term upload
// version, the precision, sizes, launch shape and every device address in the graph.
struct GraphKey {
    const void* sec = nullptr;
    int64_t version = -1;
    int prec = -1, m = 0, shape = -1, chunk = -1, reorth = -1;
    int64_t n = 0;
    const void* ptrs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool operator==(const GraphKey& o) const {
        if (sec != o.sec || version != o.version || prec != o.prec || m != o.m || shape != o.shape ||
            chunk != o.chunk || reorth != o.reorth || n != o.n)
            return false;
        for (int i = 0; i < 5; ++i)
            if (ptrs[i] != o.ptrs[i]) return false;
        return true;
    }
};
struct Workspace {
    hipStream_t stream = nullptr;
    DevBuf<double> V64, w64, xh64;    // complex interleaved doubles
    DevBuf<float> V32, w32, xh32;     // complex interleaved floats
    DevBuf<double> part, coef, S, v0;
    DevBuf<double> state;             // Lanczos coefficients, alpha/beta, step counter
    double* hcoef = nullptr;          // pinned host copy of alpha/beta and flags
    size_t hcoef_n = 0;
    hipGraphExec_t graph_exec = nullptr;  // one captured Lanczos step
    GraphKey graph_key;
    int ensure_host(size_t n_doubles);
    void drop_graph();
    void release();
    size_t device_bytes() const;
};
template <typename T>
int grow(DevBuf<T>& b, size_t n) {
    if (b.n >= n) return SED_OK;
    return b.alloc(n);
}

// Lifetime: a context outlives its catalogs and servers, a catalog its sectors.  Their
// destroy calls defer the release until the last dependent object is gone.
void context_child_added(sed_context* ctx);
void context_child_removed(sed_context* ctx);

// ---- kernel launchers (device code in the .hip files) ----
int launch_catalog(sed_catalog* cat);
int launch_catalog_amp(const sed_catalog* cat, double* d_amp);  // reference rep_amp_table on device
int launch_hop_amp(const sed_catalog* cat, sed::HopAmp* out);    // per-rank hop amplitudes (general phases)
int launch_sector(sed_sector* sec, const double* chi_dev);
int launch_hop_counts(sed_sector* sec, int64_t* gated, int64_t* in_sector);

// y = H' x in the gauge-free sector basis (H = D^-1 H' D, D = diag of the representatives'
// twist gauge factors); x, y device vectors.
// prescale: xhat = sqrt(stab) * x.   FP64: double2, FP32: float2.
int launch_matvec64(sed_sector* sec, const double* x, double* y, double* xhat, hipStream_t s);
int launch_matvec32(sed_sector* sec, const float* x, float* y, float* xhat, hipStream_t s);
// gauge D: v <- D^{+-1} v  (forward = multiply by F, backward by conj F)
int launch_gauge64(sed_sector* sec, double* v, bool forward, hipStream_t s);
// H|psi> on a prescaled xhat with the sector's chosen launch shape (shape 1 if none),
// never autotuning; and the per-device scratch array of the no-allocation matvec.
int launch_matvec_xhat_fixed(sed_sector* sec, const void* xhat, void* y, bool fp64, hipStream_t s);
int noalloc_scratch_init(int device);
int launch_gauge32(sed_sector* sec, const float* v, float* out, bool forward, hipStream_t s);

int solve_sector(sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter, int precision,
                 std::vector<double>& evals, DevBuf<double2>* out_vecs, sed_solve_stats* stats, Workspace& ws);
// device bytes a solve of this sector needs in a workspace
int64_t solve_footprint(const sed_sector* sec, int nev, int precision);

// ---- packed Lanczos engine: many small sectors, one workgroup each, per launch ----
struct PackJob {
    int64_t id = 0;
    sed_sector* sec = nullptr;
    int nev = 1, maxiter = 300, precision = SED_PREC_FP64;
    double tol = 1e-10;
    uint64_t seed = 0;
    int status = 1;
    std::string error;
    std::vector<double> evals;
    sed_solve_stats stats{};
};
class PackEngine;
// on_done is called from the engine thread, once per job, after its results are set.
PackEngine* pack_create(sed_context* ctx, int64_t arena_bytes, int max_active, void (*on_done)(void*, PackJob*),
                        void* user);
void pack_submit(PackEngine* e, PackJob* job);
int64_t pack_device_bytes(const PackEngine* e);
void pack_destroy(PackEngine* e);
bool pack_eligible(const sed_sector* sec, int nev);
int64_t pack_footprint(const sed_sector* sec, int nev, int precision);

// True if two of the first nev FP32 Ritz values (ascending) coincide within the FP32
// phase's accuracy.  A single start vector spans one direction of a degenerate
// eigenspace, so the FP64 refinement then starts from the seeded random vector (as a
// plain FP64 solve does) rather than from the sum of the FP32 Ritz vectors.
inline bool fp32_cluster(const std::vector<double>& theta, int nev) {
    const int k = std::min<int>(nev, (int)theta.size());
    for (int i = 0; i + 1 < k; ++i)
        if (std::abs(theta[i + 1] - theta[i]) <= 1e-5 * std::max(1.0, std::abs(theta[i]))) return true;
    return false;
}

// ---- host eigensolvers (sed_eigen.cpp) ----
// Real symmetric n x n (row-major, overwritten) -> ascending eigenvalues, eigenvectors in columns.
void jacobi_eigh(int n, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& evecs);
void eigh_last_row(int n, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& last);
// Complex Hermitian n x n (column-major, lower triangle read) -> all eigenvalues ascending
// and the eigenvectors of the nvec lowest (n x nvec column-major; nvec < 0: all).
void hermitian_eigh(int n, std::vector<cd>& a, std::vector<double>& evals, std::vector<cd>* evecs, int nvec = -1);

}  // namespace sed
