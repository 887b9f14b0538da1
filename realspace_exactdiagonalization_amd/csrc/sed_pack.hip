// Packed Lanczos engine: many small sectors advanced together in one persistent launch.
//
// No reference counterpart as a component: the reference solves each sector on its own
// with KrylovKit (diagonalize_block_matrixfree, src/symmetry_resolved_ed.jl:998-1030)
// inside the flux loop of ed_scan! (:1297-1356).  Here one workgroup of 1024 threads
// owns one sector and runs whole thick-restart Lanczos steps inside the kernel:
// H|psi> over all of the sector's rows, classical Gram-Schmidt against the stored basis,
// normalisation and the next prescaled vector,
// separated by workgroup barriers instead of kernel boundaries.  The Gram-Schmidt
// passes follow the per-sector solver's partial re-orthogonalisation (a local pass
// against v_{j-1}, v_j; the whole basis only when Simon's recurrence, a forced step or
// cancellation asks for it; state per request in device memory).  Every sector of a
// launch advances to its next check point; the host then reads all alpha/beta in one
// copy, diagonalises the projected matrices in parallel, and queues the resulting
// thick restarts and FP32 -> FP64 hand-overs as one batched kernel in front of the next
// launch.  Start vectors are drawn and normalised inside the Lanczos kernel itself, so
// admitting a request costs no launch.  Queued requests join at the next check point
// (continuous batching).
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "sed_internal.hpp"
#include "sed_mv.hpp"
#include "sed_vec.hpp"

namespace sed {

MvConst mv_const(const sed_sector* sec, bool fp64);

constexpr int kPackBlock = 1024;
constexpr int kPackCheck = 16;            // steps between host checks (flux scan, FP32+refine
                                          // solves/s: 8 -> 845, 16 -> 891, 32 -> 859)
constexpr int64_t kPackMaxDim = 1 << 15;   // one CU per sector: larger sectors go to the lanes
// The packed kernel always uses partial re-orthogonalisation: requests for more eigenpairs
// than it covers go to the lanes.
constexpr int kPackMaxNev = kProMaxNev;
constexpr int kPackBatch = 4;             // hops whose table reads (then gathers) issue together
                                          // (divides the term padding; flux scan, FP64 batched:
                                          // 1287-1305 solves/s, 1225 with the unpadded loop)
constexpr int kMoveRows = 64;             // rows per workgroup of the restart kernel
constexpr int kSlotDoubles = 2 * kMaxBasis + 2;   // alpha[m], beta[m], full steps, (int step, int flag)

enum PackStart { kStartNone = 0, kStartRandom = 1, kStartGiven = 2 };

struct PackItem {
    MvConst c;
    const uint64_t* rep;
    const uint8_t* stab;
    const uint32_t* table;
    void* V;
    void* w;
    void* xhat;
    double* slot;
    int64_t n;
    uint64_t seed;
    int m;
    int target;     // run until the step index reaches target
    int start;      // PackStart: draw (kStartRandom) or take (kStartGiven) V_0, normalise, step 0
    int set_step;   // >= 0: continue at this step (after a thick restart)
    ProState* pro;              // partial re-orthogonalisation state of the request
    const double* pro_init;     // after a thick restart: kept Ritz values, then arrowhead
    int pro_kk;                 // kept Ritz vectors of that restart (0: none)
    int pro_enabled;
    int pro_reset;              // first launch of the request: zero the full-step count
};

// One thick restart (kk > 0): V_i <- sum_{a<jm} V_a S[a*kk+i] for i < kk, then
// V_kk <- V_jm and xhat <- sqrt(stab) V_kk.  One FP32 -> FP64 hand-over (kk == 0):
// dst <- sum_{a<jm} V_a S[a] in FP64 (the sum of the wanted Ritz vectors).
struct PackMove {
    void* V;
    void* xhat;
    double2* dst;
    const uint8_t* stab;
    const double* S;
    int64_t n;
    int jm, kk;
};

template <typename M, typename R, bool FERMI>
__global__ void __launch_bounds__(kPackBlock) k_pack_lanczos(const PackItem* __restrict__ items) {
    using T = typename C2<R>::T;
    constexpr int NT = kPackBlock;
    __shared__ double2 sh_h[kMaxBasis + 2];
    __shared__ double2 sh_red[kBlockDots][NT / kWave];
    __shared__ double sh_redd[NT / kWave];
    __shared__ double sh_nrm;
    __shared__ int sh_lo, sh_need;
    const PackItem& it = items[blockIdx.x];
    ProState* pro = it.pro;
    const MvLds<R> L = stage_mv<R>(it.c, NT, false, true);
    T* V = reinterpret_cast<T*>(it.V);
    T* w = reinterpret_cast<T*>(it.w);
    T* xh = reinterpret_cast<T*>(it.xhat);
    const int64_t n = it.n;
    const int m = it.m;
    double* alpha = it.slot;
    double* beta = it.slot + m;
    int* state = reinterpret_cast<int*>(it.slot + 2 * m + 1);
    int j;
    if (it.start != kStartNone) {
        // V_0: drawn here or left by the hand-over kernel; normalised, then its xhat
        double nq = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            double2 a;
            if (it.start == kStartRandom) {
                const T t = from_d2<T>(gauss_entry(it.seed, r));
                V[r] = t;
                a = to_d2(t);
            } else {
                a = to_d2(V[r]);
            }
            nq += a.x * a.x + a.y * a.y;
        }
        nq = block_sum<NT>(nq, sh_redd);
        const double sc = nq > 0.0 ? 1.0 / sqrt(nq) : 0.0;
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            const double2 a = to_d2(V[r]);
            const T o = from_d2<T>(make_double2(a.x * sc, a.y * sc));
            V[r] = o;
            const double2 od = to_d2(o);
            const double s = sqrt((double)it.stab[r]);
            xh[r] = from_d2<T>(make_double2(od.x * s, od.y * s));
        }
        __syncthreads();
        j = 0;
        if (threadIdx.x == 0) {
            state[1] = 0;
            pro_start(pro, sizeof(R) == 8, it.pro_enabled != 0);
        }
    } else if (it.set_step >= 0) {
        j = it.set_step;
        if (threadIdx.x == 0 && it.pro_kk > 0)
            pro_restart(pro, it.pro_kk, it.pro_init, it.pro_init + it.pro_kk, alpha, beta);
    } else {
        j = state[0];
    }
    if (threadIdx.x == 0) {
        if (it.pro_reset) pro->nfull = 0;
        sh_lo = pro->lo;
    }
    __syncthreads();
    const int target = it.target;
    while (j < target) {
        // w = H' v_j from xhat = sqrt(stab) v_j
        for (int64_t r = threadIdx.x; r < n; r += NT) {
            const M mm = (M)it.rep[r];
            R ax = 0, ay = 0, dx = 0, dy = 0;
            row_sums_padded<M, R, FERMI, 1, kPackBatch, NT, uint32_t>(L, it.c, mm, 0, L.S + threadIdx.x, it.table, xh,
                                                                     ax, ay, dx, dy);
            const R inv = (R)1 / sqrt((R)it.stab[r]);
            const T xr = xh[r];
            T out;
            out.x = (ax + (dx * xr.x - dy * xr.y)) * inv;
            out.y = (ay + (dx * xr.y + dy * xr.x)) * inv;
            w[r] = out;
        }
        __syncthreads();
        // first pass over v_lo..v_j, then the whole basis when pro_decide asks for it
        const int lo = sh_lo;
        const int nv = j + 1;
        block_dots<T, NT>(V + (int64_t)lo * n, n, nv - lo, w, sh_h + lo, &sh_nrm, sh_red);
        const double nw = sh_nrm;
        double a = sh_h[j].x;
        double nw2 = block_update<T, NT>(V + (int64_t)lo * n, n, nv - lo, sh_h + lo, w, sh_redd);
        if (pro_decide_wg<NT>(pro, j, nw2, nw, a, alpha, beta, sh_redd, &sh_need)) {
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
            pro_advance(pro, j, a, b, beta);
            sh_lo = pro->lo;
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
    if (threadIdx.x == 0) {
        state[0] = j;
        it.slot[2 * m] = (double)pro->nfull;
    }
}

// Thick restarts and FP32 -> FP64 hand-overs of many sectors: grid (row tiles, moves).
// Each thread owns one row; a thick restart keeps the row's jm basis entries in LDS
// while it overwrites the first kk of them.
template <typename T>
__global__ void __launch_bounds__(kMoveRows) k_pack_move(const PackMove* __restrict__ moves) {
    extern __shared__ __align__(16) uint8_t smem[];
    const PackMove& mv = moves[blockIdx.y];
    const int64_t r0 = (int64_t)blockIdx.x * kMoveRows;
    if (r0 >= mv.n) return;   // whole workgroup leaves together
    const int64_t n = mv.n;
    const int64_t r = r0 + threadIdx.x;
    const bool live = r < n;
    T* V = reinterpret_cast<T*>(mv.V);
    const double* S = mv.S;
    const int jm = mv.jm, kk = mv.kk;
    if (kk == 0) {
        if (!live) return;
        double ax = 0.0, ay = 0.0;
        for (int a = 0; a < jm; ++a) {
            const double2 v = to_d2(V[(int64_t)a * n + r]);
            ax += v.x * S[a];
            ay += v.y * S[a];
        }
        mv.dst[r] = make_double2(ax, ay);
        return;
    }
    T* tile = reinterpret_cast<T*>(smem);
    if (live)
        for (int a = 0; a < jm; ++a) tile[a * kMoveRows + threadIdx.x] = V[(int64_t)a * n + r];
    if (!live) return;   // no barrier follows: each thread reads only its own tile column
    for (int i = 0; i < kk; ++i) {
        double ax = 0.0, ay = 0.0;
        for (int a = 0; a < jm; ++a) {
            const double2 v = to_d2(tile[a * kMoveRows + threadIdx.x]);
            const double s = S[a * kk + i];
            ax += v.x * s;
            ay += v.y * s;
        }
        V[(int64_t)i * n + r] = from_d2<T>(make_double2(ax, ay));
    }
    const T vj = V[(int64_t)jm * n + r];
    V[(int64_t)kk * n + r] = vj;
    const double2 od = to_d2(vj);
    const double s = sqrt((double)mv.stab[r]);
    reinterpret_cast<T*>(mv.xhat)[r] = from_d2<T>(make_double2(od.x * s, od.y * s));
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
    int64_t capacity() const { return (int64_t)buf_.bytes(); }
    void destroy() { buf_.release(); }

   private:
    DevBuf<uint8_t> buf_;
    int64_t bytes_ = 0;
    std::vector<std::pair<int64_t, int64_t>> free_;
};

// LDS of the packed kernel for this sector (dynamic part: term tables + rank scratch)
static size_t pack_lds(const sed_sector* sec) {
    size_t off[9];
    return std::max(mv_lds_layout<double>(mv_const(sec, true), off, kPackBlock, false, true),
                    mv_lds_layout<float>(mv_const(sec, false), off, kPackBlock, false, true));
}

bool pack_eligible(const sed_sector* sec, int nev) {
    if (sec->wide || sec->amp || sec->dim <= 500 || sec->dim > kPackMaxDim || nev > kPackMaxNev || nev > krylov_dim(sec->dim, nev) - 2)
        return false;
    // static LDS of the kernel is below 4 KiB
    return pack_lds(sec) + 4096 <= (size_t)sec->cat->ctx->max_lds;
}

static int64_t pack_bytes(int64_t n, int m, bool fp32) {
    int64_t b = (int64_t)(m + 1) * n * 16 + 2 * n * 16 + 1024;    // V64, w64, xh64 (+ alignment)
    if (fp32) b += (int64_t)(m + 1) * n * 8 + 2 * n * 8 + 768;    // V32, w32, xh32
    return b;
}

int64_t pack_footprint(const sed_sector* sec, int nev, int precision) {
    return pack_bytes(sec->dim, krylov_dim(sec->dim, nev), precision != SED_PREC_FP64);
}

struct PackReq {
    PackJob* job = nullptr;
    int64_t n = 0;
    int m = 0;
    int phase = 1;           // 0 = FP32 Lanczos, 1 = FP64 Lanczos
    int slot = -1;
    int64_t off = -1, bytes = 0;
    double *V64 = nullptr, *w64 = nullptr, *xh64 = nullptr;
    float *V32 = nullptr, *w32 = nullptr, *xh32 = nullptr;
    std::vector<double> T;   // m x m projected matrix
    int k = 0;               // kept Ritz vectors after the last restart
    int j = 0;               // device step index at the last check
    int start = kStartNone;  // pending start of the next launch
    int set_step = -1;       // pending restart step of the next launch
    int64_t pro_init_off = -1;   // pending restart data (Ritz values, arrowhead) in the coefficient buffer
    int pro_kk = 0;
    bool fresh = true;       // no launch yet
    int iters[2] = {0, 0};
    int cycles[2] = {0, 0};  // thick restarts per precision phase
    int restarts = 0;
    std::chrono::steady_clock::time_point t0;
};

// What one check point decided for one request (pure host computation).
struct Decision {
    int action = 0;          // 0 continue, 1 done, 2 hand over FP32 -> FP64, 3 thick restart,
                             // 4 FP64 phase from the random start, -1 fail
    int rc = SED_OK;
    std::string err;
    int jm = 0, kk = 0;
    std::vector<double> theta, coef;   // Ritz values; restart / hand-over coefficients
    double resmax = 0.0;
    int conv = 0;
};

class PackEngine {
   public:
    PackEngine(sed_context* ctx, int max_active, void (*cb)(void*, PackJob*), void* user)
        : ctx_(ctx), max_active_(max_active), cb_(cb), user_(user) {
        const char* e = std::getenv("SED_PRO");   // SED_PRO=0: the whole basis at every step
        pro_on_ = !(e && e[0] == '0');
    }

    int start(int64_t arena_bytes) {
        SED_HIP(hipSetDevice(ctx_->device));
        SED_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        SED_TRY(arena_.init(arena_bytes));
        const size_t coef_doubles = (size_t)max_active_ * kMaxBasis * kMaxBasis;
        SED_TRY(slots_.alloc((size_t)max_active_ * kSlotDoubles));
        SED_TRY(pros_.alloc((size_t)max_active_ * sizeof(ProState)));
        SED_TRY(items_.alloc((size_t)max_active_ * sizeof(PackItem)));
        SED_TRY(moves_.alloc((size_t)max_active_ * sizeof(PackMove)));
        SED_TRY(coef_.alloc(coef_doubles));
        SED_HIP(hipHostMalloc((void**)&h_slots_, (size_t)max_active_ * kSlotDoubles * 8, hipHostMallocDefault));
        SED_HIP(hipHostMalloc((void**)&h_items_, (size_t)max_active_ * sizeof(PackItem), hipHostMallocDefault));
        SED_HIP(hipHostMalloc((void**)&h_moves_, (size_t)max_active_ * sizeof(PackMove), hipHostMallocDefault));
        SED_HIP(hipHostMalloc((void**)&h_coef_, coef_doubles * 8, hipHostMallocDefault));
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
        if (st_) (void)hipStreamSynchronize(st_);
        if (st_) (void)hipStreamDestroy(st_);
        if (h_slots_) (void)hipHostFree(h_slots_);
        if (h_items_) (void)hipHostFree(h_items_);
        if (h_moves_) (void)hipHostFree(h_moves_);
        if (h_coef_) (void)hipHostFree(h_coef_);
        slots_.release();
        pros_.release();
        items_.release();
        moves_.release();
        coef_.release();
        arena_.destroy();
    }

   private:
    double* slot_ptr(int s) { return slots_.p + (size_t)s * kSlotDoubles; }
    ProState* pro_ptr(int s) { return reinterpret_cast<ProState*>(pros_.p) + s; }

    void finish(PackReq& q, int rc, const std::string& err) {
        PackJob* job = q.job;
        job->status = rc == SED_OK ? 0 : rc;
        job->error = err;
        job->stats.iters_fp32 = q.iters[0];
        job->stats.iters_fp64 = q.iters[1];
        job->stats.matvecs = q.iters[0] + q.iters[1];
        job->stats.full_orth_steps = q.slot >= 0 ? (int)h_slots_[(size_t)q.slot * kSlotDoubles + 2 * q.m] : 0;
        job->stats.restarts = q.restarts;
        job->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - q.t0).count();
        if (q.off >= 0) arena_.release(q.off, q.bytes);
        if (q.slot >= 0) free_slots_.push_back(q.slot);
        q.off = -1;
        q.slot = -1;
        cb_(user_, job);
    }

    // Host-side admission: arena space and a slot; the start vector is drawn by the
    // first launch.  Returns 1 when the arena has no room yet.
    int admit(PackJob* job, PackReq& q) {
        q.job = job;
        q.t0 = std::chrono::steady_clock::now();
        q.n = job->sec->dim;
        q.m = krylov_dim(q.n, job->nev);
        const bool fp32 = job->precision != SED_PREC_FP64;
        q.bytes = pack_bytes(q.n, q.m, fp32);
        q.off = arena_.alloc(q.bytes);
        if (q.off < 0) return 1;
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
        if (fp32) {
            q.V32 = reinterpret_cast<float*>(take((int64_t)(q.m + 1) * q.n * 8));
            q.w32 = reinterpret_cast<float*>(take(q.n * 8));
            q.xh32 = reinterpret_cast<float*>(take(q.n * 8));
        }
        q.phase = fp32 ? 0 : 1;
        q.start = kStartRandom;
        q.set_step = -1;
        q.pro_kk = 0;
        q.fresh = true;
        q.T.assign((size_t)q.m * q.m, 0.0);
        q.k = 0;
        q.j = 0;
        return 0;
    }

    void decide(PackReq& q, Decision& d) {
        PackJob* job = q.job;
        const int m = q.m;
        const double* hs = h_slots_ + (size_t)q.slot * kSlotDoubles;
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
        std::vector<double> a((size_t)jm * jm), Sm, last;
        auto load = [&] {
            for (int r = 0; r < jm; ++r)
                for (int c = 0; c < jm; ++c) a[(size_t)r * jm + c] = q.T[(size_t)r * m + c];
        };
        // eigenvalues and the last eigenvector row decide; the whole eigenvector matrix
        // only for a hand-over or a thick restart
        load();
        eigh_last_row(jm, a, d.theta, last);
        auto full = [&] {
            load();
            std::vector<double> th;
            jacobi_eigh(jm, a, th, Sm);
        };
        const int nev = std::min(job->nev, jm);
        const double tol = q.phase == 0 ? (job->precision == SED_PREC_FP32 ? std::max(job->tol, 1e-5) : 1e-4)
                                        : job->tol;
        for (int i = 0; i < nev; ++i) {
            const double res = beta * std::fabs(last[i]);
            if (res <= tol && d.conv == i) ++d.conv;
            d.resmax = std::max(d.resmax, res);
        }
        d.jm = jm;
        if (d.conv >= job->nev || breakdown) {
            if (d.conv < job->nev) {
                d.action = -1;
                d.rc = SED_ENOCONV;
                d.err = "Lanczos (packed) did not converge the requested eigenpairs";
            } else if (q.phase == 0 && job->precision == SED_PREC_FP32_REFINE && fp32_cluster(d.theta, job->nev)) {
                d.action = 4;   // degenerate FP32 Ritz values: FP64 phase from the seeded random vector
            } else if (q.phase == 0 && job->precision == SED_PREC_FP32_REFINE) {
                // FP64 start vector = sum of the wanted FP32 Ritz vectors
                full();
                d.action = 2;
                d.kk = 0;
                d.coef.assign(jm, 0.0);
                for (int r = 0; r < jm; ++r)
                    for (int i = 0; i < nev; ++i) d.coef[r] += Sm[(size_t)r * jm + i];
            } else {
                d.action = 1;
            }
            return;
        }
        if (jm == m && q.cycles[q.phase] >= job->maxiter) {
            d.action = -1;
            d.rc = SED_ENOCONV;
            d.err = "Lanczos (packed) reached maxiter restarts";
            return;
        }
        if (jm == m) {
            full();
            d.action = 3;
            const int kk = std::min(m - 2, job->nev + std::max(4, (m - job->nev) / 3));
            d.kk = kk;
            d.coef.resize((size_t)jm * kk);
            for (int r = 0; r < jm; ++r)
                for (int i = 0; i < kk; ++i) d.coef[(size_t)r * kk + i] = Sm[(size_t)r * jm + i];
            // the restart's projected matrix: Ritz values and their couplings to v_m
            std::fill(q.T.begin(), q.T.end(), 0.0);
            for (int i = 0; i < kk; ++i) {
                q.T[(size_t)i * m + i] = d.theta[i];
                const double sc = beta * Sm[(size_t)(jm - 1) * jm + i];
                q.T[(size_t)kk * m + i] = sc;
                q.T[(size_t)i * m + kk] = sc;
            }
        }
    }

    // Host side of a decision; restarts and hand-overs are queued as moves for the
    // next cycle.  Returns 1 when the request left the engine.
    int apply(PackReq& q, Decision& d, std::vector<std::pair<PackMove, int>>& moves, size_t& coef_used) {
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
            case 4:
                q.phase = 1;
                q.start = kStartRandom;
                q.T.assign((size_t)q.m * q.m, 0.0);
                q.k = 0;
                q.j = 0;
                return 0;
            case 2:
            case 3: {
                PackMove mv{};
                double* dS = coef_.p + coef_used;
                std::memcpy(h_coef_ + coef_used, d.coef.data(), d.coef.size() * 8);
                coef_used += (d.coef.size() + 31) & ~(size_t)31;
                mv.S = dS;
                mv.n = q.n;
                mv.jm = d.jm;
                mv.kk = d.kk;
                mv.stab = job->sec->d_stab.p;
                const bool f32 = q.phase == 0;
                mv.V = f32 ? (void*)q.V32 : (void*)q.V64;
                mv.xhat = f32 ? (void*)q.xh32 : (void*)q.xh64;
                if (d.action == 2) {
                    mv.dst = reinterpret_cast<double2*>(q.V64);
                    q.phase = 1;
                    q.start = kStartGiven;
                    q.T.assign((size_t)q.m * q.m, 0.0);
                    q.k = 0;
                    q.j = 0;
                } else {
                    q.set_step = d.kk;
                    q.k = d.kk;
                    q.j = d.kk;
                    ++q.restarts;
                    ++q.cycles[q.phase];
                    // kept Ritz values and their couplings to v_kk for the device's
                    // re-orthogonalisation state
                    double* h = h_coef_ + coef_used;
                    for (int i = 0; i < d.kk; ++i) {
                        h[i] = d.theta[i];
                        h[d.kk + i] = q.T[(size_t)d.kk * q.m + i];
                    }
                    q.pro_init_off = (int64_t)coef_used;
                    q.pro_kk = d.kk;
                    coef_used += (2 * (size_t)d.kk + 31) & ~(size_t)31;
                }
                moves.push_back({mv, f32 ? 0 : 1});
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
            it.seed = q.job->seed;
            it.m = q.m;
            it.start = q.start;
            it.set_step = q.set_step;
            it.pro = pro_ptr(q.slot);
            it.pro_kk = q.set_step >= 0 ? q.pro_kk : 0;
            it.pro_init = it.pro_kk > 0 ? coef_.p + q.pro_init_off : nullptr;
            it.pro_enabled = pro_on_ ? 1 : 0;
            it.pro_reset = q.fresh ? 1 : 0;
            q.fresh = false;
            q.pro_kk = 0;
            const int j0 = q.start != kStartNone ? 0 : (q.set_step >= 0 ? q.set_step : q.j);
            const int first = q.job->nev + 2;
            it.target = std::min(q.m, j0 < first ? first : j0 + kPackCheck);
            q.start = kStartNone;
            q.set_step = -1;
            size_t off[9];
            lds = std::max(lds, mv_lds_layout<R>(it.c, off, kPackBlock, false, true));
        }
        PackItem* d_items = reinterpret_cast<PackItem*>(items_.p) + offset;
        SED_HIP(hipMemcpyAsync(d_items, h_items_ + offset, idx.size() * sizeof(PackItem), hipMemcpyHostToDevice, st_));
        hipLaunchKernelGGL((k_pack_lanczos<M, R, F>), dim3((unsigned)idx.size()), dim3(kPackBlock), lds, st_, d_items);
        SED_HIP(hipGetLastError());
        return SED_OK;
    }

    // queued restarts / hand-overs, batched per basis precision
    int launch_moves(std::vector<std::pair<PackMove, int>>& moves, size_t coef_used) {
        if (moves.empty()) return SED_OK;
        SED_HIP(hipMemcpyAsync(coef_.p, h_coef_, coef_used * 8, hipMemcpyHostToDevice, st_));
        int cnt = 0;
        for (int prec = 0; prec < 2; ++prec) {
            const int first = cnt;
            int64_t max_n = 0;
            int max_jm = 0;
            for (auto& mv : moves)
                if (mv.second == prec) {
                    h_moves_[cnt++] = mv.first;
                    max_n = std::max(max_n, mv.first.n);
                    max_jm = std::max(max_jm, mv.first.jm);
                }
            const int nm = cnt - first;
            if (nm == 0) continue;
            PackMove* d_mv = reinterpret_cast<PackMove*>(moves_.p) + first;
            SED_HIP(hipMemcpyAsync(d_mv, h_moves_ + first, nm * sizeof(PackMove), hipMemcpyHostToDevice, st_));
            const dim3 grid((unsigned)((max_n + kMoveRows - 1) / kMoveRows), (unsigned)nm);
            if (prec == 0) {
                const size_t lds = (size_t)max_jm * kMoveRows * sizeof(float2);
                hipLaunchKernelGGL(k_pack_move<float2>, grid, dim3(kMoveRows), lds, st_, d_mv);
            } else {
                const size_t lds = (size_t)max_jm * kMoveRows * sizeof(double2);
                hipLaunchKernelGGL(k_pack_move<double2>, grid, dim3(kMoveRows), lds, st_, d_mv);
            }
            SED_HIP(hipGetLastError());
        }
        moves.clear();
        return SED_OK;
    }

    int cycle(std::vector<PackReq>& act, std::vector<std::pair<PackMove, int>>& moves, size_t coef_used) {
        SED_TRY(launch_moves(moves, coef_used));
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
        SED_HIP(hipMemcpyAsync(h_slots_, slots_.p, (size_t)max_active_ * kSlotDoubles * 8, hipMemcpyDeviceToHost, st_));
        SED_HIP(hipStreamSynchronize(st_));
        return SED_OK;
    }

    void loop() {
        (void)hipSetDevice(ctx_->device);
        std::vector<PackReq> act;
        std::vector<std::pair<PackMove, int>> moves;
        size_t coef_used = 0;
        const int nthr_max = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        while (true) {
            std::vector<std::pair<PackReq, std::pair<int, std::string>>> failed;
            bool stopping = false;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !queue_.empty() || !act.empty(); });
                stopping = stop_;
                while (!stopping && !queue_.empty() && !free_slots_.empty()) {
                    PackJob* job = queue_.front();
                    PackReq q;
                    if (admit(job, q) == 1) {
                        if (act.empty()) {   // larger than the whole arena: can never run
                            queue_.pop_front();
                            q.job = job;
                            failed.push_back({q, {SED_ENOMEM, "request larger than the packed engine's arena"}});
                            continue;
                        }
                        break;   // wait for a retirement
                    }
                    queue_.pop_front();
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
            const int rc = cycle(act, moves, coef_used);
            coef_used = 0;
            const auto tc1 = std::chrono::steady_clock::now();
            if (rc != SED_OK) {
                const std::string err = sed_last_error();
                for (PackReq& q : act) finish(q, rc, err);
                act.clear();
                moves.clear();
                continue;
            }
            // projected-matrix work of every request in parallel (pure computation)
            std::vector<Decision> dec(act.size());
            {
                const int nthr = std::min<int>((int)act.size(), nthr_max);
                std::vector<std::thread> pool;
                for (int t = 1; t < nthr; ++t)
                    pool.emplace_back([&, t] {
                        for (size_t i = t; i < act.size(); i += nthr) decide(act[i], dec[i]);
                    });
                for (size_t i = 0; i < act.size(); i += nthr) decide(act[i], dec[i]);
                for (auto& th : pool) th.join();
            }
            const auto tc2 = std::chrono::steady_clock::now();
            std::vector<PackReq> keep;
            keep.reserve(act.size());
            for (size_t i = 0; i < act.size(); ++i)
                if (apply(act[i], dec[i], moves, coef_used) == 0) keep.push_back(std::move(act[i]));
            act.swap(keep);
            if (trace_) {
                const auto tc3 = std::chrono::steady_clock::now();
                auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
                fprintf(stderr, "[pack] active=%zu moves=%zu gpu=%.2fms decide=%.2fms apply=%.2fms\n", dec.size(),
                        moves.size(), ms(tc0, tc1), ms(tc1, tc2), ms(tc2, tc3));
            }
        }
    }

    bool trace_ = std::getenv("SED_PACK_TRACE") != nullptr;
    sed_context* ctx_;
    int max_active_;
    void (*cb_)(void*, PackJob*);
    void* user_;
    hipStream_t st_ = nullptr;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
   public:
    // device memory the engine holds from start() to its destruction
    int64_t device_bytes() const {
        return (int64_t)(arena_.capacity() + slots_.bytes() + pros_.bytes() + items_.bytes() + moves_.bytes() +
                         coef_.bytes());
    }

   private:
    std::deque<PackJob*> queue_;
    bool stop_ = false;
    Arena arena_;
    DevBuf<double> slots_;
    DevBuf<uint8_t> pros_;      // ProState per slot
    DevBuf<uint8_t> items_, moves_;
    DevBuf<double> coef_;
    double* h_slots_ = nullptr;
    PackItem* h_items_ = nullptr;
    PackMove* h_moves_ = nullptr;
    double* h_coef_ = nullptr;
    std::vector<int> free_slots_;
    bool pro_on_ = true;
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
int64_t pack_device_bytes(const PackEngine* e) { return e ? e->device_bytes() : 0; }
void pack_destroy(PackEngine* e) {
    if (!e) return;
    e->stop();
    delete e;
}

}  // namespace sed
