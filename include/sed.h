/*
 * sed.h -- C ABI of the MI355X (gfx950) matrix-free symmetry-sector solver.
 *
 * Drop-in for the matrix-free mode of GrothendieckUniverse/RealSpace_ExactDiagonalization:
 * a host shim keeps the reference's model / symmetry-group / sector-basis types and
 * forwards the operator calls here.  Every entry point below names the reference
 * interface it replaces (src/symmetry_resolved_ed.jl unless stated otherwise).
 *
 * Conventions (the reference's Julia layout):
 *   - complex numbers are two contiguous doubles (re, im), i.e. Julia ComplexF64;
 *   - site indices are 1-based, bit (i-1) of a mask is site i;
 *   - group perm[g*n_site + (i-1)] is the 1-based target of site i under element g;
 *   - table indices (combinadic ranks, orbit indices, group indices) are 1-based
 *     where the reference's are.
 * Ownership: host buffers are borrowed for the duration of the call and never
 * retained; device memory belongs to the handles and is released by *_destroy.
 * Errors: every function returns SED_OK (0) or a negative code; the message is in
 * sed_last_error() (thread-local).  No C++ exception crosses this ABI.
 * Threading: a handle is not thread-safe; distinct handles on distinct host
 * threads are independent.  Calls are synchronous with respect to host buffers.
 */
#ifndef SED_H
#define SED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SED_OK 0
#define SED_EINVAL (-1)       /* bad argument */
#define SED_ENOMEM (-2)       /* device allocation failed / over budget */
#define SED_EDEVICE (-3)      /* HIP runtime error */
#define SED_ENOCONV (-4)      /* eigensolver did not converge */
#define SED_EUNSUPPORTED (-5) /* input outside what the device path handles */
#define SED_ESTATE (-6)       /* call out of order (e.g. matvec before set_terms) */

#define SED_BOSONIC 0
#define SED_FERMIONIC 1

/* arithmetic of a matvec / solve */
#define SED_PREC_FP64 0        /* complex FP64 vectors and matvec */
#define SED_PREC_FP32 1        /* complex FP32 vectors and matvec, FP32 accumulation */
#define SED_PREC_FP32_REFINE 2 /* FP32 Lanczos, then FP64 Lanczos restarted from its Ritz vectors */

typedef struct sed_context sed_context;
typedef struct sed_catalog sed_catalog;
typedef struct sed_sector sed_sector;
typedef struct sed_server sed_server;

typedef struct sed_solve_stats {
    int32_t converged;        /* number of converged eigenpairs */
    int32_t iters_fp32;       /* Lanczos steps in FP32 */
    int32_t iters_fp64;       /* Lanczos steps in FP64 */
    int32_t restarts;         /* thick restarts */
    int64_t matvecs;          /* H|psi> applications */
    double residual_max;      /* max residual norm of the returned pairs */
    double seconds;           /* wall time of the solve */
    int32_t full_orth_steps;  /* steps orthogonalised against the whole Krylov basis */
    int32_t repeated;         /* 1 if the FP64 phase failed its true-residual check and was
                                 repeated with whole-basis passes (sed_solve, nev <= 5) */
} sed_solve_stats;

/* Thread-local message of the last failing call on this thread. */
const char* sed_last_error(void);
/* ABI version of this library (major*10000 + minor*100 + patch). */
int sed_version(void);

/* One context per GPU (device ordinal), owning a HIP stream.  No reference
 * counterpart: the reference runs on CPU threads. */
int sed_create(int device, sed_context** out);
/* Destroy calls may come in any order: a context still used by live catalogs or
 * servers, and a catalog still used by live sectors, is released together with the
 * last object that uses it. */
int sed_destroy(sed_context* ctx);

/* build_symmetry_orbit_catalog(; n_filled, symmetry_group, particle_statistics)
 * (:371-487) together with Finite_Symmetry_Group (:90-104) and
 * _inverse_operation_indices (:109-117).  Site phases of gauge form
 * eta_g(i) = f(i)/f(perm_g(i)) (build_translation_group at any twist, :1374-1405) run on
 * sign-only tables.  Other site phases are accepted when the group's operators form a
 * representation (the permutations are closed under composition and
 * eta_{g o h}(i) = eta_h(i) eta_g(perm_h(i)), e.g. a phase on a site that rotations
 * fix): the catalog then also keeps a per-rank hop amplitude (24 B per rank, catalogs of
 * at most 2^31 - 1 ranks) and H|psi> runs one launch shape; phases that are neither give
 * SED_EUNSUPPORTED.  Requires n_site <= 64,
 * C(n_site, n_filled) <= 2^32 - 2 (colex ranks are 32-bit), at most 2^31 orbits and
 * n_group <= 64.  Sectors built on it hold at most 2^31 - 1 rows. */
int sed_catalog_build(sed_context* ctx, int n_site, int n_filled, int statistics,
                      int n_group, int identity_idx, const int32_t* perm, const double* phases,
                      sed_catalog** out);
/* length(rep_rank_table) and length(representative_mask_list) of the catalog. */
int sed_catalog_info(const sed_catalog* cat, int64_t* n_total, int64_t* n_orbits);
/* Julia-layout dump of Symmetry_Orbit_Catalog (:305-327): rep_rank_table (1-based
 * orbit index), rep_sym_table (1-based inverse group index), rep_amp_table
 * (ComplexF64), representative_mask_list and stabilizer_order_list.  Any pointer
 * may be NULL.  Arrays are n_total / n_orbits long. */
int sed_catalog_copy(const sed_catalog* cat, int32_t* rep_rank, int32_t* rep_sym, double* rep_amp,
                     uint64_t* reps, int64_t* stab_order);
/* stabilizer_g_indices_list / stabilizer_phases_list (:309-310) as padded
 * n_orbits x n_group arrays (1-based indices, 0 = padding). */
int sed_catalog_stabilizers(const sed_catalog* cat, int32_t* stab_gidx, double* stab_phase);
/* update_orbit_stabilizer_phases!(catalog, new_group, particle_statistics) (:338-369):
 * replace the group's site phases (n_group x n_site ComplexF64, same permutations, e.g.
 * another flux twist) without re-enumerating orbits: the orbit tables depend on the
 * permutations only, and with phases of gauge form f(i)/f(perm_g(i)) a stabilizer
 * element's phase on its representative is its reordering sign, so sector membership
 * does not change either (with other phases the per-rank hop amplitudes are recomputed
 * and membership follows the new stabilizer phases).  Later dumps (sed_catalog_copy,
 * sed_catalog_stabilizers) and sectors built afterwards use the new phases; sectors built
 * before keep theirs.  SED_EUNSUPPORTED (catalog unchanged) when the phases are neither of
 * gauge form nor a representation, or would change the inverse elements. */
int sed_catalog_update_phases(sed_catalog* cat, const double* phases);
int sed_catalog_destroy(sed_catalog* cat);

/* build_symmetry_sector_basis(catalog, irrep; atol) (:544-556) with the irrep's
 * characters chi (n_group ComplexF64, OneDim_Irrep.values :501-504).  The catalog
 * must outlive the sector.  Sectors hold at most 2^31 - 1 rows; one whose row index does
 * not fit beside ceil(log2 n_group) + 1 group and sign bits in 4 bytes (more than 32 sites
 * only) keeps an 8-byte entry per rank, and is refused with SED_EUNSUPPORTED when the
 * catalog's site phases are not of gauge form. */
int sed_sector_build(sed_catalog* cat, const double* chi, double atol, sed_sector** out, int64_t* dim);
/* representative_mask_list and stabilizer_order_list of the sector basis (:522-528). */
int sed_sector_basis(const sed_sector* sec, uint64_t* reps, int64_t* stab_order);
/* hamiltonian_linear_operator(basis, bilinear_terms, density_terms, cmap)
 * (:891-897) / MatrixFreeHamiltonian(...) (:779-791): attaches the term lists
 * (i_from, i_to, t) and (i, j, V), 1-based sites, t and V ComplexF64. */
int sed_sector_set_terms(sed_sector* sec, int n_bilinear, const int32_t* i_from, const int32_t* i_to,
                         const double* t, int n_density, const int32_t* di, const int32_t* dj,
                         const double* v);
/* Exact hop counts of one matvec: gated hops (occupancy gate passed, one rep-table
 * read each) and in-sector hops (one x gather each); they define the algorithmic
 * bytes of the H|psi> kernel. */
int sed_sector_hop_counts(sed_sector* sec, int64_t* hops_gated, int64_t* hops_in_sector);
/* (H::MatrixFreeHamiltonian)(x) (:880-884) / _apply_table! (:731-765): y = H x for
 * host vectors of sector_dim ComplexF64; precision SED_PREC_FP64 or SED_PREC_FP32. */
int sed_matvec(sed_sector* sec, const double* x, double* y, int precision);
/* Same on device buffers of sector_dim complex elements (double2 for FP64, float2
 * for FP32) in the sector basis, with or without twisted phases; ordered on the
 * context's stream. */
int sed_matvec_device(sed_sector* sec, const void* x_dev, void* y_dev, int precision);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * sed_matvec_device with no allocation and no host synchronisation.  Its scratch (one
 * vector) is a fixed 64 MiB device array per GPU, set up by sed_create and held until
 * the process ends, so dim * 16 bytes (FP64) or dim * 8 bytes (FP32) must fit: dim <=
 * 4,194,304 (FP64) or 8,388,608 (FP32), else SED_EINVAL.  It uses the launch shape an
 * earlier matvec or solve chose for the sector (shape 1 if none) and never autotunes.
 * Ordered on the context's stream; calls that overlap on one GPU (other streams,
 * other contexts) race on the shared scratch array. */
int sed_matvec_device_noalloc(sed_sector* sec, const void* x_dev, void* y_dev, int precision);
/* Time n_calls device matvecs on resident random vectors; returns the mean
 * kernel time in ms measured with HIP events on the context's stream. */
int sed_matvec_benchmark(sed_sector* sec, int precision, int n_calls, double* ms_per_call);
/* Launch shape of the H|psi> kernel (no reference counterpart; the reference has one
 * threaded loop).  Shapes are indexed 0..n_shapes-1; each has `lanes` lanes per row
 * (64/lanes rows per wavefront), `batch` hops in flight per lane and a workgroup size
 * (128, 256 or 512 threads; a shape whose LDS does not fit is skipped).  xcd_chunk is the
 * run length in which each of the 8 XCDs takes consecutive row blocks (1 = natural
 * order, 0 = one contiguous range per XCD).  shape = -1 re-enables the autotuner
 * (xcd_chunk ignored), which on the first matvec of a sector of dimension >= 2^17 times
 * every shape, then the run lengths for the fastest one, then the shapes again at the
 * fastest run length, and keeps the fastest pair; sectors of 512 <= dimension < 2^17 get
 * the fastest shape at run length 1, each shape timed over back-to-back launches, and
 * smaller ones shape 1 at run length 1. */
int sed_sector_set_shape(sed_sector* sec, int precision, int shape, int xcd_chunk);
/* Chosen shape (-1 if not chosen yet), its lanes/batch and XCD run length, the numbers
 * of shapes and of run lengths tried, the autotuner's times in ms (may be NULL;
 * 2 * n_shapes + n_chunks floats: every shape at run length 1, every run length for the
 * fastest of those, the shapes within 20 % of the fastest again at the fastest run
 * length; 0 if not timed, -1 if the shape does not fit or was not timed again), and the
 * shape's workgroup size (may be NULL). */
int sed_sector_shape(const sed_sector* sec, int precision, int* shape, int* lanes, int* batch, int* xcd_chunk,
                     int* n_shapes, int* n_chunks, float* tune_ms, int* workgroup);
/* diagonalize_block_matrixfree(H_op, n; nev) (:998-1030): lowest nev eigenvalues
 * (ascending) and optionally eigenvectors (sector_dim x nev column-major ComplexF64,
 * may be NULL).  Sectors of dim <= 500 are solved densely, as the reference does;
 * larger ones by thick-restart Lanczos from a seeded Gaussian start vector, residual
 * tolerance tol (reference: 1e-10), at most maxiter thick restarts per precision phase
 * (the reference's KrylovKit maxiter = 300 likewise counts restarts of its Krylov basis).
 * For nev <= 5 the basis is kept semi-orthogonal by partial re-orthogonalisation and the
 * FP64 phase checks the true residuals of its pairs (one matvec each), repeating the
 * solve with whole-basis passes if one exceeds tol; larger nev orthogonalises every step
 * against the whole basis (SED_PRO=0 in the environment does so for every nev,
 * SED_PRO_MAX_NEV moves the limit).  stats->residual_max is the largest checked residual
 * (the Lanczos estimate where none was checked).  Returns SED_ENOCONV if fewer than nev
 * converge. */
int sed_solve(sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter, int precision,
              double* evals, double* evecs, sed_solve_stats* stats);
/* Site densities <n_i> of eigenvector `which` (0-based) of the last sed_solve on this
 * sector, n_site doubles (observables/static_structure_factor.jl:62-83 convention). */
int sed_sector_density(sed_sector* sec, int which, double* density);
/* Density-density correlation of eigenvector `which` of the last sed_solve:
 * nn[i * n_site + j] = <n_i n_j> (0-based sites, i == j included) and
 * density[i] = <n_i>, formed as the orbit average of
 * _build_density_correlation_matrix! (observables/static_structure_factor.jl:41-104)
 * before its disconnected part is subtracted. */
int sed_sector_density_correlation(sed_sector* sec, int which, double* nn, double* density);
/* One-body density matrix rho[i * n_site + j] = <a_i^dag a_j> of eigenvector `which`
 * of the last sed_solve (complex interleaved, 0-based sites), the quantity of
 * off_diagonal_long_range_order (observables/off_diagonal_long_range_order.jl:55-126). */
int sed_sector_one_body(sed_sector* sec, int which, double* rho);
/* Matrix mode: the sector block of the Hamiltonian set by sed_sector_set_terms as the
 * reference's SparseMatrixCSC{ComplexF64,Int} (build_ed_Hamiltonian_symmetry_block,
 * src/symmetry_resolved_ed.jl:903-963): 1-based colptr (dim+1) and rowval (nnz), nzval
 * (nnz ComplexF64), duplicates summed, explicit zeros dropped.  Call once with colptr,
 * rowval and nzval NULL to get *nnz (the matrix is built and kept), then again with
 * arrays of that size to receive it (the kept copy is released). */
int sed_sector_matrix_csc(sed_sector* sec, int64_t* nnz, int64_t* colptr, int64_t* rowval, double* nzval);
/* Fock amplitudes of a sector state c (dim ComplexF64, reference basis), as
 * _expand_sector_state_to_fock_amplitudes (observables/entanglement_spectrum.jl:46-79):
 * representatives with |c| <= atol skipped, amplitudes with |a| <= atol dropped, the
 * rest normalised.  Writes the n_out nonzero (mask, amplitude) pairs in ascending mask
 * order when capacity >= n_out; capacity 0 only reports n_out.  SED_EINVAL when
 * 0 < capacity < n_out or the expanded norm is <= atol (the reference errors). */
int sed_sector_fock_expansion(sed_sector* sec, const double* c, double atol, int64_t capacity, int64_t* n_out,
                              uint64_t* masks, double* amps);
/* Projected position operator <chi_to| U_d |chi_from> between two sectors of one
 * catalog (_position_operator_matrix, observables/charge_pump.jl:47-91) with site
 * phases phi (n_site ComplexF64): column col of the from-sector has its one possible
 * entry in rows[col] (1-based row of the to-sector, 0 when the orbit is not there)
 * with value vals[col] (ComplexF64). */
int sed_sector_position_operator(sed_sector* to, sed_sector* from, const double* site_phases, int64_t* rows,
                                 double* vals);
int sed_sector_destroy(sed_sector* sec);

/* Continuous-batching service of many sector solves (no reference counterpart; it
 * replaces the per-sector loop of ed_scan! / _ed_scan_sectors! :1199-1356 for the
 * flux scan).  Requests are admitted while their HBM footprint fits the budget
 * (hbm_budget_bytes <= 0: 80% of free memory) and one of max_slots solver lanes is
 * free; each lane has its own HIP stream, so admitted requests overlap on the
 * device, and a finished request frees its lane for the queue at once.  The budget
 * covers all device memory the server allocates: the packed engine's arena (a quarter
 * of the budget, at most 16 GiB, taken at creation), the workspaces lanes keep between
 * requests (released when a queued request needs the room) and the running requests'
 * footprints; catalogs and sectors belong to the caller and are not counted.  A request
 * whose footprint exceeds what the arena leaves of the budget fails with SED_ENOMEM.
 * Two requests on the same sector object of one server never run at the same time. */
int sed_server_create(sed_context* ctx, int64_t hbm_budget_bytes, int max_slots, sed_server** out);
/* Queue one solve of `sec` (which must stay alive until its result is taken). */
int sed_server_submit(sed_server* srv, sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter,
                      int precision, int64_t* request_id);
/* Drive the service from the caller's thread: admit queued requests that fit, then
 * wait: max_steps < 0 returns at once, max_steps == 0 waits until every request
 * submitted so far has finished, max_steps > 0 waits until that many more have
 * finished.  Admitted requests keep running (and queued ones keep being admitted)
 * between calls.  n_done receives the number of requests finished so far. */
int sed_server_run(sed_server* srv, int max_steps, int64_t* n_done);
/* Result of a request: status 0 = done, 1 = pending, 2 = running, <0 = failed (error
 * code); evals (nev doubles) and stats are filled when done.  A result can be read any
 * number of times: the server keeps every request's record (its eigenvalues, stats and
 * timestamps, about 200 bytes + 8 per eigenvalue) until it is destroyed, so a
 * long-running service recreates its server from time to time. */
int sed_server_result(sed_server* srv, int64_t request_id, int* status, double* evals, sed_solve_stats* stats,
                      double* latency_s);
/* Memory accounting of a server (no reference counterpart): its HBM budget, the bytes
 * taken at creation (packed engine), the bytes charged now (creation + lane workspaces +
 * running footprints) and the largest such charge so far; any pointer may be NULL. */
int sed_server_memory(sed_server* srv, int64_t* budget_bytes, int64_t* fixed_bytes, int64_t* in_use_bytes,
                      int64_t* peak_bytes);
int sed_server_destroy(sed_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* SED_H */
