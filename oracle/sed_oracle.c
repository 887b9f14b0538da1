/*
 * sed_oracle.c -- CPU restatement of the reference's matrix-free symmetry-sector
 * path (GrothendieckUniverse/RealSpace_ExactDiagonalization, Julia).
 *
 * TEST INFRASTRUCTURE.  This file is the parity checker and the CPU baseline:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product path (libsed.so) never links, calls or falls back to it.
 *
 * It follows the reference function by function, in the reference's own order of
 * operations, so that its tables are the reference's tables:
 *   apply_operation_to_mask           src/symmetry_resolved_ed.jl:132-163
 *   get_canonical_representative      :174-187
 *   _gosper_next / first mask         :193-202
 *   rank_combination                  :231-243
 *   _lin_table / lin_index            :262-299
 *   _inverse_operation_indices        :109-117
 *   build_symmetry_orbit_catalog      :371-487 (last writer wins on the tables)
 *   _is_orbit_compatible / sector     :534-556
 *   get_canonical / project_to_sector :584-594, :627-640
 *   hopping_phase_for_stats           :646-654
 *   build_matrixfree_projection_table :680-728 (also over a sample of the columns,
 *                                      for the host baseline's bounded timing)
 *   _apply_table!                     :731-765 (per-thread buffers kept between calls,
 *                                      zeroed per call, static split; optionally timed by part)
 *   apply_hamiltonian!                :813-878 (on-the-fly, threaded over columns)
 *   update_orbit_stabilizer_phases!   :338-369
 *   _build_density_correlation_matrix! observables/static_structure_factor.jl:41-104
 *   off_diagonal_long_range_order     observables/off_diagonal_long_range_order.jl:74-126
 *   _expand_sector_state_to_fock_amplitudes observables/entanglement_spectrum.jl:46-79
 *   _position_operator_matrix         observables/charge_pump.jl:47-91
 * and two checkers for sectors too large for its catalog: a sector basis against its
 * definition (oc_check_sector_basis) and sampled rows of H x without any table
 * (oc_sample_rows, the projection-table formula of :718 evaluated per row).
 *
 * How each is pinned (tests/test_oracle.py, tests/test_checkpoint.py, tests/test_topology.py):
 *   ranks, Gosper order, LinTable      the design notebook's cells 22/24; Gosper's sequence
 *                                      visits ranks 0, 1, 2, ... over whole C(n, k) spaces
 *   catalog tables, sector bases       the benchmark CSV's orbit counts and sector dims,
 *                                      notebook cell 26, the character formula for every
 *                                      momentum, the sectors partitioning C(N, k)
 *   matrix elements, both matvecs      the CSV energies (36 rows), notebook cell 43 (nnz 756),
 *                                      free-fermion closed forms (XX chain, Hubbard U=0, twisted),
 *                                      brute-force Fock-space spectra (union of sectors =
 *                                      full spectrum, also for site phases not of gauge form),
 *                                      twist periodicity
 *   observables, Fock expansion,       literal operators in the full Fock space applied to the
 *   position operator                  embedded sector state; the charge-pump notebook's
 *                                      stored charges, gaps and Chern number
 *   phase refresh                      equal to a catalog rebuilt at the new phases
 *   sampled rows, basis checker        equal to the projection table on small sectors; planted
 *                                      defects in a basis are each reported
 *   table of sampled columns           exactly those columns' entries of the whole table, in
 *                                      order; _apply_table! over it equals the on-the-fly
 *                                      matvec on those columns
 *
 * Complex products are evaluated as the reference writes them, left to right;
 * build with -ffp-contract=off so no FMA contraction changes the rounding.
 * The sector-basis Dict{Mask,Int} is replaced by a binary search over the
 * ascending representative list; both return the same row.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t mask_t;
typedef double _Complex cplx;

/* ------------------------------------------------------------------ */
/* binomials and colex rank                                            */
/* ------------------------------------------------------------------ */
static int64_t binom_tab[65][65];
static int binom_ready = 0;

static void binom_init(void) {
    if (binom_ready) return;
    for (int n = 0; n <= 64; ++n) {
        binom_tab[n][0] = 1;
        for (int k = 1; k <= 64; ++k) {
            if (k > n) binom_tab[n][k] = 0;
            else binom_tab[n][k] = binom_tab[n - 1][k - 1] + (k <= n - 1 ? binom_tab[n - 1][k] : 0);
        }
    }
    binom_ready = 1;
}

static inline int64_t binom(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    return binom_tab[n][k];
}

int64_t oc_binomial(int n, int k) { binom_init(); return binom(n, k); }

/* rank_combination (:231-243): sum_j C(b_j, j+1) over set bits b_0<b_1<... */
int64_t oc_rank_combination(mask_t m) {
    binom_init();
    int64_t idx = 0;
    int j = 1;
    mask_t tmp = m;
    while (tmp != 0) {
        mask_t lsb = tmp & (~tmp + 1);
        int b = __builtin_ctzll(lsb);
        idx += binom(b, j);
        j += 1;
        tmp ^= lsb;
    }
    return idx;
}

/* _gosper_next (:193-197) */
mask_t oc_gosper_next(mask_t x) {
    mask_t c = x & (~x + 1);
    mask_t r = x + c;
    return (((r ^ x) >> 2) / c) | r;
}

/* _lin_table (:262-292); left/right are caller arrays of 2^n_left, 2^n_right */
int oc_lin_table(int n_site, int n_filled, int64_t* left, int64_t* right) {
    binom_init();
    int n_right = n_site / 2;
    int n_left = n_site - n_right;
    int64_t size_right = (int64_t)1 << n_right;
    int64_t size_left = (int64_t)1 << n_left;
    for (int64_t v = 0; v < size_right; ++v) right[v] = oc_rank_combination((mask_t)v);
    int64_t acc = 0;
    for (int64_t u = 0; u < size_left; ++u) {
        left[u] = acc;
        int pc = __builtin_popcountll((mask_t)u);
        int rem = n_filled - pc;
        if (0 <= rem && rem <= n_right) acc += binom(n_right, rem);
    }
    return n_right;
}

/* lin_index (:295-299) */
int64_t oc_lin_index(mask_t m, const int64_t* left, const int64_t* right, int n_right) {
    mask_t hi = m >> n_right;
    mask_t lo = m & (((mask_t)1 << n_right) - 1);
    return left[hi] + right[lo];
}

/* ------------------------------------------------------------------ */
/* group action                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int n_site, n_group, fermionic;
    int32_t* perm;     /* n_group x n_site, 1-based */
    cplx* phases;      /* n_group x n_site */
    int32_t* inv;      /* n_group, 1-based */
} oc_group;

/* apply_operation_to_mask, Bosonic (:132-144) and Fermionic (:146-163) */
static inline mask_t apply_op(mask_t m, const oc_group* G, int g0, cplx* alpha) {
    const int32_t* perm = G->perm + (size_t)g0 * G->n_site;
    const cplx* ph = G->phases + (size_t)g0 * G->n_site;
    mask_t tmp = m, nm = 0;
    cplx phase = 1.0;
    int parity = 1;
    while (tmp != 0) {
        mask_t lsb = tmp & (~tmp + 1);
        int idx = __builtin_ctzll(lsb) + 1;
        int p = perm[idx - 1];
        if (G->fermionic) {
            mask_t above = p >= 64 ? 0 : (nm >> p);
            if (__builtin_popcountll(above) & 1) parity = -parity;
        }
        nm |= (mask_t)1 << (p - 1);
        phase = phase * ph[idx - 1];
        tmp ^= lsb;
    }
    *alpha = G->fermionic ? phase * (double)parity : phase;
    return nm;
}

/* _inverse_operation_indices (:109-117): first h with h o g = id and matching phases */
int oc_inverse_indices(int n_site, int n_group, const int32_t* perm, const double* phases, int32_t* out) {
    const cplx* ph = (const cplx*)phases;
    for (int g = 0; g < n_group; ++g) {
        int found = -1;
        for (int h = 0; h < n_group && found < 0; ++h) {
            int ok = 1;
            for (int i = 0; i < n_site && ok; ++i) {
                int gi = perm[(size_t)g * n_site + i];                 /* 1-based */
                if (perm[(size_t)h * n_site + gi - 1] != i + 1) ok = 0;
            }
            for (int i = 0; i < n_site && ok; ++i) {
                int gi = perm[(size_t)g * n_site + i];
                cplx prod = ph[(size_t)g * n_site + i] * ph[(size_t)h * n_site + gi - 1];
                if (cabs(prod - 1.0) > 1e-12) ok = 0;
            }
            if (ok) found = h;
        }
        if (found < 0) return -1;
        out[g] = found + 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* orbit catalog (:371-487)                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    oc_group G;
    int n_filled;
    int64_t n_total, n_orbits, cap;
    mask_t* reps;
    int64_t* stab_order;
    int32_t* stab_gidx;   /* n_orbits x n_group, 1-based, 0 padded */
    cplx* stab_phase;     /* n_orbits x n_group */
    int32_t* rep_rank;    /* n_total, 1-based orbit index, 0 = unset */
    int32_t* rep_sym;     /* n_total, 1-based inverse op index */
    cplx* rep_amp;        /* n_total */
    int borrowed;         /* tables/reps owned by the caller (oc_catalog_wrap) */
} oc_catalog;

static void cat_push(oc_catalog* c, mask_t m, int64_t so, const int32_t* gi, const cplx* gp) {
    if (c->n_orbits == c->cap) {
        int64_t nc = c->cap ? 2 * c->cap : 1024;
        c->reps = realloc(c->reps, nc * sizeof(mask_t));
        c->stab_order = realloc(c->stab_order, nc * sizeof(int64_t));
        c->stab_gidx = realloc(c->stab_gidx, nc * (size_t)c->G.n_group * sizeof(int32_t));
        c->stab_phase = realloc(c->stab_phase, nc * (size_t)c->G.n_group * sizeof(cplx));
        c->cap = nc;
    }
    int64_t o = c->n_orbits++;
    c->reps[o] = m;
    c->stab_order[o] = so;
    for (int j = 0; j < c->G.n_group; ++j) {
        c->stab_gidx[o * c->G.n_group + j] = j < so ? gi[j] : 0;
        c->stab_phase[o * c->G.n_group + j] = j < so ? gp[j] : 0.0;
    }
}

void oc_catalog_free(oc_catalog* c) {
    if (!c) return;
    free(c->G.perm); free(c->G.phases); free(c->G.inv);
    if (!c->borrowed) {
        free(c->reps); free(c->stab_order); free(c->stab_gidx); free(c->stab_phase);
        free(c->rep_rank); free(c->rep_sym); free(c->rep_amp);
    }
    free(c);
}

/* A catalog over tables produced elsewhere (borrowed, not copied): lets the CPU
 * baseline time the reference's matvec on sizes whose serial catalog build would
 * take minutes.  Only the lookup tables and the representative list are needed by
 * get_canonical / project_to_sector. */
oc_catalog* oc_catalog_wrap(int n_site, int n_filled, int n_group, const int32_t* perm, const double* phases,
                            int fermionic, int64_t n_orbits, uint64_t* reps, int32_t* rep_rank, int32_t* rep_sym,
                            double* rep_amp) {
    binom_init();
    oc_catalog* c = calloc(1, sizeof(oc_catalog));
    c->G.n_site = n_site; c->G.n_group = n_group; c->G.fermionic = fermionic;
    c->G.perm = malloc((size_t)n_group * n_site * sizeof(int32_t));
    c->G.phases = malloc((size_t)n_group * n_site * sizeof(cplx));
    c->G.inv = malloc((size_t)n_group * sizeof(int32_t));
    memcpy(c->G.perm, perm, (size_t)n_group * n_site * sizeof(int32_t));
    memcpy(c->G.phases, phases, (size_t)n_group * n_site * sizeof(cplx));
    if (oc_inverse_indices(n_site, n_group, perm, phases, c->G.inv) != 0) { oc_catalog_free(c); return NULL; }
    c->n_filled = n_filled;
    c->n_total = binom(n_site, n_filled);
    c->n_orbits = n_orbits;
    c->reps = reps;
    c->rep_rank = rep_rank;
    c->rep_sym = rep_sym;
    c->rep_amp = (cplx*)rep_amp;
    c->borrowed = 1;
    return c;
}

oc_catalog* oc_catalog_build(int n_site, int n_filled, int n_group, const int32_t* perm,
                             const double* phases, int fermionic) {
    binom_init();
    if (n_site < 1 || n_site > 63 || n_filled < 0 || n_filled > n_site || n_group < 1) return NULL;
    oc_catalog* c = calloc(1, sizeof(oc_catalog));
    c->G.n_site = n_site; c->G.n_group = n_group; c->G.fermionic = fermionic;
    c->G.perm = malloc((size_t)n_group * n_site * sizeof(int32_t));
    c->G.phases = malloc((size_t)n_group * n_site * sizeof(cplx));
    c->G.inv = malloc((size_t)n_group * sizeof(int32_t));
    memcpy(c->G.perm, perm, (size_t)n_group * n_site * sizeof(int32_t));
    memcpy(c->G.phases, phases, (size_t)n_group * n_site * sizeof(cplx));
    if (oc_inverse_indices(n_site, n_group, perm, phases, c->G.inv) != 0) { oc_catalog_free(c); return NULL; }
    c->n_filled = n_filled;
    c->n_total = binom(n_site, n_filled);
    c->rep_rank = calloc(c->n_total, sizeof(int32_t));
    c->rep_sym = calloc(c->n_total, sizeof(int32_t));
    c->rep_amp = calloc(c->n_total, sizeof(cplx));
    const oc_group* G = &c->G;
    int32_t* gi = malloc(n_group * sizeof(int32_t));
    cplx* gp = malloc(n_group * sizeof(cplx));

    if (n_filled == 0 || n_filled == n_site) {
        mask_t m = n_filled == 0 ? 0 : (((mask_t)1 << n_filled) - 1);
        int64_t so = 0;
        for (int g = 0; g < n_group; ++g) {
            cplx a; mask_t s = apply_op(m, G, g, &a);
            if (s != m) continue;
            gi[so] = g + 1; gp[so] = a; ++so;
        }
        cat_push(c, m, so, gi, gp);
        for (int g = 0; g < n_group; ++g) {
            cplx a; mask_t s = apply_op(m, G, g, &a);
            int64_t idx = oc_rank_combination(s);
            c->rep_rank[idx] = 1;
            c->rep_sym[idx] = G->inv[g];
            c->rep_amp[idx] = conj(a);
        }
    } else {
        mask_t x = ((mask_t)1 << n_filled) - 1;
        for (int64_t it = 0; it < c->n_total; ++it, x = oc_gosper_next(x)) {
            mask_t m = x;
            int is_rep = 1;
            for (int g = 0; g < n_group; ++g) {
                cplx a; mask_t s = apply_op(m, G, g, &a);
                if (s < m) { is_rep = 0; break; }
            }
            if (!is_rep) continue;
            int32_t rep_idx = (int32_t)(c->n_orbits + 1);
            int64_t so = 0;
            for (int g = 0; g < n_group; ++g) {
                cplx a; mask_t s = apply_op(m, G, g, &a);
                int64_t idx = oc_rank_combination(s);
                c->rep_rank[idx] = rep_idx;
                c->rep_sym[idx] = G->inv[g];
                c->rep_amp[idx] = conj(a);
                if (s == m) { gi[so] = g + 1; gp[so] = a; ++so; }
            }
            cat_push(c, m, so, gi, gp);
        }
    }
    free(gi); free(gp);
    return c;
}

int64_t oc_catalog_n_orbits(const oc_catalog* c) { return c->n_orbits; }
int64_t oc_catalog_n_total(const oc_catalog* c) { return c->n_total; }

void oc_catalog_tables(const oc_catalog* c, int32_t* rep_rank, int32_t* rep_sym, double* rep_amp) {
    memcpy(rep_rank, c->rep_rank, c->n_total * sizeof(int32_t));
    memcpy(rep_sym, c->rep_sym, c->n_total * sizeof(int32_t));
    memcpy(rep_amp, c->rep_amp, c->n_total * sizeof(cplx));
}

void oc_catalog_orbits(const oc_catalog* c, uint64_t* reps, int64_t* stab_order,
                       int32_t* stab_gidx, double* stab_phase) {
    memcpy(reps, c->reps, c->n_orbits * sizeof(mask_t));
    memcpy(stab_order, c->stab_order, c->n_orbits * sizeof(int64_t));
    if (stab_gidx) memcpy(stab_gidx, c->stab_gidx, c->n_orbits * (size_t)c->G.n_group * sizeof(int32_t));
    if (stab_phase) memcpy(stab_phase, c->stab_phase, c->n_orbits * (size_t)c->G.n_group * sizeof(cplx));
}

/* get_canonical_representative (:174-187) */
uint64_t oc_canonical_representative(const oc_catalog* c, uint64_t m, int32_t* g_out, double* amp_out) {
    mask_t repr = m;
    int best_g = 1;
    cplx best = 1.0;
    for (int g = 0; g < c->G.n_group; ++g) {
        cplx a; mask_t s = apply_op(m, &c->G, g, &a);
        if (s < repr) { repr = s; best_g = g + 1; best = a; }
    }
    *g_out = best_g;
    amp_out[0] = creal(best); amp_out[1] = cimag(best);
    return repr;
}

/* apply_operation_to_mask exposed for the tests */
uint64_t oc_apply_operation(const oc_catalog* c, uint64_t m, int g1, double* amp_out) {
    cplx a; mask_t s = apply_op(m, &c->G, g1 - 1, &a);
    amp_out[0] = creal(a); amp_out[1] = cimag(a);
    return s;
}

/* ------------------------------------------------------------------ */
/* sector basis (:534-556) and projection (:584-640)                   */
/* ------------------------------------------------------------------ */
typedef struct {
    const oc_catalog* cat;
    cplx* chi;            /* n_group */
    int64_t dim;
    mask_t* reps;         /* ascending */
    int64_t* stab;
    /* projection table */
    int64_t nnz;
    int64_t* row;         /* 1-based */
    int64_t* col;         /* 1-based */
    cplx* vals;
    cplx* hdiag;
    /* _apply_table!'s per-thread buffers (MatrixFreeHamiltonian's y_threads), kept
     * between calls as the reference keeps them */
    cplx* ybuf;
    int ybuf_threads;
} oc_sector;

static int is_orbit_compatible(const oc_catalog* c, int64_t o, const cplx* chi, double atol) {
    for (int j = 0; j < c->stab_order[o]; ++j) {
        int g = c->stab_gidx[o * c->G.n_group + j];
        cplx ph = c->stab_phase[o * c->G.n_group + j];
        if (!(cabs(chi[g - 1] - ph) <= atol)) return 0;
    }
    return 1;
}

oc_sector* oc_sector_build(const oc_catalog* c, const double* chi, double atol) {
    oc_sector* s = calloc(1, sizeof(oc_sector));
    s->cat = c;
    s->chi = malloc(c->G.n_group * sizeof(cplx));
    memcpy(s->chi, chi, c->G.n_group * sizeof(cplx));
    s->reps = malloc((c->n_orbits + 1) * sizeof(mask_t));
    s->stab = malloc((c->n_orbits + 1) * sizeof(int64_t));
    for (int64_t o = 0; o < c->n_orbits; ++o) {
        if (is_orbit_compatible(c, o, s->chi, atol)) {
            s->reps[s->dim] = c->reps[o];
            s->stab[s->dim] = c->stab_order[o];
            s->dim++;
        }
    }
    return s;
}

int64_t oc_sector_dim(const oc_sector* s) { return s->dim; }

/* A sector over a given basis (representatives ascending, stabilizer orders). */
oc_sector* oc_sector_from_basis(const oc_catalog* c, const double* chi, int64_t dim, const uint64_t* reps,
                                const int64_t* stab) {
    oc_sector* s = calloc(1, sizeof(oc_sector));
    s->cat = c;
    s->chi = malloc(c->G.n_group * sizeof(cplx));
    memcpy(s->chi, chi, c->G.n_group * sizeof(cplx));
    s->dim = dim;
    s->reps = malloc((dim + 1) * sizeof(mask_t));
    s->stab = malloc((dim + 1) * sizeof(int64_t));
    memcpy(s->reps, reps, dim * sizeof(mask_t));
    memcpy(s->stab, stab, dim * sizeof(int64_t));
    return s;
}

void oc_sector_basis(const oc_sector* s, uint64_t* reps, int64_t* stab) {
    memcpy(reps, s->reps, s->dim * sizeof(mask_t));
    memcpy(stab, s->stab, s->dim * sizeof(int64_t));
}

void oc_sector_free(oc_sector* s) {
    if (!s) return;
    free(s->chi); free(s->reps); free(s->stab);
    free(s->row); free(s->col); free(s->vals); free(s->hdiag);
    free(s->ybuf);
    free(s);
}

/* Dict{Mask,Int} probe (:530-532): 1-based row or 0 */
static int64_t basis_index(const oc_sector* s, mask_t r) {
    int64_t lo = 0, hi = s->dim - 1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (s->reps[mid] == r) return mid + 1;
        if (s->reps[mid] < r) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* get_canonical (:584-594) */
static mask_t get_canonical(const oc_catalog* c, mask_t m, int* g_idx, cplx* alpha) {
    int64_t idx = oc_rank_combination(m);
    int32_t rep_idx = (idx >= 0 && idx < c->n_total) ? c->rep_rank[idx] : 0;
    if (rep_idx == 0) {
        /* fallback for masks outside the fixed-filling basis */
        mask_t repr = m; int best_g = 1; cplx best = 1.0;
        for (int g = 0; g < c->G.n_group; ++g) {
            cplx a; mask_t t = apply_op(m, &c->G, g, &a);
            if (t < repr) { repr = t; best_g = g + 1; best = a; }
        }
        *g_idx = best_g; *alpha = best;
        return repr;
    }
    *g_idx = c->rep_sym[idx];
    *alpha = c->rep_amp[idx];
    return c->reps[rep_idx - 1];
}

/* project_to_sector (:627-632): returns 1-based row or 0 */
static int64_t project(const oc_sector* s, mask_t m, cplx* coeff) {
    int g; cplx a;
    mask_t repr = get_canonical(s->cat, m, &g, &a);
    int64_t idx = basis_index(s, repr);
    if (idx == 0) return 0;
    *coeff = a * conj(s->chi[g - 1]);
    return idx;
}

int64_t oc_project_to_sector(const oc_sector* s, uint64_t m, double* coeff) {
    cplx cf = 0;
    int64_t r = project(s, m, &cf);
    coeff[0] = creal(cf); coeff[1] = cimag(cf);
    return r;
}

/* project_to_unnormalized_sector (:635-640) */
int64_t oc_project_to_unnormalized_sector(const oc_sector* s, uint64_t m, double* coeff) {
    int32_t g; double a2[2];
    mask_t repr = oc_canonical_representative(s->cat, m, &g, a2);
    int64_t idx = basis_index(s, repr);
    if (idx == 0) return 0;
    cplx cf = (a2[0] + I * a2[1]) * conj(s->chi[g - 1]);
    coeff[0] = creal(cf); coeff[1] = cimag(cf);
    return idx;
}

/* hopping_phase_for_stats (:646-654) */
static inline double hop_phase(int fermionic, mask_t m, int i_from, int i_to) {
    if (!fermionic || i_from == i_to) return 1.0;
    int lo = i_from < i_to ? i_from : i_to;
    int hi = i_from < i_to ? i_to : i_from;
    mask_t between = hi - lo <= 1 ? 0 : ((((mask_t)1 << (hi - lo - 1)) - 1) << lo);
    return (__builtin_popcountll(m & between) & 1) ? -1.0 : 1.0;
}

/* build_matrixfree_projection_table (:680-728) over the columns c of [c0, c1) with
 * (c - c0) mod period < block, in ascending order (the reference builds all of them:
 * c0 = 0, c1 = dim, period = block = 1; runs of columns spread over the sector let a host
 * baseline time the serial build and _apply_table! on a bounded sample of a sector whose
 * whole table would take minutes); h_diag is filled for those columns.  Returns nnz. */
int64_t oc_table_build_strided(oc_sector* s, int n_bil, const int32_t* ifrom, const int32_t* ito,
                               const double* t_, int n_den, const int32_t* di, const int32_t* dj,
                               const double* V_, int64_t c0, int64_t c1, int64_t period, int64_t block) {
    const cplx* t = (const cplx*)t_;
    const cplx* V = (const cplx*)V_;
    const oc_catalog* c = s->cat;
    free(s->row); free(s->col); free(s->vals); free(s->hdiag);
    int64_t cap = 1024, nnz = 0;
    s->row = malloc(cap * sizeof(int64_t));
    s->col = malloc(cap * sizeof(int64_t));
    s->vals = malloc(cap * sizeof(cplx));
    s->hdiag = calloc(s->dim > 0 ? s->dim : 1, sizeof(cplx));
    mask_t* hop = malloc((n_bil + 1) * sizeof(mask_t));
    mask_t* from = malloc((n_bil + 1) * sizeof(mask_t));
    for (int k = 0; k < n_bil; ++k) {
        hop[k] = ((mask_t)1 << (ifrom[k] - 1)) | ((mask_t)1 << (ito[k] - 1));
        from[k] = (mask_t)1 << (ifrom[k] - 1);
    }
    if (c0 < 0) c0 = 0;
    if (c1 > s->dim) c1 = s->dim;
    if (period < 1) period = 1;
    if (block < 1) block = 1;
    for (int64_t col = c0; col < c1; ++col) {
        if ((col - c0) % period >= block) {
            col += period - (col - c0) % period - 1;   /* to the next run */
            continue;
        }
        mask_t m = s->reps[col];
        int64_t stab_col = s->stab[col];
        cplx hd = 0.0;
        for (int q = 0; q < n_den; ++q) {
            if (((m >> (di[q] - 1)) & 1) && ((m >> (dj[q] - 1)) & 1)) hd = hd + V[q];
        }
        s->hdiag[col] = hd;
        for (int k = 0; k < n_bil; ++k) {
            if ((m & hop[k]) != from[k]) continue;
            mask_t nm = m ^ hop[k];
            cplx coeff;
            int64_t row = project(s, nm, &coeff);
            if (row == 0) continue;
            int64_t stab_row = s->stab[row - 1];
            double hp = hop_phase(c->G.fermionic, m, ifrom[k], ito[k]);
            double nrm = sqrt((double)stab_row / (double)stab_col);
            cplx h = ((t[k] * hp) * coeff) * nrm;
            if (nnz == cap) {
                cap *= 2;
                s->row = realloc(s->row, cap * sizeof(int64_t));
                s->col = realloc(s->col, cap * sizeof(int64_t));
                s->vals = realloc(s->vals, cap * sizeof(cplx));
            }
            s->row[nnz] = row; s->col[nnz] = col + 1; s->vals[nnz] = h; ++nnz;
        }
    }
    free(hop); free(from);
    s->nnz = nnz;
    return nnz;
}

int64_t oc_table_build_cols(oc_sector* s, int n_bil, const int32_t* ifrom, const int32_t* ito, const double* t_,
                            int n_den, const int32_t* di, const int32_t* dj, const double* V_, int64_t c0,
                            int64_t c1) {
    return oc_table_build_strided(s, n_bil, ifrom, ito, t_, n_den, di, dj, V_, c0, c1, 1, 1);
}

/* build_matrixfree_projection_table (:680-728); returns nnz */
int64_t oc_table_build(oc_sector* s, int n_bil, const int32_t* ifrom, const int32_t* ito, const double* t_,
                       int n_den, const int32_t* di, const int32_t* dj, const double* V_) {
    return oc_table_build_strided(s, n_bil, ifrom, ito, t_, n_den, di, dj, V_, 0, s->dim, 1, 1);
}

void oc_table_copy(const oc_sector* s, int64_t* row, int64_t* col, double* vals, double* hdiag) {
    if (row) memcpy(row, s->row, s->nnz * sizeof(int64_t));
    if (col) memcpy(col, s->col, s->nnz * sizeof(int64_t));
    if (vals) memcpy(vals, s->vals, s->nnz * sizeof(cplx));
    if (hdiag) memcpy(hdiag, s->hdiag, s->dim * sizeof(cplx));
}

static double wall_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* _apply_table! (:731-765): per-thread dense buffers, serial diagonal, static edge
 * split over nthreads, serial reduction over the buffers.  With phases != NULL the wall
 * time of its four parts is stored there: buffer allocation and zeroing, diagonal,
 * entries, reduction (so a host baseline can scale the entry part of a partial table). */
void oc_apply_table_timed(oc_sector* s, const double* x_, double* y_, int nthreads, double* phases) {
    const cplx* x = (const cplx*)x_;
    cplx* y = (cplx*)y_;
    int64_t n = s->dim;
    if (nthreads < 1) nthreads = 1;
    if (!s->ybuf || s->ybuf_threads != nthreads) {
        free(s->ybuf);
        s->ybuf = calloc((size_t)nthreads * (n > 0 ? n : 1), sizeof(cplx));
        s->ybuf_threads = nthreads;
    }
    cplx* buf = s->ybuf;
    double t0 = wall_s();
    /* fill!(y_threads[tid], 0) for every buffer, one after another (:739-741) */
    memset(buf, 0, (size_t)nthreads * (n > 0 ? n : 1) * sizeof(cplx));
    double t1 = wall_s();
    for (int64_t col = 0; col < n; ++col) {
        cplx xc = x[col];
        if (xc == 0) continue;
        buf[col] += s->hdiag[col] * xc;
    }
    double t2 = wall_s();
    int64_t nnz = s->nnz;
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num(); nt = omp_get_num_threads();
#endif
        int64_t chunk = (nnz + nt - 1) / nt;
        int64_t e0 = (int64_t)tid * chunk, e1 = e0 + chunk < nnz ? e0 + chunk : nnz;
        cplx* yt = buf + (size_t)tid * n;
        for (int64_t e = e0; e < e1; ++e) {
            cplx xc = x[s->col[e] - 1];
            if (xc == 0) continue;
            yt[s->row[e] - 1] += s->vals[e] * xc;
        }
    }
    double t3 = wall_s();
    for (int64_t i = 0; i < n; ++i) y[i] = 0;
    for (int t = 0; t < nthreads; ++t) {
        const cplx* yt = buf + (size_t)t * n;
        for (int64_t i = 0; i < n; ++i) y[i] += yt[i];
    }
    if (phases) {
        phases[0] = t1 - t0;
        phases[1] = t2 - t1;
        phases[2] = t3 - t2;
        phases[3] = wall_s() - t3;
    }
}

void oc_apply_table(oc_sector* s, const double* x_, double* y_, int nthreads) {
    oc_apply_table_timed(s, x_, y_, nthreads, NULL);
}

/* apply_hamiltonian! (:813-878): on-the-fly, threaded over columns [c0, c1) */
void oc_apply_hamiltonian(const oc_sector* s, int n_bil, const int32_t* ifrom, const int32_t* ito,
                          const double* t_, int n_den, const int32_t* di, const int32_t* dj,
                          const double* V_, const double* x_, double* y_, int nthreads,
                          int64_t c0, int64_t c1) {
    const cplx* t = (const cplx*)t_;
    const cplx* V = (const cplx*)V_;
    const cplx* x = (const cplx*)x_;
    cplx* y = (cplx*)y_;
    const oc_catalog* c = s->cat;
    int64_t n = s->dim;
    if (nthreads < 1) nthreads = 1;
    if (c1 > n) c1 = n;
    if (c0 < 0) c0 = 0;
    cplx* buf = calloc((size_t)nthreads * (n > 0 ? n : 1), sizeof(cplx));
    mask_t* hop = malloc((n_bil + 1) * sizeof(mask_t));
    mask_t* from = malloc((n_bil + 1) * sizeof(mask_t));
    for (int k = 0; k < n_bil; ++k) {
        hop[k] = ((mask_t)1 << (ifrom[k] - 1)) | ((mask_t)1 << (ito[k] - 1));
        from[k] = (mask_t)1 << (ifrom[k] - 1);
    }
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num(); nt = omp_get_num_threads();
#endif
        int64_t len = c1 - c0;
        int64_t chunk = (len + nt - 1) / nt;
        int64_t a = c0 + (int64_t)tid * chunk, b = a + chunk < c1 ? a + chunk : c1;
        cplx* yt = buf + (size_t)tid * n;
        for (int64_t col = a; col < b; ++col) {
            cplx xc = x[col];
            if (xc == 0) continue;
            mask_t m = s->reps[col];
            double inv_sqrt_stab_col = 1.0 / sqrt((double)s->stab[col]);
            cplx hd = 0.0;
            for (int q = 0; q < n_den; ++q)
                if (((m >> (di[q] - 1)) & 1) && ((m >> (dj[q] - 1)) & 1)) hd = hd + V[q];
            yt[col] += hd * xc;
            for (int k = 0; k < n_bil; ++k) {
                if ((m & hop[k]) != from[k]) continue;
                mask_t nm = m ^ hop[k];
                cplx coeff;
                int64_t row = project(s, nm, &coeff);
                if (row == 0) continue;
                double hp = hop_phase(c->G.fermionic, m, ifrom[k], ito[k]);
                cplx h = (((t[k] * hp) * coeff) * sqrt((double)s->stab[row - 1])) * inv_sqrt_stab_col;
                yt[row - 1] += h * xc;
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) y[i] = 0;
    for (int t2 = 0; t2 < nthreads; ++t2) {
        const cplx* yt = buf + (size_t)t2 * n;
        for (int64_t i = 0; i < n; ++i) y[i] += yt[i];
    }
    free(buf); free(hop); free(from);
}

/* ------------------------------------------------------------------ */
/* sector observables (observables/ directory)                         */
/* ------------------------------------------------------------------ */

/* _build_density_correlation_matrix! (observables/static_structure_factor.jl:41-104)
 * with every site in both flavours and the disconnected part kept: for each
 * representative with amplitude c_col, each of the |G| images g(rep) adds
 * |c_col|^2/|G| to density[i] for its occupied i and to C[i][j] (row-major, 0-based)
 * for its occupied pairs (i, j), i == j included. */
void oc_density_correlation(const oc_sector* s, const double* c_, double* C, double* density) {
    const cplx* c = (const cplx*)c_;
    const oc_catalog* cat = s->cat;
    const int n = cat->G.n_site, nG = cat->G.n_group;
    for (int i = 0; i < n * n; ++i) C[i] = 0.0;
    for (int i = 0; i < n; ++i) density[i] = 0.0;
    const double nG_inv = 1.0 / nG;
    for (int64_t col = 0; col < s->dim; ++col) {
        const double w = creal(c[col]) * creal(c[col]) + cimag(c[col]) * cimag(c[col]);
        if (w == 0.0) continue;
        const double w_eff = w * nG_inv;
        for (int g = 0; g < nG; ++g) {
            cplx a;
            const mask_t shifted = apply_op(s->reps[col], &cat->G, g, &a);
            for (mask_t t = shifted; t; t &= t - 1) density[__builtin_ctzll(t)] += w_eff;
            for (mask_t t2 = shifted; t2; t2 &= t2 - 1) {
                const int i = __builtin_ctzll(t2);
                for (mask_t t3 = shifted; t3; t3 &= t3 - 1) C[i * n + __builtin_ctzll(t3)] += w_eff;
            }
        }
    }
}

/* off_diagonal_long_range_order (observables/off_diagonal_long_range_order.jl:74-126):
 * rho[i][j] = <psi| a_i^dag a_j |psi> (row-major, 0-based, complex interleaved) by the
 * orbit-basis contraction: every image U_g|rep> of every representative, every
 * occupied j and empty i, the scattered mask projected back onto the sector. */
void oc_one_body(const oc_sector* s, const double* c_, double* rho_) {
    const cplx* c = (const cplx*)c_;
    cplx* rho = (cplx*)rho_;
    const oc_catalog* cat = s->cat;
    const int n = cat->G.n_site, nG = cat->G.n_group;
    for (int i = 0; i < n * n; ++i) rho[i] = 0.0;
    const double inv_nG = 1.0 / nG, inv_sqrt_nG = 1.0 / sqrt((double)nG);
    for (int64_t col = 0; col < s->dim; ++col) {
        const cplx c_col = c[col];
        if (c_col == 0) continue;
        const double diag_weight = (creal(c_col) * creal(c_col) + cimag(c_col) * cimag(c_col)) * inv_nG;
        const double ket_norm = inv_sqrt_nG / sqrt((double)s->stab[col]);
        for (int g = 0; g < nG; ++g) {
            cplx ket_group_phase;
            const mask_t ket = apply_op(s->reps[col], &cat->G, g, &ket_group_phase);
            const cplx ket_basis_amp = (ket_norm * conj(s->chi[g])) * ket_group_phase;
            const cplx ket_amp = c_col * ket_basis_amp;
            for (mask_t tj = ket; tj; tj &= tj - 1) {
                const int j = __builtin_ctzll(tj) + 1;
                rho[(j - 1) * n + (j - 1)] += diag_weight;
                for (int i = 1; i <= n; ++i) {
                    if (i == j || ((ket >> (i - 1)) & 1)) continue;
                    const mask_t nm = (ket & ~((mask_t)1 << (j - 1))) | ((mask_t)1 << (i - 1));
                    cplx proj_coeff;
                    const int64_t row = project(s, nm, &proj_coeff);
                    if (row == 0) continue;
                    const cplx bra_basis_amp = (sqrt((double)s->stab[row - 1]) * inv_sqrt_nG) * conj(proj_coeff);
                    const cplx bra_amp = c[row - 1] * bra_basis_amp;
                    const double phase = hop_phase(cat->G.fermionic, ket, j, i);
                    rho[(i - 1) * n + (j - 1)] += (conj(bra_amp) * ket_amp) * phase;
                }
            }
        }
    }
}

/* _expand_sector_state_to_fock_amplitudes (observables/entanglement_spectrum.jl:46-79)
 * into a rank-indexed dense array out[C(N,k)] (complex): representatives with
 * |c| <= atol skipped, each image adds c/sqrt(|G| stab) conj(chi(g)) alpha_g(rep);
 * returns the number of amplitudes with |a| > atol and normalises those (the others
 * are set to zero), or -1 when the norm is <= atol (the reference errors). */
int64_t oc_fock_expand(const oc_sector* s, const double* c_, double atol, double* out_) {
    const cplx* c = (const cplx*)c_;
    cplx* out = (cplx*)out_;
    const oc_catalog* cat = s->cat;
    const int nG = cat->G.n_group;
    for (int64_t k = 0; k < cat->n_total; ++k) out[k] = 0.0;
    for (int64_t col = 0; col < s->dim; ++col) {
        const cplx coeff = c[col];
        if (cabs(coeff) <= atol) continue;
        const cplx prefactor = coeff / sqrt((double)nG * (double)s->stab[col]);
        for (int g = 0; g < nG; ++g) {
            cplx phase;
            const mask_t shifted = apply_op(s->reps[col], &cat->G, g, &phase);
            const cplx amp = (prefactor * conj(s->chi[g])) * phase;
            out[oc_rank_combination(shifted)] += amp;
        }
    }
    double norm2 = 0.0;
    int64_t cnt = 0;
    for (int64_t k = 0; k < cat->n_total; ++k) {
        if (cabs(out[k]) <= atol) { out[k] = 0.0; continue; }
        norm2 += creal(out[k]) * creal(out[k]) + cimag(out[k]) * cimag(out[k]);
        ++cnt;
    }
    if (!(norm2 > atol)) return -1;
    const double invnorm = 1.0 / sqrt(norm2);
    for (int64_t k = 0; k < cat->n_total; ++k) out[k] *= invnorm;
    return cnt;
}

/* _position_operator_matrix (observables/charge_pump.jl:47-91): per from-column, the
 * to-sector row of the same representative (1-based, 0 if absent) and
 * (1/|G|) sum_g chi_to(g) conj(chi_from(g)) z(g(rep)), z(mask) = prod phi_i. */
void oc_position_operator(const oc_sector* to, const oc_sector* from, const double* phi_, int64_t* rows,
                          double* vals_) {
    const cplx* phi = (const cplx*)phi_;
    cplx* vals = (cplx*)vals_;
    const oc_catalog* cat = from->cat;
    const int nG = cat->G.n_group;
    const double inv_nG = 1.0 / nG;
    for (int64_t col = 0; col < from->dim; ++col) {
        const mask_t repr = from->reps[col];
        const int64_t row = basis_index(to, repr);
        rows[col] = row;
        vals[col] = 0.0;
        if (row == 0) continue;
        cplx elem = 0.0;
        for (int g = 0; g < nG; ++g) {
            cplx a;
            const mask_t shifted = apply_op(repr, &cat->G, g, &a);
            cplx z = 1.0;
            for (mask_t t = shifted; t; t &= t - 1) z *= phi[__builtin_ctzll(t)];
            elem += (to->chi[g] * conj(from->chi[g])) * z;
        }
        vals[col] = elem * inv_nG;
    }
}

/* update_orbit_stabilizer_phases! (src/symmetry_resolved_ed.jl:338-369): the group's
 * site phases replaced in place (same permutations); stabilizer phases recomputed, and
 * for every representative and group element (in order) whose image lies in the
 * representative's orbit the rep_sym / rep_amp entries rewritten.  Returns 0, or -1
 * when the new group is not closed under inversion. */
int oc_catalog_update_phases(oc_catalog* c, const double* phases) {
    oc_group* G = &c->G;
    const int nG = G->n_group, n = G->n_site;
    int32_t* inv = malloc((size_t)nG * sizeof(int32_t));
    if (oc_inverse_indices(n, nG, G->perm, phases, inv) != 0) { free(inv); return -1; }
    memcpy(G->phases, phases, (size_t)nG * n * sizeof(cplx));
    memcpy(G->inv, inv, (size_t)nG * sizeof(int32_t));
    free(inv);
    for (int64_t o = 0; o < c->n_orbits; ++o)
        for (int j = 0; j < nG && c->stab_gidx[o * nG + j] != 0; ++j) {
            cplx a;
            (void)apply_op(c->reps[o], G, c->stab_gidx[o * nG + j] - 1, &a);
            c->stab_phase[o * nG + j] = a;
        }
    for (int64_t o = 0; o < c->n_orbits; ++o)
        for (int g = 0; g < nG; ++g) {
            cplx a;
            const mask_t s = apply_op(c->reps[o], G, g, &a);
            const int64_t idx = oc_rank_combination(s);
            if (c->rep_rank[idx] == (int32_t)(o + 1)) {
                c->rep_sym[idx] = G->inv[g];
                c->rep_amp[idx] = conj(a);
            }
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Sectors too large for the catalog tables: a listed basis checked    */
/* against its definition, and sampled rows of H x                     */
/* ------------------------------------------------------------------ */

/* A group view over caller arrays (no inverse table: only apply_op is used). */
static oc_group group_view(int n_site, int n_group, const int32_t* perm, const double* phases, int fermionic) {
    oc_group G;
    G.n_site = n_site; G.n_group = n_group; G.fermionic = fermionic;
    G.perm = (int32_t*)perm; G.phases = (cplx*)phases; G.inv = NULL;
    return G;
}

/* The sector basis of build_symmetry_sector_basis (:544-556) is the ascending list of
 * catalog representatives (masks with n_filled bits that no group element maps lower,
 * :442-453) whose stabilizer phases equal the characters (_is_orbit_compatible,
 * :534-542), each with its stabilizer order.  Checks a listed basis entry by entry
 * against that definition: strictly ascending, n_filled bits, canonical, stabilizer
 * order, compatibility.  Together with a length equal to the character-formula
 * dimension this fixes the list.  Returns the number of entries that fail (and the
 * first one, 0-based, in *first_bad; -1 if none). */
int64_t oc_check_sector_basis(int n_site, int n_filled, int n_group, const int32_t* perm, const double* phases,
                              int fermionic, const double* chi_, double atol, int64_t dim, const uint64_t* reps,
                              const int64_t* stab, int nthreads, int64_t* first_bad) {
    const oc_group G = group_view(n_site, n_group, perm, phases, fermionic);
    const cplx* chi = (const cplx*)chi_;
    int64_t bad = 0, first = -1;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static, 4096) reduction(+ : bad)
    for (int64_t r = 0; r < dim; ++r) {
        const mask_t m = reps[r];
        int ok = __builtin_popcountll(m) == n_filled && (n_site == 64 || (m >> n_site) == 0);
        if (ok && r > 0 && !(reps[r - 1] < m)) ok = 0;
        int64_t so = 0;
        for (int g = 0; g < n_group && ok; ++g) {
            cplx a;
            const mask_t s = apply_op(m, &G, g, &a);
            if (s < m) ok = 0;
            else if (s == m) {
                ++so;
                if (!(cabs(chi[g] - a) <= atol)) ok = 0;
            }
        }
        if (ok && so != stab[r]) ok = 0;
        if (!ok) {
            ++bad;
#pragma omp critical
            if (first < 0 || r < first) first = r;
        }
    }
    *first_bad = first;
    return bad;
}

static int64_t find_rep(const mask_t* reps, int64_t dim, mask_t m) {
    int64_t lo = 0, hi = dim - 1;
    while (lo <= hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (reps[mid] == m) return mid;
        if (reps[mid] < m) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static int is_canonical(mask_t m, const oc_group* G) {
    for (int g = 0; g < G->n_group; ++g) {
        cplx a;
        if (apply_op(m, G, g, &a) < m) return 0;
    }
    return 1;
}

/* (H x)[r] for the listed rows r (0-based) of a sector given by its basis (checked with
 * oc_check_sector_basis), without rep tables.  H[r, c] is the entry
 * build_matrixfree_projection_table (:680-728) writes for column c: term k = (i_from,
 * i_to, t) hops the column's representative m_c to nm = m_c xor hop_k when i_from is
 * occupied and i_to empty (:709-710), nm projects onto row r when its canonical
 * representative (get_canonical_representative, :174-187, via g with alpha = the
 * phase of g on nm) is rep_r, and the entry is ((t * hop_phase) * alpha conj(chi(g))) *
 * sqrt(stab_r / stab_c) (:718).  So the columns reaching row r are found from the row:
 * for each distinct image s of rep_r and each term with i_to occupied and i_from empty
 * in s, m_c = s xor hop_k must itself be a representative in the basis.  The diagonal
 * is the sum of V over density terms with both sites occupied (:700-706).  No use is
 * made of Hermiticity. */
void oc_sample_rows(int n_site, int n_group, const int32_t* perm, const double* phases, int fermionic,
                    const double* chi_, int64_t dim, const uint64_t* reps, const int64_t* stab, int n_bil,
                    const int32_t* ifrom, const int32_t* ito, const double* t_, int n_den, const int32_t* di,
                    const int32_t* dj, const double* V_, const double* x_, int64_t n_rows, const int64_t* rows,
                    double* y_, int nthreads) {
    const oc_group G = group_view(n_site, n_group, perm, phases, fermionic);
    const cplx* chi = (const cplx*)chi_;
    const cplx* t = (const cplx*)t_;
    const cplx* V = (const cplx*)V_;
    const cplx* x = (const cplx*)x_;
    cplx* y = (cplx*)y_;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
    for (int64_t q = 0; q < n_rows; ++q) {
        const int64_t r = rows[q];
        const mask_t mr = reps[r];
        cplx acc = 0.0;
        cplx hd = 0.0;
        for (int p = 0; p < n_den; ++p)
            if (((mr >> (di[p] - 1)) & 1) && ((mr >> (dj[p] - 1)) & 1)) hd = hd + V[p];
        acc = hd * x[r];
        /* distinct images of rep_r */
        mask_t img[64];
        int n_img = 0;
        for (int g = 0; g < n_group; ++g) {
            cplx a;
            const mask_t s = apply_op(mr, &G, g, &a);
            int seen = 0;
            for (int u = 0; u < n_img && !seen; ++u) seen = img[u] == s;
            if (!seen) img[n_img++] = s;
        }
        for (int u = 0; u < n_img; ++u) {
            const mask_t s = img[u];
            /* get_canonical_representative of s: first minimal element (strict <) */
            mask_t repr = s;
            int gbest = 0;
            cplx abest = 1.0;
            for (int g = 0; g < n_group; ++g) {
                cplx a;
                const mask_t w = apply_op(s, &G, g, &a);
                if (w < repr) { repr = w; gbest = g; abest = a; }
            }
            const cplx coeff = abest * conj(chi[gbest]);
            for (int k = 0; k < n_bil; ++k) {
                const int f = ifrom[k] - 1, to = ito[k] - 1;
                if (f == to) continue;
                if (!((s >> to) & 1) || ((s >> f) & 1)) continue;
                const mask_t mc = s ^ (((mask_t)1 << f) | ((mask_t)1 << to));
                if (!is_canonical(mc, &G)) continue;
                const int64_t c = find_rep(reps, dim, mc);
                if (c < 0) continue;
                const double hp = hop_phase(fermionic, mc, ifrom[k], ito[k]);
                const double nrm = sqrt((double)stab[r] / (double)stab[c]);
                const cplx h = ((t[k] * hp) * coeff) * nrm;
                acc += h * x[c];
            }
        }
        y[q] = acc;
    }
}
