"""The benchmark workloads at their full sizes, through the C ABI.

The oracle's catalog cannot be built at these sizes (C(48, 8) = 377M ranks), so parity
rests on the table-free oracle and on size-independent properties:
  * the GPU's sector basis checked entry by entry against its definition (canonical
    representative, stabilizer order, stabilizer phases = characters, ascending) with
    the character-formula length, which together fix the list;
  * (H x)[r] on 10,000 sampled rows from the oracle's row restatement of the projection
    table (oc_sample_rows: no rep tables, no Hermiticity) against the GPU matvec, FP64
    and FP32, for every launch shape on the checkerboard 4x6 workload;
  * eigenpairs: the oracle's rows of H v equal lambda v on the sampled rows, plus
    residuals, orthonormality, other seeds and FP32+refine against FP64;
  * sector dimensions against the character formula (closed form C1) and their sum
    against C(N, k); Hermiticity and linearity of the matvec; FP32 against FP64.
The XXZ N=16 sector (dim 810) is small enough for the oracle's dense spectrum.
"""
import math

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import _lib
from realspace_exactdiagonalization_amd import lattice as L
from tests.test_oracle import character_dim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ed():
    from realspace_exactdiagonalization_amd import ed as E
    return E


@pytest.fixture(scope="module")
def ctx(ed):
    c = ed.Context(0)
    yield c
    c.close()


N_ROWS = 10000


class FullSector:
    """Catalog, sector and terms of one workload at its full size, with the basis dump
    checked against its definition once."""

    def __init__(self, ed, ctx, m, theta):
        self.m = L.twisted(m, theta)
        self.G = L.model_group(self.m, theta)
        self.ir = L.translation_irreps(self.G, self.m.lattice)[0]
        self.cat = catalog(ed, ctx, self.m, self.G)
        self.b = ed.build_symmetry_sector_basis(self.cat, self.ir)
        self.H = ed.MatrixFreeHamiltonian(self.b, self.m.bilinear_terms, self.m.density_terms)
        self.reps = self.b.representative_mask_list
        self.stab = self.b.stabilizer_order_list
        self.checked = None

    def check_basis(self, oracle_mod):
        if self.checked is None:
            bad = oracle_mod.check_sector_basis(self.G, self.m.n_filled, self.m.statistics == L.FERMIONIC,
                                                self.ir.values, self.reps, self.stab)
            self.checked = (bad, len(self.reps) == round(character_dim(self.m, self.G, self.ir.values)))
        return self.checked

    def rows(self, seed=0):
        rng = np.random.default_rng(seed)
        r = rng.choice(self.b.dim, size=min(N_ROWS, self.b.dim), replace=False)
        return np.unique(np.concatenate([r, [0, self.b.dim - 1]]))

    def oracle_rows(self, oracle_mod, x, rows):
        return oracle_mod.sample_rows(self.m, self.G, self.ir.values, self.reps, self.stab, x, rows)

    def close(self):
        self.b.close()
        self.cat.close()


FULL = {"checker4x6": lambda: L.workload("checker4x6"), "hubbard4x4": lambda: L.workload("hubbard4x4"),
        "xxz32": lambda: L.heisenberg_chain(32), "xxz34": lambda: L.heisenberg_chain(34)}


@pytest.fixture(scope="module")
def full(ed, ctx):
    cache = {}

    def get(name, theta):
        key = (name, tuple(theta))
        if key not in cache:
            for v in cache.values():   # one workload on the device at a time
                v.close()
            cache.clear()
            print(f"[full] building {name} theta={theta}", flush=True)
            cache[key] = FullSector(ed, ctx, FULL[name](), theta)
        return cache[key]

    yield get
    for v in cache.values():
        v.close()


# XXZ N=34 at half filling: C(34, 17) = 2,333,606,220 ranks (above 2^31) and a k = 0
# sector of 68,635,478 rows, which takes the 8-byte sector table
FULL_CASES = [("checker4x6", (0.0, 0.0)), ("checker4x6", (0.3, 0.7)), ("hubbard4x4", (0.0, 0.0)),
              ("xxz32", (0.0, 0.0)), ("xxz34", (0.0, 0.0))]


@pytest.mark.parametrize("name,theta", FULL_CASES)
def test_full_size_sector_basis_matches_its_definition(oracle_mod, full, name, theta):
    fs = full(name, theta)
    print(f"[full-basis] {name} theta={theta} dim={fs.b.dim}", flush=True)
    (bad, first), length_ok = fs.check_basis(oracle_mod)
    assert (bad, first) == (0, -1)
    assert length_ok


@pytest.mark.parametrize("name,theta", FULL_CASES)
def test_full_size_sampled_rows_match_the_oracle(ed, oracle_mod, full, name, theta):
    """GPU H x on the autotuned launch shape against the oracle's rows, FP64 (1e-12) and
    FP32 (5e-6), relative to the largest sampled |(H x)[r]|."""
    fs = full(name, theta)
    assert fs.check_basis(oracle_mod) == ((0, -1), True)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(fs.b.dim) + 1j * rng.standard_normal(fs.b.dim)
    rows = fs.rows(1)
    ref = fs.oracle_rows(oracle_mod, x, rows)
    scale = np.abs(ref).max()
    for prec, tol in ((ed.PREC_FP64, 1e-12), (ed.PREC_FP32, 5e-6)):
        fs.H.precision = prec
        y = fs.H(x)
        err = np.abs(y[rows] - ref).max() / scale
        print(f"[full-rows] {name} theta={theta} prec={prec} shape={fs.b.shape(prec)['shape']} "
              f"rel_err={err:.2e}", flush=True)
        assert err <= tol
    fs.H.precision = ed.PREC_FP64


def test_full_size_every_launch_shape_matches_the_oracle(ed, oracle_mod, full):
    """Every launch shape on the checkerboard 4x6 workload (48 sites, 64-bit masks,
    fermions) against the oracle's rows.  The flat hop loop needs C(47, 9) < 2^32 and is
    refused there with SED_EUNSUPPORTED."""
    fs = full("checker4x6", (0.0, 0.0))
    assert fs.check_basis(oracle_mod) == ((0, -1), True)
    rng = np.random.default_rng(23)
    x = rng.standard_normal(fs.b.dim) + 1j * rng.standard_normal(fs.b.dim)
    rows = fs.rows(2)
    ref = fs.oracle_rows(oracle_mod, x, rows)
    scale = np.abs(ref).max()
    n_shapes = len(fs.b.shape(ed.PREC_FP64)["tune_ms"])
    ran = 0
    for prec, tol in ((ed.PREC_FP64, 1e-12), (ed.PREC_FP32, 5e-6)):
        fs.H.precision = prec
        for sh in range(n_shapes):
            fs.b.set_shape(prec, sh, 64)
            try:
                y = fs.H(x)
            except _lib.SedError as e:
                assert e.code == _lib.SED_EUNSUPPORTED, (sh, e)
                print(f"[full-shape] prec={prec} shape={sh} refused: {e}", flush=True)
                continue
            err = np.abs(y[rows] - ref).max() / scale
            print(f"[full-shape] prec={prec} shape={sh} rel_err={err:.2e}", flush=True)
            assert err <= tol, (prec, sh)
            ran += 1
        fs.b.set_shape(prec, -1)
    fs.H.precision = ed.PREC_FP64
    assert ran >= 2 * 9


@pytest.mark.parametrize("name", ["checker4x6", "hubbard4x4", "xxz34"])
def test_full_size_eigenpairs_satisfy_the_oracle_rows(ed, oracle_mod, full, name):
    """The lowest three eigenpairs of the k = (0,0) sector, FP64 and FP32+refine: on the
    sampled rows the oracle's (H v)[r] equals lambda v[r] (2e-10, the solver's residual
    tolerance being 1e-10), the vectors are orthonormal, FP32+refine and a second seed
    give the FP64 eigenvalues to 1e-10 relative."""
    fs = full(name, (0.0, 0.0))
    assert fs.check_basis(oracle_mod) == ((0, -1), True)
    rows = fs.rows(3)
    out = {}
    for prec, seed in ((ed.PREC_FP64, 1234), (ed.PREC_FP32_REFINE, 1234), (ed.PREC_FP64, 99)):
        st = {}
        vals, vecs = ed.diagonalize_block_matrixfree(fs.H, fs.b.dim, nev=3, precision=prec, seed=seed, stats=st)
        print(f"[full-eig] {name} prec={prec} seed={seed} vals={vals.tolist()} "
              f"steps={st['iters_fp32']}+{st['iters_fp64']} resid={st['residual_max']:.2e}", flush=True)
        gram = vecs.conj().T @ vecs
        assert np.abs(gram - np.eye(3)).max() <= 1e-10
        for j in range(3):
            hv = fs.oracle_rows(oracle_mod, vecs[:, j], rows)
            assert np.abs(hv - vals[j] * vecs[rows, j]).max() <= 2e-10, (prec, j)
        out[(prec, seed)] = vals
    ref = out[(ed.PREC_FP64, 1234)]
    for v in out.values():
        assert np.abs(v - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["xxz34", "hubbard4x4", "checker4x6"])
def test_full_size_observables_reproduce_the_energy(ed, full, name):
    """Observables of the ground state at full size, including the 8-byte-table sector of
    XXZ N=34: the one-body matrix and <n_i n_j> reproduce the eigenvalue through the model's
    own terms, E = sum_k t_k <a^dag_to a_from> + sum_d V_d <n_i n_j> (1e-9 relative); the
    one-body matrix is Hermitian with the site densities on its diagonal and trace N; the
    densities are translation invariant per sublattice orbit and sum to N; and
    sum_j <n_i n_j> = N <n_i>."""
    fs = full(name, (0.0, 0.0))
    m = fs.m
    vals, _ = ed.diagonalize_block_matrixfree(fs.H, fs.b.dim, nev=1)
    e0 = float(vals[0])
    d = ed.sector_density(fs.H, 0)
    C, d2 = ed.density_correlation_matrix(fs.H, 0, subtract_disconnected=False)
    rho = ed.off_diagonal_long_range_order(fs.H, 0)
    k = m.n_filled
    assert np.abs(d - d2).max() < 1e-12 and abs(d.sum() - k) < 1e-9
    assert np.abs(np.diag(rho).real - d).max() < 1e-10 and abs(np.trace(rho) - k) < 1e-9
    assert np.abs(rho - rho.conj().T).max() < 1e-12
    assert np.abs(C.sum(axis=1) - k * d).max() < 1e-9
    for g in range(fs.G.order):   # a translation-invariant state: <n_g(i)> = <n_i>
        assert np.abs(d[fs.G.perm[g] - 1] - d).max() < 1e-10
    e = sum(complex(t) * rho[b - 1, a - 1] for a, b, t in m.bilinear_terms)
    e += sum(complex(v) * C[i - 1, j - 1] for i, j, v in m.density_terms)
    print(f"[full-obs] {name} E0={e0:.12f} from observables {e.real:.12f} (imag {e.imag:.1e})", flush=True)
    assert abs(e.imag) < 1e-9 * abs(e0) and abs(e.real - e0) < 1e-9 * abs(e0)


def test_xxz34_every_launch_shape_on_the_8_byte_table(ed, oracle_mod, full):
    """The k = 0 sector of XXZ N=34 (68,635,478 rows at |G| = 34) does not fit the 4-byte
    table entry (row < 2^25 beside 7 group and parity bits) and runs on the 8-byte table:
    every launch shape (site by site and flat: C(33, 18) < 2^32) against the oracle's rows,
    FP64 and FP32."""
    fs = full("xxz34", (0.0, 0.0))
    assert fs.b.dim == 68635478 and fs.b.dim >= 1 << 25
    assert fs.check_basis(oracle_mod) == ((0, -1), True)
    rng = np.random.default_rng(29)
    x = rng.standard_normal(fs.b.dim) + 1j * rng.standard_normal(fs.b.dim)
    rows = fs.rows(4)
    ref = fs.oracle_rows(oracle_mod, x, rows)
    scale = np.abs(ref).max()
    n_shapes = len(fs.b.shape(ed.PREC_FP64)["tune_ms"])
    ran, refused = 0, 0
    for prec, tol in ((ed.PREC_FP64, 1e-12), (ed.PREC_FP32, 5e-6)):
        fs.H.precision = prec
        for sh in range(n_shapes):
            fs.b.set_shape(prec, sh, 0)
            try:
                y = fs.H(x)
            except _lib.SedError as e:
                assert e.code == _lib.SED_EUNSUPPORTED, (sh, e)
                refused += 1
                continue
            err = np.abs(y[rows] - ref).max() / scale
            print(f"[xxz34-shape] prec={prec} shape={sh} rel_err={err:.2e}", flush=True)
            assert err <= tol, (prec, sh)
            ran += 1
        fs.b.set_shape(prec, -1)
    fs.H.precision = ed.PREC_FP64
    assert ran == 2 * n_shapes and refused == 0


def catalog(ed, ctx, m, G):
    return ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                           particle_statistics=m.statistics, ctx=ctx)


@pytest.mark.parametrize("name", ["xxz16", "haldane3x4", "hubbard4x4", "checker4x6"])
def test_every_sector_dimension_matches_the_character_formula(ed, ctx, name):
    m = L.workload(name)
    G = L.model_group(m)
    print(f"[dims] {name} N={m.n_site} k={m.n_filled} |G|={G.order}", flush=True)
    cat = catalog(ed, ctx, m, G)
    assert cat.n_total == math.comb(m.n_site, m.n_filled)
    total = 0
    for ir in L.translation_irreps(G, m.lattice):
        b = ed.build_symmetry_sector_basis(cat, ir)
        assert b.dim == round(character_dim(m, G, ir.values)), ir.label
        total += b.dim
        b.close()
    # every state has exactly one momentum: the sectors partition the fixed-N space
    assert total == cat.n_total
    cat.close()


@pytest.mark.parametrize("name,expected_dim", [("hubbard4x4", 4031196), ("checker4x6", 15724214)])
def test_full_size_matvec_is_hermitian_linear_and_fp32_consistent(ed, ctx, name, expected_dim):
    m = L.workload(name)
    G = L.model_group(m)
    print(f"[matvec] {name}", flush=True)
    cat = catalog(ed, ctx, m, G)
    b = ed.build_symmetry_sector_basis(cat, L.translation_irreps(G, m.lattice)[0])
    assert b.dim == expected_dim
    H = ed.MatrixFreeHamiltonian(b, m.bilinear_terms, m.density_terms)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)
    y = rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)
    Hx, Hy = H(x), H(y)
    # <y, H x> = <H y, x>
    lhs, rhs = np.vdot(y, Hx), np.vdot(Hy, x)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Hx) * np.linalg.norm(y)
    # linearity with complex coefficients
    a, c = 0.3 - 1.1j, -0.7 + 0.2j
    H2 = H(a * x + c * y)
    assert np.linalg.norm(H2 - (a * Hx + c * Hy)) <= 1e-12 * np.linalg.norm(H2)
    # FP32 vectors and accumulation against FP64
    H.precision = ed.PREC_FP32
    Hx32 = H(x)
    assert np.linalg.norm(Hx32 - Hx) <= 1e-5 * np.linalg.norm(Hx)
    b.close()
    cat.close()


def test_xxz16_lowest_four_match_the_oracle(ed, ctx, oracle_mod):
    m = L.workload("xxz16")
    G = L.model_group(m)
    ir = L.translation_irreps(G, m.lattice)[0]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, False)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    ref = np.linalg.eigvalsh(os_.dense())[:4]
    cat = catalog(ed, ctx, m, G)
    b = ed.build_symmetry_sector_basis(cat, ir)
    assert b.dim == 810
    H, n = ed.hamiltonian_linear_operator(b, m.bilinear_terms, m.density_terms)
    for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
        vals = ed.diagonalize_block_matrixfree(H, n, nev=4, tol=1e-12, precision=prec, return_vectors=False)[0]
        assert np.abs(vals - ref).max() <= 1e-12 * np.abs(ref).max(), (prec, vals, ref)


def test_hubbard4x4_fp32_refine_matches_fp64(ed, ctx):
    m = L.workload("hubbard4x4")
    G = L.model_group(m)
    print("[solve] hubbard4x4 k=(0,0)", flush=True)
    cat = catalog(ed, ctx, m, G)
    b = ed.build_symmetry_sector_basis(cat, L.translation_irreps(G, m.lattice)[0])
    H, n = ed.hamiltonian_linear_operator(b, m.bilinear_terms, m.density_terms)
    v64 = ed.diagonalize_block_matrixfree(H, n, nev=3, precision=ed.PREC_FP64, return_vectors=False)[0]
    v32 = ed.diagonalize_block_matrixfree(H, n, nev=3, precision=ed.PREC_FP32_REFINE, return_vectors=False)[0]
    assert np.abs(v32 - v64).max() <= 1e-10 * np.abs(v64).max(), (v32, v64)
    b.close()
    cat.close()


def test_xxz34_catalog_above_2e31_ranks_and_every_sector_dimension(ed, ctx):
    """XXZ N=34 at half filling: C(34, 17) = 2,333,606,220 ranks, above the int32 rank range
    (the catalog stores only representative flags in its first pass).  Every momentum
    sector's dimension equals the character formula, the sectors partition the ranks, and
    the orbit count equals the number of necklaces (Burnside over the 34 translations)."""
    m = L.heisenberg_chain(34)
    G = L.model_group(m)
    print(f"[xxz34] catalog N=34 k=17 |G|={G.order}", flush=True)
    cat = catalog(ed, ctx, m, G)
    assert cat.n_total == math.comb(34, 17) == 2333606220
    necklaces = sum(math.comb(math.gcd(d, 34), 17 * math.gcd(d, 34) // 34) for d in range(34)
                    if (17 * math.gcd(d, 34)) % 34 == 0) // 34
    assert cat.n_orbits == necklaces
    total = 0
    for ir in L.translation_irreps(G, m.lattice):
        b = ed.build_symmetry_sector_basis(cat, ir)
        want = round(character_dim(m, G, ir.values))
        print(f"[xxz34] sector {ir.label} dim={b.dim} formula={want}", flush=True)
        assert b.dim == want, ir.label
        total += b.dim
        b.close()
    assert total == cat.n_total
    cat.close()
