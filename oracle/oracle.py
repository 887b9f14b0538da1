"""ctypes front end of the CPU oracle (``oracle/sed_oracle.c``).

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module.  It checks the HIP path;
it is never part of it.

Each function mirrors the reference function of the same name
(``src/symmetry_resolved_ed.jl``); eigenvalues come from LAPACK (dense) or
ARPACK via scipy (sparse, the eigensolver of the reference's matrix mode,
``:978-995``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsed_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oc_binomial.restype = C.c_int64
        L.oc_binomial.argtypes = [C.c_int, C.c_int]
        L.oc_rank_combination.restype = C.c_int64
        L.oc_rank_combination.argtypes = [C.c_uint64]
        L.oc_gosper_next.restype = C.c_uint64
        L.oc_gosper_next.argtypes = [C.c_uint64]
        L.oc_lin_table.restype = C.c_int
        L.oc_lin_table.argtypes = [C.c_int, C.c_int, _i64p, _i64p]
        L.oc_lin_index.restype = C.c_int64
        L.oc_lin_index.argtypes = [C.c_uint64, _i64p, _i64p, C.c_int]
        L.oc_inverse_indices.restype = C.c_int
        L.oc_inverse_indices.argtypes = [C.c_int, C.c_int, _i32p, _f64p, _i32p]
        L.oc_catalog_build.restype = C.c_void_p
        L.oc_catalog_build.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, C.c_int]
        L.oc_catalog_free.argtypes = [C.c_void_p]
        L.oc_catalog_n_orbits.restype = C.c_int64
        L.oc_catalog_n_orbits.argtypes = [C.c_void_p]
        L.oc_catalog_n_total.restype = C.c_int64
        L.oc_catalog_n_total.argtypes = [C.c_void_p]
        L.oc_catalog_tables.argtypes = [C.c_void_p, _i32p, _i32p, _f64p]
        L.oc_catalog_orbits.argtypes = [C.c_void_p, _u64p, _i64p, _i32p, _f64p]
        L.oc_canonical_representative.restype = C.c_uint64
        L.oc_canonical_representative.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32), _f64p]
        L.oc_apply_operation.restype = C.c_uint64
        L.oc_apply_operation.argtypes = [C.c_void_p, C.c_uint64, C.c_int, _f64p]
        L.oc_catalog_wrap.restype = C.c_void_p
        L.oc_catalog_wrap.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, C.c_int, C.c_int64, _u64p, _i32p,
                                      _i32p, _f64p]
        L.oc_sector_from_basis.restype = C.c_void_p
        L.oc_sector_from_basis.argtypes = [C.c_void_p, _f64p, C.c_int64, _u64p, _i64p]
        L.oc_sector_build.restype = C.c_void_p
        L.oc_sector_build.argtypes = [C.c_void_p, _f64p, C.c_double]
        L.oc_sector_free.argtypes = [C.c_void_p]
        L.oc_sector_dim.restype = C.c_int64
        L.oc_sector_dim.argtypes = [C.c_void_p]
        L.oc_sector_basis.argtypes = [C.c_void_p, _u64p, _i64p]
        L.oc_project_to_sector.restype = C.c_int64
        L.oc_project_to_sector.argtypes = [C.c_void_p, C.c_uint64, _f64p]
        L.oc_project_to_unnormalized_sector.restype = C.c_int64
        L.oc_project_to_unnormalized_sector.argtypes = [C.c_void_p, C.c_uint64, _f64p]
        L.oc_table_build.restype = C.c_int64
        L.oc_table_build.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p, _f64p]
        L.oc_table_build_cols.restype = C.c_int64
        L.oc_table_build_cols.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p, _f64p,
                                          C.c_int64, C.c_int64]
        L.oc_table_build_strided.restype = C.c_int64
        L.oc_table_build_strided.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p,
                                             _f64p, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
        L.oc_table_copy.argtypes = [C.c_void_p, _i64p, _i64p, _f64p, _f64p]
        L.oc_apply_table.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int]
        L.oc_apply_table_timed.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, _f64p]
        L.oc_apply_hamiltonian.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p,
                                           _f64p, _f64p, _f64p, C.c_int, C.c_int64, C.c_int64]
        L.oc_density_correlation.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.oc_one_body.argtypes = [C.c_void_p, _f64p, _f64p]
        L.oc_catalog_update_phases.restype = C.c_int
        L.oc_catalog_update_phases.argtypes = [C.c_void_p, _f64p]
        L.oc_fock_expand.restype = C.c_int64
        L.oc_fock_expand.argtypes = [C.c_void_p, _f64p, C.c_double, _f64p]
        L.oc_position_operator.argtypes = [C.c_void_p, C.c_void_p, _f64p, _i64p, _f64p]
        L.oc_check_sector_basis.restype = C.c_int64
        L.oc_check_sector_basis.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, C.c_int, _f64p, C.c_double,
                                            C.c_int64, _u64p, _i64p, C.c_int, C.POINTER(C.c_int64)]
        L.oc_sample_rows.argtypes = [C.c_int, C.c_int, _i32p, _f64p, C.c_int, _f64p, C.c_int64, _u64p, _i64p,
                                     C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p, _f64p, _f64p,
                                     C.c_int64, _i64p, _f64p, C.c_int]
        _lib = L
    return _lib


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


def binomial(n: int, k: int) -> int:
    return int(lib().oc_binomial(n, k))


def rank_combination(m: int) -> int:
    return int(lib().oc_rank_combination(int(m)))


def gosper_next(x: int) -> int:
    return int(lib().oc_gosper_next(int(x)))


def lin_table(n_site: int, n_filled: int):
    n_right = n_site // 2
    left = np.zeros(1 << (n_site - n_right), dtype=np.int64)
    right = np.zeros(1 << n_right, dtype=np.int64)
    nr = lib().oc_lin_table(n_site, n_filled, left, right)
    return left, right, nr


def lin_index(m: int, left, right, n_right: int) -> int:
    return int(lib().oc_lin_index(int(m), left, right, n_right))


def inverse_operation_indices(group) -> np.ndarray:
    perm = np.ascontiguousarray(group.perm, dtype=np.int32)
    out = np.zeros(group.order, dtype=np.int32)
    if lib().oc_inverse_indices(group.n_site, group.order, perm, _c128(group.phases), out) != 0:
        raise RuntimeError("Symmetry operations are not closed under inversion.")
    return out


class Catalog:
    """``build_symmetry_orbit_catalog`` (``src/symmetry_resolved_ed.jl:371-487``)."""

    def __init__(self, n_site: int, n_filled: int, group, fermionic: bool):
        self.group = group
        self.n_site = n_site
        self.n_filled = n_filled
        self.fermionic = bool(fermionic)
        perm = np.ascontiguousarray(group.perm, dtype=np.int32)
        self._h = lib().oc_catalog_build(n_site, n_filled, group.order, perm, _c128(group.phases),
                                         1 if fermionic else 0)
        if not self._h:
            raise RuntimeError("oracle catalog build failed")
        L = lib()
        self.n_orbits = int(L.oc_catalog_n_orbits(self._h))
        self.n_total = int(L.oc_catalog_n_total(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oc_catalog_free(self._h)
            self._h = None

    def tables(self):
        """(rep_rank_table int32 1-based, rep_sym_table int32 1-based, rep_amp_table c128)."""
        rr = np.zeros(self.n_total, dtype=np.int32)
        rs = np.zeros(self.n_total, dtype=np.int32)
        ra = np.zeros(self.n_total, dtype=np.complex128)
        lib().oc_catalog_tables(self._h, rr, rs, ra.view(np.float64))
        return rr, rs, ra

    def orbits(self):
        """(reps u64, stab_order i64, stab_gidx (n_orbits, |G|) 1-based 0-padded, stab_phase)."""
        nG = self.group.order
        reps = np.zeros(self.n_orbits, dtype=np.uint64)
        so = np.zeros(self.n_orbits, dtype=np.int64)
        gi = np.zeros((self.n_orbits, nG), dtype=np.int32)
        gp = np.zeros((self.n_orbits, nG), dtype=np.complex128)
        lib().oc_catalog_orbits(self._h, reps, so, gi.reshape(-1), gp.view(np.float64).reshape(-1))
        return reps, so, gi, gp

    def update_phases(self, new_group):
        """``update_orbit_stabilizer_phases!`` (``src/symmetry_resolved_ed.jl:338-369``)."""
        if lib().oc_catalog_update_phases(self._h, _c128(new_group.phases)) != 0:
            raise RuntimeError("Symmetry operations are not closed under inversion.")
        self.group = new_group
        return self

    def canonical_representative(self, m: int):
        g = C.c_int32(0)
        a = np.zeros(2)
        r = lib().oc_canonical_representative(self._h, int(m), C.byref(g), a)
        return int(r), int(g.value), complex(a[0], a[1])

    def apply_operation(self, m: int, g1: int):
        a = np.zeros(2)
        r = lib().oc_apply_operation(self._h, int(m), int(g1), a)
        return int(r), complex(a[0], a[1])


class WrappedCatalog(Catalog):
    """Catalog over externally produced tables in the reference layout (borrowed)."""

    def __init__(self, n_site, n_filled, group, fermionic, reps, rep_rank, rep_sym, rep_amp):
        self.group = group
        self.n_site = n_site
        self.n_filled = n_filled
        self.fermionic = bool(fermionic)
        self._arrays = (np.ascontiguousarray(reps, dtype=np.uint64), np.ascontiguousarray(rep_rank, dtype=np.int32),
                        np.ascontiguousarray(rep_sym, dtype=np.int32),
                        np.ascontiguousarray(rep_amp, dtype=np.complex128))
        r, rr, rs, ra = self._arrays
        perm = np.ascontiguousarray(group.perm, dtype=np.int32)
        self._h = lib().oc_catalog_wrap(n_site, n_filled, group.order, perm, _c128(group.phases),
                                        1 if fermionic else 0, len(r), r, rr, rs, ra.view(np.float64))
        if not self._h:
            raise RuntimeError("oracle catalog wrap failed")
        self.n_orbits = len(r)
        self.n_total = len(rr)


class Sector:
    """``build_symmetry_sector_basis`` + projection table + matvecs."""

    def __init__(self, catalog: Catalog, chi, atol: float = 1e-12, basis=None):
        self.catalog = catalog
        self.chi = np.asarray(chi, dtype=np.complex128)
        if basis is None:
            self._h = lib().oc_sector_build(catalog._h, _c128(self.chi), atol)
        else:
            reps = np.ascontiguousarray(basis[0], dtype=np.uint64)
            stab = np.ascontiguousarray(basis[1], dtype=np.int64)
            self._h = lib().oc_sector_from_basis(catalog._h, _c128(self.chi), len(reps), reps, stab)
        self.dim = int(lib().oc_sector_dim(self._h))
        self.nnz = None
        self._terms = None

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oc_sector_free(self._h)
            self._h = None

    def basis(self):
        reps = np.zeros(self.dim, dtype=np.uint64)
        stab = np.zeros(self.dim, dtype=np.int64)
        lib().oc_sector_basis(self._h, reps, stab)
        return reps, stab

    def project_to_sector(self, m: int):
        c = np.zeros(2)
        r = lib().oc_project_to_sector(self._h, int(m), c)
        return None if r == 0 else (int(r), complex(c[0], c[1]))

    def project_to_unnormalized_sector(self, m: int):
        c = np.zeros(2)
        r = lib().oc_project_to_unnormalized_sector(self._h, int(m), c)
        return None if r == 0 else (int(r), complex(c[0], c[1]))

    def set_terms(self, i_from, i_to, t, di, dj, V, build_table: bool = True):
        """Attach the term lists; with build_table the projection table of
        build_matrixfree_projection_table is built (needed by apply_table)."""
        self._terms = (np.ascontiguousarray(i_from, dtype=np.int32), np.ascontiguousarray(i_to, dtype=np.int32),
                       _c128(t), np.ascontiguousarray(di, dtype=np.int32), np.ascontiguousarray(dj, dtype=np.int32),
                       _c128(V))
        if not build_table:
            return None
        a, b, tt, c, d, vv = self._terms
        self.nnz = int(lib().oc_table_build(self._h, len(a), a, b, tt, len(c), c, d, vv))
        return self.nnz

    def build_table_cols(self, c0: int, c1: int, period: int = 1, block: int = 1) -> int:
        """The projection table of the columns c of [c0, c1) with (c - c0) mod period <
        block only (replacing any table built before); returns its number of entries."""
        a, b, tt, c, d, vv = self._terms
        self.nnz = int(lib().oc_table_build_strided(self._h, len(a), a, b, tt, len(c), c, d, vv, int(c0), int(c1),
                                                    int(period), int(block)))
        return self.nnz

    def table(self):
        """``build_matrixfree_projection_table`` output: (row, col) 1-based, vals, h_diag."""
        row = np.zeros(self.nnz, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int64)
        vals = np.zeros(self.nnz, dtype=np.complex128)
        hd = np.zeros(self.dim, dtype=np.complex128)
        lib().oc_table_copy(self._h, row, col, vals.view(np.float64), hd.view(np.float64))
        return row, col, vals, hd

    def apply_table(self, x, nthreads: int = 1) -> np.ndarray:
        """``_apply_table!`` with ``nthreads`` per-thread buffers."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.dim, dtype=np.complex128)
        lib().oc_apply_table(self._h, x.view(np.float64), y.view(np.float64), nthreads)
        return y

    def apply_table_timed(self, x, nthreads: int = 1):
        """``_apply_table!`` and the wall seconds of its parts: (y, {"zero", "diagonal",
        "entries", "reduction"})."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.dim, dtype=np.complex128)
        ph = np.zeros(4)
        lib().oc_apply_table_timed(self._h, x.view(np.float64), y.view(np.float64), nthreads, ph)
        return y, dict(zip(("zero", "diagonal", "entries", "reduction"), ph.tolist()))

    def apply_hamiltonian(self, x, nthreads: int = 1, c0: int = 0, c1: Optional[int] = None) -> np.ndarray:
        """On-the-fly ``apply_hamiltonian!`` over columns [c0, c1)."""
        a, b, tt, c, d, vv = self._terms
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.dim, dtype=np.complex128)
        lib().oc_apply_hamiltonian(self._h, len(a), a, b, tt, len(c), c, d, vv, x.view(np.float64),
                                   y.view(np.float64), nthreads, c0, self.dim if c1 is None else c1)
        return y

    def sparse(self):
        """The sector block as a scipy CSR matrix (duplicates summed, as ``sparse()`` does)."""
        import scipy.sparse as sp
        row, col, vals, hd = self.table()
        n = self.dim
        r = np.concatenate([row - 1, np.arange(n)])
        c = np.concatenate([col - 1, np.arange(n)])
        v = np.concatenate([vals, hd])
        return sp.csr_matrix((v, (r, c)), shape=(n, n))

    def dense(self) -> np.ndarray:
        return self.sparse().toarray()

    def density_correlation(self, c):
        """(<n_i n_j> as an n_site x n_site matrix, <n_i>) of the sector state c, as
        _build_density_correlation_matrix! forms them (static_structure_factor.jl:41-104)
        before the disconnected part is subtracted."""
        n = self.catalog.n_site
        c = np.ascontiguousarray(c, dtype=np.complex128)
        C = np.zeros(n * n)
        d = np.zeros(n)
        lib().oc_density_correlation(self._h, c.view(np.float64), C, d)
        return C.reshape(n, n), d

    def fock_expand(self, c, atol: float = 1e-12):
        """Rank-indexed (= ascending-mask) normalised Fock amplitudes of the sector state c
        (_expand_sector_state_to_fock_amplitudes, entanglement_spectrum.jl:46-79); None when
        the expanded norm is <= atol."""
        c = np.ascontiguousarray(c, dtype=np.complex128)
        out = np.zeros(self.catalog.n_total, dtype=np.complex128)
        cnt = lib().oc_fock_expand(self._h, c.view(np.float64), atol, out.view(np.float64))
        return None if cnt < 0 else out

    def position_operator(self, sec_from, phi):
        """(rows 1-based / 0, values) per column of sec_from for <chi_self| U_d |chi_from>
        (_position_operator_matrix, charge_pump.jl:47-91)."""
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        rows = np.zeros(sec_from.dim, dtype=np.int64)
        vals = np.zeros(sec_from.dim, dtype=np.complex128)
        lib().oc_position_operator(self._h, sec_from._h, phi.view(np.float64), rows, vals.view(np.float64))
        return rows, vals

    def one_body(self, c):
        """rho[i, j] = <psi| a_i^dag a_j |psi> of the sector state c
        (off_diagonal_long_range_order.jl:74-126)."""
        n = self.catalog.n_site
        c = np.ascontiguousarray(c, dtype=np.complex128)
        rho = np.zeros(n * n, dtype=np.complex128)
        lib().oc_one_body(self._h, c.view(np.float64), rho.view(np.float64))
        return rho.reshape(n, n)


def sector_for(model, group, chi) -> Sector:
    """Catalog + sector + projection table for a model (lattice builder output)."""
    from realspace_exactdiagonalization_amd.lattice import term_arrays, FERMIONIC
    cat = Catalog(model.n_site, model.n_filled, group, model.statistics == FERMIONIC)
    sec = Sector(cat, chi)
    sec.set_terms(*term_arrays(model))
    return sec


def lowest_eigenvalues(sec: Sector, nev: int, tol: float = 1e-13) -> np.ndarray:
    """Lowest ``nev`` eigenvalues: dense LAPACK for small blocks, ARPACK (scipy) otherwise."""
    n = sec.dim
    if n == 0:
        return np.zeros(0)
    if n <= 2000:
        H = sec.dense()
        return np.linalg.eigvalsh(0.5 * (H + H.conj().T))[:nev]
    from scipy.sparse.linalg import eigsh
    H = sec.sparse()
    rng = np.random.default_rng(1234)
    v0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    vals = eigsh(H, k=nev, which="SA", tol=tol, v0=v0, return_eigenvectors=False, ncv=max(2 * nev + 20, 40))
    return np.sort(vals.real)


def _threads(nthreads: int) -> int:
    return nthreads if nthreads > 0 else max(1, min(16, os.cpu_count() or 1))


def check_sector_basis(group, n_filled: int, fermionic: bool, chi, reps, stab, atol: float = 1e-12,
                       nthreads: int = 0):
    """Check a listed sector basis against the definition of ``build_symmetry_sector_basis``
    (``src/symmetry_resolved_ed.jl:534-556``) without building the catalog: every entry a
    canonical representative with ``n_filled`` bits (``:442-453``), strictly ascending,
    its stabilizer order, and its stabilizer phases equal to the characters
    (``_is_orbit_compatible`` ``:534-542``).  Returns (number of failing entries, first
    failing index or -1)."""
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    stab = np.ascontiguousarray(stab, dtype=np.int64)
    perm = np.ascontiguousarray(group.perm, dtype=np.int32)
    ph = _c128(group.phases)
    chi = _c128(chi)
    first = C.c_int64(-1)
    bad = lib().oc_check_sector_basis(group.n_site, n_filled, group.order, perm, ph, 1 if fermionic else 0, chi,
                                      atol, len(reps), reps, stab, _threads(nthreads), C.byref(first))
    return int(bad), int(first.value)


def sample_rows(model, group, chi, reps, stab, x, rows, nthreads: int = 0) -> np.ndarray:
    """``(H x)[rows]`` (0-based rows) of the sector with the listed basis, from the entries
    ``build_matrixfree_projection_table`` (``src/symmetry_resolved_ed.jl:680-728``) would
    write, found from each row without rep tables or Hermiticity (``oc_sample_rows``)."""
    from realspace_exactdiagonalization_amd.lattice import term_arrays, FERMIONIC
    a, b, t, di, dj, V = term_arrays(model)
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    stab = np.ascontiguousarray(stab, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    assert len(x) == len(reps) == len(stab) and (len(rows) == 0 or (rows.min() >= 0 and rows.max() < len(reps)))
    perm = np.ascontiguousarray(group.perm, dtype=np.int32)
    y = np.zeros(len(rows), dtype=np.complex128)
    lib().oc_sample_rows(group.n_site, group.order, perm, _c128(group.phases),
                         1 if model.statistics == FERMIONIC else 0, _c128(chi), len(reps), reps, stab,
                         len(a), a, b, _c128(t), len(di), di, dj, _c128(V), x.view(np.float64), len(rows), rows,
                         y.view(np.float64), _threads(nthreads))
    return y
