"""Reference-named host API over the C ABI of ``libsed.so``.

Mirrors the operator/plugin interface of the reference's matrix-free path
(``src/symmetry_resolved_ed.jl``) with the same names, argument meaning and error
behaviour (exceptions on failure):

    build_symmetry_orbit_catalog   :371-487
    build_symmetry_sector_basis    :544-556
    hamiltonian_linear_operator    :891-897   -> (MatrixFreeHamiltonian, n)
    MatrixFreeHamiltonian(x)       :880-884
    diagonalize_block_matrixfree   :998-1030
    build_ed_data                  :1460-1475
    ed_scan_at_irrep_matrixfree!   :1092-1128   (ed_scan_at_irrep_matrixfree)
    ed_scan!                       :1297-1356   (ed_scan; flux mode also as ed_scan_flux,
                                                 batched through the server)
    ed_scan_at_irrep_matrix!       :1061-1090   (ed_scan_at_irrep_matrix)
    resume_ed_scan!                :1266-1277   (resume_ed_scan; checkpoints in checkpoint.py)

Every call runs on the GPU through the HIP library; there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from . import checkpoint as _ckpt
from ._lib import PREC_FP32, PREC_FP32_REFINE, PREC_FP64, SolveStats, check, lib, ptr
from .lattice import (FERMIONIC, Irrep, Model, SymmetryGroup, identity_group, model_group, term_arrays,
                      translation_irreps, twisted)

__all__ = [
    "Context", "SymmetryOrbitCatalog", "SymmetrySectorBasis", "MatrixFreeHamiltonian",
    "build_symmetry_orbit_catalog", "build_symmetry_sector_basis", "hamiltonian_linear_operator",
    "diagonalize_block_matrixfree", "build_ed_data", "ed_scan_at_irrep_matrixfree", "ed_scan_flux",
    "SectorServer", "PREC_FP64", "PREC_FP32", "PREC_FP32_REFINE", "sector_density",
    "density_correlation_matrix", "off_diagonal_long_range_order", "build_ed_Hamiltonian_symmetry_block",
    "expand_sector_state_to_fock_amplitudes", "fock_amplitudes", "position_operator_matrix",
    "update_orbit_stabilizer_phases", "ed_scan", "ed_scan_at_irrep_matrix", "resume_ed_scan",
]

_default_ctx = None


class Context:
    """One GPU (device ordinal) with its HIP stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().sed_create(device, C.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().sed_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


class SymmetryOrbitCatalog:
    """Device-resident ``Symmetry_Orbit_Catalog`` (``src/symmetry_resolved_ed.jl:305-327``)."""

    def __init__(self, ctx: Context, n_site: int, n_filled: int, group: SymmetryGroup, statistics: int):
        self.ctx = ctx
        self.n_site = n_site
        self.n_filled = n_filled
        self.symmetry_group = group
        self.statistics = statistics
        h = C.c_void_p()
        perm = np.ascontiguousarray(group.perm, dtype=np.int32).reshape(-1)
        check(lib().sed_catalog_build(ctx._h, n_site, n_filled, statistics, group.order, group.identity_idx,
                                      perm, _c128(group.phases).reshape(-1), C.byref(h)))
        self._h = h
        nt, no = C.c_int64(), C.c_int64()
        check(lib().sed_catalog_info(h, C.byref(nt), C.byref(no)))
        self.n_total = nt.value
        self.n_orbits = no.value

    def tables(self):
        """(rep_rank_table, rep_sym_table, rep_amp_table) in the reference's layout."""
        rr = np.zeros(self.n_total, dtype=np.int32)
        rs = np.zeros(self.n_total, dtype=np.int32)
        ra = np.zeros(self.n_total, dtype=np.complex128)
        check(lib().sed_catalog_copy(self._h, ptr(rr), ptr(rs), ptr(ra), None, None))
        return rr, rs, ra

    def representatives(self) -> np.ndarray:
        """representative_mask_list (ascending)."""
        reps = np.zeros(self.n_orbits, dtype=np.uint64)
        check(lib().sed_catalog_copy(self._h, None, None, None, ptr(reps), None))
        return reps

    def orbits(self):
        """(representative_mask_list, stabilizer_order_list, stabilizer g indices, phases)."""
        nG = self.symmetry_group.order
        reps = np.zeros(self.n_orbits, dtype=np.uint64)
        so = np.zeros(self.n_orbits, dtype=np.int64)
        check(lib().sed_catalog_copy(self._h, None, None, None, ptr(reps), ptr(so)))
        gi = np.zeros((self.n_orbits, nG), dtype=np.int32)
        gp = np.zeros((self.n_orbits, nG), dtype=np.complex128)
        check(lib().sed_catalog_stabilizers(self._h, ptr(gi), ptr(gp)))
        return reps, so, gi, gp

    def update_phases(self, new_group: SymmetryGroup):
        """``update_orbit_stabilizer_phases!`` (``src/symmetry_resolved_ed.jl:338-369``): take the
        site phases of ``new_group`` (same permutations) without re-enumerating orbits."""
        if new_group.order != self.symmetry_group.order or not np.array_equal(
                np.asarray(new_group.perm), np.asarray(self.symmetry_group.perm)):
            raise ValueError("new_group must have the catalog's permutations")
        check(lib().sed_catalog_update_phases(self._h, _c128(new_group.phases).reshape(-1)))
        self.symmetry_group = new_group
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib().sed_catalog_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_symmetry_orbit_catalog(*, n_site: int, n_filled: int, symmetry_group: SymmetryGroup,
                                 particle_statistics: int, ctx: Optional[Context] = None) -> SymmetryOrbitCatalog:
    return SymmetryOrbitCatalog(ctx or default_context(), n_site, n_filled, symmetry_group, particle_statistics)


def update_orbit_stabilizer_phases(catalog: SymmetryOrbitCatalog, new_group: SymmetryGroup,
                                   particle_statistics: Optional[int] = None) -> SymmetryOrbitCatalog:
    """``update_orbit_stabilizer_phases!(catalog, new_group, particle_statistics)`` (``:338-369``)."""
    return catalog.update_phases(new_group)


class SymmetrySectorBasis:
    """``Symmetry_Sector_Basis`` of one 1D irrep (``src/symmetry_resolved_ed.jl:522-528``)."""

    def __init__(self, catalog: SymmetryOrbitCatalog, irrep: Irrep, atol: float = 1e-12):
        self.catalog = catalog
        self.irrep = irrep
        chi = _c128(irrep.values).reshape(-1)
        h = C.c_void_p()
        d = C.c_int64()
        check(lib().sed_sector_build(catalog._h, chi, atol, C.byref(h), C.byref(d)))
        self._h = h
        self.dim = d.value
        self._terms_set = False

    @property
    def representative_mask_list(self) -> np.ndarray:
        reps = np.zeros(self.dim, dtype=np.uint64)
        check(lib().sed_sector_basis(self._h, ptr(reps), None))
        return reps

    @property
    def stabilizer_order_list(self) -> np.ndarray:
        so = np.zeros(self.dim, dtype=np.int64)
        check(lib().sed_sector_basis(self._h, None, ptr(so)))
        return so

    def set_terms(self, bilinear_terms, density_terms):
        i_from = np.array([a for a, _, _ in bilinear_terms], dtype=np.int32)
        i_to = np.array([b for _, b, _ in bilinear_terms], dtype=np.int32)
        t = np.array([c for _, _, c in bilinear_terms], dtype=np.complex128)
        di = np.array([a for a, _, _ in density_terms], dtype=np.int32)
        dj = np.array([b for _, b, _ in density_terms], dtype=np.int32)
        v = np.array([c for _, _, c in density_terms], dtype=np.complex128)
        self.set_term_arrays(i_from, i_to, t, di, dj, v)

    def set_term_arrays(self, i_from, i_to, t, di, dj, v):
        i_from = np.ascontiguousarray(i_from, dtype=np.int32)
        i_to = np.ascontiguousarray(i_to, dtype=np.int32)
        di = np.ascontiguousarray(di, dtype=np.int32)
        dj = np.ascontiguousarray(dj, dtype=np.int32)
        check(lib().sed_sector_set_terms(self._h, len(i_from), i_from, i_to, _c128(t).reshape(-1) if len(t) else
                                         np.zeros(2), len(di), di, dj, _c128(v).reshape(-1) if len(v) else
                                         np.zeros(2)))
        self._terms_set = True

    def set_shape(self, precision: int, shape: int, xcd_chunk: int = 1):
        """Force H|psi> launch shape `shape` and XCD run length (shape -1: autotune)."""
        check(lib().sed_sector_set_shape(self._h, precision, shape, xcd_chunk))

    def shape(self, precision: int) -> dict:
        sh, ln, bt, ch, ns, nc, wg = (C.c_int() for _ in range(7))
        ms = np.zeros(64, dtype=np.float32)
        check(lib().sed_sector_shape(self._h, precision, C.byref(sh), C.byref(ln), C.byref(bt), C.byref(ch),
                                     C.byref(ns), C.byref(nc), ptr(ms), C.byref(wg)))
        return {"shape": sh.value, "lanes_per_row": ln.value, "batch": bt.value, "workgroup": wg.value,
                "xcd_chunk": ch.value,
                "tune_ms": [round(float(v), 4) for v in ms[:ns.value]],
                "tune_ms_xcd_chunk": [round(float(v), 4) for v in ms[ns.value:ns.value + nc.value]],
                "tune_ms_at_chunk": [round(float(v), 4) for v in ms[ns.value + nc.value:2 * ns.value + nc.value]]}

    def hop_counts(self) -> Tuple[int, int]:
        g, s = C.c_int64(), C.c_int64()
        check(lib().sed_sector_hop_counts(self._h, C.byref(g), C.byref(s)))
        return g.value, s.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sed_sector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_symmetry_sector_basis(catalog: SymmetryOrbitCatalog, irrep: Irrep, atol: float = 1e-12) -> SymmetrySectorBasis:
    return SymmetrySectorBasis(catalog, irrep, atol)


class MatrixFreeHamiltonian:
    """Callable sector operator ``H(x) -> y`` (``src/symmetry_resolved_ed.jl:767-791, 880-884``)."""

    def __init__(self, basis: SymmetrySectorBasis, bilinear_terms, density_terms, precision: int = PREC_FP64):
        self.basis = basis
        basis.set_terms(bilinear_terms, density_terms)
        self.n = basis.dim
        self.precision = precision

    def __call__(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape != (self.n,):
            raise ValueError(f"expected a vector of length {self.n}")
        y = np.zeros(self.n, dtype=np.complex128)
        check(lib().sed_matvec(self.basis._h, x.view(np.float64), y.view(np.float64), self.precision))
        return y

    def benchmark(self, precision: int = PREC_FP64, n_calls: int = 10) -> float:
        """Mean device time (ms) of the H|psi> kernel over n_calls back-to-back launches."""
        ms = C.c_double()
        check(lib().sed_matvec_benchmark(self.basis._h, precision, n_calls, C.byref(ms)))
        return ms.value


def hamiltonian_linear_operator(basis: SymmetrySectorBasis, bilinear_terms, density_terms, cmap=None):
    """``(H_op, n)`` as the reference returns it; ``cmap`` is accepted and unused (the
    representative tables live on the device)."""
    H = MatrixFreeHamiltonian(basis, bilinear_terms, density_terms)
    return H, H.n


def diagonalize_block_matrixfree(H_op: MatrixFreeHamiltonian, n: int, nev: int = 5, *, tol: float = 1e-10,
                                 seed: int = 1234, maxiter: int = 300, precision: int = PREC_FP64,
                                 return_vectors: bool = True, stats: Optional[dict] = None):
    """Lowest ``nev`` eigenpairs ``(vals, vecs)`` of a sector operator.

    Raises ``SedError`` (code ENOCONV) when fewer than ``nev`` pairs converge, as the
    reference errors (``:1023``)."""
    if n == 0:
        return np.zeros(0), np.zeros((0, 0), dtype=np.complex128)
    k = min(nev, n)
    vals = np.zeros(k, dtype=np.float64)
    vecs = np.zeros((k, n), dtype=np.complex128) if return_vectors else None
    st = SolveStats()
    rc = lib().sed_solve(H_op.basis._h, k, tol, seed, maxiter, precision, vals, ptr(vecs) if vecs is not None else None,
                         C.byref(st))
    if stats is not None:
        stats.update(st.as_dict())
    check(rc)
    return vals, (vecs.T.copy() if vecs is not None else None)


def sector_density(H_op: MatrixFreeHamiltonian, which: int = 0) -> np.ndarray:
    """Site densities of eigenvector ``which`` of the last solve on this sector."""
    d = np.zeros(H_op.basis.catalog.n_site, dtype=np.float64)
    check(lib().sed_sector_density(H_op.basis._h, which, d))
    return d


def build_ed_Hamiltonian_symmetry_block(basis: SymmetrySectorBasis, bilinear_terms, density_terms, cmap=None):
    """Matrix mode (``src/symmetry_resolved_ed.jl:903-963``): the sector block as a
    ``scipy.sparse.csc_matrix`` (0-based; duplicates summed, explicit zeros dropped),
    built on the GPU from the device tables.  ``cmap`` is accepted and unused."""
    import scipy.sparse as sp
    basis.set_terms(bilinear_terms, density_terms)
    nnz = C.c_int64()
    check(lib().sed_sector_matrix_csc(basis._h, C.byref(nnz), None, None, None))
    n = basis.dim
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowval = np.zeros(nnz.value, dtype=np.int64)
    nzval = np.zeros(nnz.value, dtype=np.complex128)
    check(lib().sed_sector_matrix_csc(basis._h, C.byref(nnz), colptr.ctypes.data, rowval.ctypes.data,
                                      nzval.ctypes.data))
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))


def expand_sector_state_to_fock_amplitudes(c, basis: SymmetrySectorBasis, atol: float = 1e-12) -> Dict[int, complex]:
    """``_expand_sector_state_to_fock_amplitudes`` (observables/entanglement_spectrum.jl:46-79):
    ``{mask: amplitude}`` of the sector state ``c`` in the fixed-N Fock basis, normalised,
    amplitudes with ``|a| <= atol`` dropped.  Expanded on the GPU."""
    masks, amps = fock_amplitudes(c, basis, atol)
    return {int(m): complex(a) for m, a in zip(masks, amps)}


def fock_amplitudes(c, basis: SymmetrySectorBasis, atol: float = 1e-12):
    """Arrays ``(masks, amplitudes)`` (ascending masks) of the same expansion."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    if c.shape != (basis.dim,):
        raise ValueError(f"expected a vector of length {basis.dim}")
    n = C.c_int64()
    check(lib().sed_sector_fock_expansion(basis._h, c.ctypes.data, atol, 0, C.byref(n), None, None))
    masks = np.zeros(n.value, dtype=np.uint64)
    amps = np.zeros(n.value, dtype=np.complex128)
    check(lib().sed_sector_fock_expansion(basis._h, c.ctypes.data, atol, n.value, C.byref(n), masks.ctypes.data,
                                          amps.ctypes.data))
    return masks, amps


def position_operator_matrix(basis_to: SymmetrySectorBasis, basis_from: SymmetrySectorBasis, site_phases):
    """``_position_operator_matrix`` (observables/charge_pump.jl:47-91): the projected
    position operator <chi_to| U_d |chi_from> as a ``scipy.sparse.csc_matrix`` (entries
    with ``|elem| > 1e-13``), formed on the GPU."""
    import scipy.sparse as sp
    phi = np.ascontiguousarray(site_phases, dtype=np.complex128)
    if phi.shape != (basis_from.catalog.n_site,):
        raise ValueError("site_phases must have n_site entries")
    nf = basis_from.dim
    rows = np.zeros(nf, dtype=np.int64)
    vals = np.zeros(nf, dtype=np.complex128)
    check(lib().sed_sector_position_operator(basis_to._h, basis_from._h, phi.ctypes.data, rows.ctypes.data,
                                             vals.ctypes.data))
    keep = (rows > 0) & (np.abs(vals) > 1e-13)
    cols = np.nonzero(keep)[0]
    return sp.csc_matrix((vals[keep], (rows[keep] - 1, cols)), shape=(basis_to.dim, nf))


def density_correlation_matrix(H_op: MatrixFreeHamiltonian, which: int = 0, subtract_disconnected: bool = True):
    """``(C, density)`` of eigenvector ``which`` of the last solve: C[i, j] = <n_i n_j>
    minus <n_i><n_j> when ``subtract_disconnected`` (0-based sites), as
    ``_build_density_correlation_matrix!`` (observables/static_structure_factor.jl:41-113)
    forms them with every site in both flavours."""
    n = H_op.basis.catalog.n_site
    nn = np.zeros(n * n, dtype=np.float64)
    d = np.zeros(n, dtype=np.float64)
    check(lib().sed_sector_density_correlation(H_op.basis._h, which, nn, d))
    C = nn.reshape(n, n)
    if subtract_disconnected:
        C = C - np.outer(d, d)
    return C, d


def off_diagonal_long_range_order(H_op: MatrixFreeHamiltonian, which: int = 0) -> np.ndarray:
    """One-body density matrix rho[i, j] = <a_i^dag a_j> (0-based sites) of eigenvector
    ``which`` of the last solve (observables/off_diagonal_long_range_order.jl:55-126)."""
    n = H_op.basis.catalog.n_site
    rho = np.zeros(n * n, dtype=np.complex128)
    check(lib().sed_sector_one_body(H_op.basis._h, which, rho.view(np.float64)))
    return rho.reshape(n, n)


@dataclass
class EDData:
    """``Symmetry_Resolved_ED_Data`` (``src/symmetry_resolved_ed.jl:1051-1060``).  The
    catalog is None only for scans driven by a caller-supplied sector solver."""
    model: Model
    n_filled: int
    symmetry_group: SymmetryGroup
    irrep_list: List[Irrep]
    orbit_catalog: Optional[SymmetryOrbitCatalog]
    sector_dims: List[int]
    ed_scan_res: Dict[int, Tuple[np.ndarray, np.ndarray]] = field(default_factory=dict)
    checkpoint_path: Optional[str] = None

    @property
    def filling_fraction(self) -> Fraction:
        """Particles per vertex, n_filled / n_site (the reference's filling_fraction)."""
        return Fraction(self.n_filled, self.model.n_site)


def build_ed_data(model: Model, symmetry_group: Optional[SymmetryGroup] = None,
                  ctx: Optional[Context] = None, build_catalog: bool = True) -> EDData:
    G = symmetry_group if symmetry_group is not None else model_group(model)
    irreps = translation_irreps(G, model.lattice)
    cat = None
    if build_catalog:
        cat = build_symmetry_orbit_catalog(n_site=model.n_site, n_filled=model.n_filled, symmetry_group=G,
                                           particle_statistics=model.statistics, ctx=ctx)
    return EDData(model, model.n_filled, G, irreps, cat, [0] * len(irreps))


def ed_scan_at_irrep_matrixfree(irrep_label, ed_data: EDData, nev: int = 5, **solve_kw):
    """Per-sector driver (``src/symmetry_resolved_ed.jl:1092-1128``)."""
    idx = next(i for i, ir in enumerate(ed_data.irrep_list) if tuple(ir.label) == tuple(irrep_label))
    basis = build_symmetry_sector_basis(ed_data.orbit_catalog, ed_data.irrep_list[idx])
    ed_data.sector_dims[idx] = basis.dim
    if basis.dim == 0:
        ed_data.ed_scan_res[idx] = (np.zeros(0), np.zeros((0, 0), dtype=np.complex128))
        return ed_data.ed_scan_res[idx]
    H, n = hamiltonian_linear_operator(basis, ed_data.model.bilinear_terms, ed_data.model.density_terms)
    vals, vecs = diagonalize_block_matrixfree(H, n, nev=nev, **solve_kw)
    ed_data.ed_scan_res[idx] = (vals, vecs)
    return vals, vecs


def ed_scan_at_irrep_matrix(irrep_label, ed_data: EDData, nev: int = 5, **solve_kw):
    """Matrix-mode per-sector driver (``ed_scan_at_irrep_matrix!``, ``:1061-1090``).  The
    reference builds the sector block as a sparse matrix and hands it to Arpack (dense
    ``eigen`` up to dimension 500, ``:978-995``); the block is the operator the matrix-free
    path applies, so its eigenpairs come from the same GPU solver here (dense up to 500,
    Lanczos above; Arpack itself is not reproduced).  ``build_ed_Hamiltonian_symmetry_block``
    returns the block itself."""
    return ed_scan_at_irrep_matrixfree(irrep_label, ed_data, nev=nev, **solve_kw)


def _state(ed_data: EDData) -> _ckpt.ScanState:
    return _ckpt.ScanState(ed_data.model, ed_data.n_filled, ed_data.filling_fraction, ed_data.symmetry_group.name,
                           [tuple(ir.label) for ir in ed_data.irrep_list], ed_data.sector_dims, ed_data.ed_scan_res)


def _scan_sectors(ed_data: EDData, nev: int, mode: str, scanned_sectors, checkpoint_path: Optional[str] = None,
                  overwrite: bool = False, _checkpoint_loaded: bool = False, solve=None, **solve_kw) -> EDData:
    """``_ed_scan_sectors!`` (``:1199-1256``).  With ``checkpoint_path``: unless
    ``overwrite``, an existing checkpoint is loaded first (it must be of the same filling,
    sector labels and Hamiltonian, or the scan stops with an error) and its finished
    sectors are skipped; after each solved sector the checkpoint is rewritten atomically.
    ``solve(ed_data, irrep_index, nev) -> (vals, vecs, dim)`` replaces the GPU sector solve
    (for drivers and tests that bring their own solver)."""
    if mode not in ("matrix", "matrixfree"):
        raise ValueError(f"Unknown ED scan mode: {mode}. Use 'matrix' or 'matrixfree'.")
    if overwrite:
        ed_data.ed_scan_res.clear()
    elif checkpoint_path is not None and os.path.isfile(checkpoint_path) and not _checkpoint_loaded:
        st = _ckpt.load_checkpoint(checkpoint_path)
        if (st.n_filled != ed_data.n_filled or st.filling_fraction != ed_data.filling_fraction
                or st.irrep_labels != [tuple(ir.label) for ir in ed_data.irrep_list]
                or not _ckpt.same_flux_problem(st.model, ed_data.model)):
            raise RuntimeError("ED checkpoint is incompatible with the requested filling or symmetry-sector labels.")
        ed_data.ed_scan_res = st.ed_scan_res
        ed_data.sector_dims = list(st.sector_dims)
    if checkpoint_path is not None:
        ed_data.checkpoint_path = checkpoint_path
    wanted = None if scanned_sectors is None else {tuple(l) for l in scanned_sectors}
    for idx, ir in enumerate(ed_data.irrep_list):
        if wanted is not None and tuple(ir.label) not in wanted:
            continue
        if idx in ed_data.ed_scan_res:
            dim = ed_data.sector_dims[idx]
            if dim == 0 or len(ed_data.ed_scan_res[idx][0]) >= min(nev, dim):
                continue
        if solve is not None:
            vals, vecs, dim = solve(ed_data, idx, nev)
            ed_data.sector_dims[idx] = int(dim)
            ed_data.ed_scan_res[idx] = (np.asarray(vals), np.asarray(vecs))
        elif mode == "matrixfree":
            ed_scan_at_irrep_matrixfree(ir.label, ed_data, nev=nev, **solve_kw)
        else:
            ed_scan_at_irrep_matrix(ir.label, ed_data, nev=nev, **solve_kw)
        if checkpoint_path is not None:
            _ckpt.save_checkpoint(_state(ed_data), checkpoint_path)
    return ed_data


def resume_ed_scan(checkpoint_path: str, nev: int = 5, mode: str = "matrix", scanned_sectors=None,
                   overwrite: bool = False, ctx: Optional[Context] = None, solve=None, **solve_kw) -> EDData:
    """``resume_ed_scan!`` (``:1266-1277``): the scan state is read from the checkpoint
    (model, filling, group, sectors finished so far), the orbit catalog rebuilt, and the
    missing or undersized sectors solved with per-sector checkpoints."""
    if not os.path.isfile(checkpoint_path):
        raise RuntimeError(f"ED checkpoint does not exist: {checkpoint_path}")
    st = _ckpt.load_checkpoint(checkpoint_path)
    m = st.model
    m.n_filled = st.n_filled
    G = identity_group(m.n_site) if st.group_name == "identity" else model_group(m, m.theta)
    d = build_ed_data(m, symmetry_group=G, ctx=ctx, build_catalog=solve is None)
    if [tuple(ir.label) for ir in d.irrep_list] != st.irrep_labels:
        raise RuntimeError("ED checkpoint is incompatible with the requested filling or symmetry-sector labels.")
    d.ed_scan_res = st.ed_scan_res
    d.sector_dims = list(st.sector_dims)
    return _scan_sectors(d, nev, mode, scanned_sectors, checkpoint_path=checkpoint_path, overwrite=overwrite,
                         _checkpoint_loaded=True, solve=solve, **solve_kw)


def ed_scan(ed_data: EDData, nev: int = 5, mode: str = "matrix", scanned_sectors=None, *, flux_direction: int = 1,
            twisted_phases_over_2pi_list: Optional[Sequence[float]] = None, ctx: Optional[Context] = None,
            checkpoint_path: Optional[str] = None, checkpoint_dir: Optional[str] = None, overwrite: bool = False,
            solve=None, **solve_kw):
    """``ed_scan!`` (``:1297-1356``, sector loop ``_ed_scan_sectors!`` ``:1199-1256``).

    Without a flux list: every momentum sector of ``ed_data`` (or those whose labels are in
    ``scanned_sectors``) is solved for ``nev`` eigenpairs into ``ed_data.ed_scan_res``;
    sectors already holding enough pairs are skipped, as a resumed scan skips them.
    Returns ``ed_data``.  With ``twisted_phases_over_2pi_list``: for each twist theta along
    ``flux_direction`` (1-based) the model is re-twisted, the flux-aware translation group
    (or the identity group when ``ed_data`` uses it) and a fresh orbit catalog built, and the
    sectors scanned; returns the list of per-twist ``EDData``.
    ``mode`` is ``'matrix'`` (the reference's default) or ``'matrixfree'``; both solve on the
    GPU (``ed_scan_at_irrep_matrix``).

    Checkpoints (``:1137-1196``): a plain scan with ``checkpoint_path`` loads and extends that
    file; a flux scan with ``checkpoint_dir`` writes one file per twist, named by
    ``ed_scan_checkpoint_filename`` (the model's parameter digest), and a twist whose file
    exists is resumed (after checking it holds the same Hamiltonian) instead of restarted.
    ``overwrite`` discards stored results.  Each returned ``EDData`` names its file in
    ``checkpoint_path``."""
    if twisted_phases_over_2pi_list is None:
        return _scan_sectors(ed_data, nev, mode, scanned_sectors, checkpoint_path=checkpoint_path,
                             overwrite=overwrite, solve=solve, **solve_kw)
    model = ed_data.model
    if not 1 <= flux_direction <= 2:
        raise ValueError("flux_direction must be 1 or 2")
    out = []
    if ctx is None and solve is None:
        ctx = ed_data.orbit_catalog.ctx
    identity = ed_data.symmetry_group.name == "identity"
    for t in twisted_phases_over_2pi_list:
        flux = [0.0, 0.0]
        flux[flux_direction - 1] = float(t)
        m = twisted(model, flux)
        m.n_filled = ed_data.n_filled
        path = None
        if checkpoint_dir is not None:
            path = os.path.join(checkpoint_dir, _ckpt.ed_scan_checkpoint_filename(
                m, flux, Fraction(ed_data.n_filled, m.n_site), identity))
            if os.path.isfile(path) and not overwrite:
                cached = _ckpt.load_checkpoint(path)
                if not _ckpt.same_flux_problem(cached.model, m):
                    raise RuntimeError("Flux checkpoint belongs to a different Hamiltonian.")
                out.append(resume_ed_scan(path, nev, mode, scanned_sectors, ctx=ctx, solve=solve, **solve_kw))
                continue
        G = identity_group(m.n_site) if identity else model_group(m, flux)
        d = build_ed_data(m, symmetry_group=G, ctx=ctx, build_catalog=solve is None)
        out.append(_scan_sectors(d, nev, mode, scanned_sectors, checkpoint_path=path, overwrite=overwrite,
                                 solve=solve, **solve_kw))
    return out


class SectorServer:
    """Continuous-batching service of sector solves (``sed_server_*``)."""

    def __init__(self, ctx: Optional[Context] = None, hbm_budget_bytes: int = 0, max_slots: int = 8):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().sed_server_create(self.ctx._h, int(hbm_budget_bytes), int(max_slots), C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, basis: SymmetrySectorBasis, nev: int, tol: float = 1e-10, seed: int = 1234,
               maxiter: int = 300, precision: int = PREC_FP64) -> int:
        rid = C.c_int64()
        check(lib().sed_server_submit(self._h, basis._h, nev, tol, seed, maxiter, precision, C.byref(rid)))
        self._keep[rid.value] = (basis, nev)
        return rid.value

    def run(self, max_steps: int = 0) -> int:
        nd = C.c_int64()
        check(lib().sed_server_run(self._h, max_steps, C.byref(nd)))
        return nd.value

    def result(self, rid: int):
        """(status, evals, stats dict, latency seconds); status 0 done, 1 pending, 2 running, <0 error."""
        basis, nev = self._keep[rid]
        st = C.c_int()
        ev = np.full(nev, np.nan)
        stats = SolveStats()
        lat = C.c_double(float("nan"))
        check(lib().sed_server_result(self._h, rid, C.byref(st), ev, C.byref(stats), C.byref(lat)))
        return st.value, ev, stats.as_dict(), lat.value

    def memory(self) -> dict:
        """Device-memory accounting: budget, bytes taken at creation (packed engine), bytes
        charged now and the largest charge so far."""
        b, f, u, p = (C.c_int64() for _ in range(4))
        check(lib().sed_server_memory(self._h, C.byref(b), C.byref(f), C.byref(u), C.byref(p)))
        return {"budget": b.value, "fixed": f.value, "in_use": u.value, "peak": p.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().sed_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ed_scan_flux(model: Model, thetas: Sequence[Tuple[float, float]], nev: int = 3, *,
                 sectors: Optional[Sequence[Tuple[int, int]]] = None, precision: int = PREC_FP64,
                 tol: float = 1e-10, seed: int = 1234, max_slots: int = 8, ctx: Optional[Context] = None,
                 reuse_catalog: bool = True):
    """Flux scan (``ed_scan!`` with ``twisted_phases_over_2pi_list``, ``:1320-1355``): every
    (twist, momentum) sector solve is submitted to one continuous-batching server.

    Returns {(theta, k_label): evals}.  The orbit partition does not depend on the twist:
    with ``reuse_catalog`` one catalog is built and its phases refreshed per twist
    (``update_orbit_stabilizer_phases!``, ``:338-369``; the reference's own loop rebuilds it,
    ``:1348``); every twist's sectors keep the phases they were built with."""
    ctx = ctx or default_context()
    srv = SectorServer(ctx, max_slots=max_slots)
    keep = []
    ids = {}
    cat = None
    for th in thetas:
        m = twisted(model, th)
        G = model_group(m, th)
        irreps = translation_irreps(G, m.lattice)
        if cat is not None and reuse_catalog:
            cat.update_phases(G)
        else:
            cat = build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                               particle_statistics=m.statistics, ctx=ctx)
        for ir in irreps:
            if sectors is not None and tuple(ir.label) not in set(map(tuple, sectors)):
                continue
            b = build_symmetry_sector_basis(cat, ir)
            if b.dim == 0:
                continue
            b.set_terms(m.bilinear_terms, m.density_terms)
            keep.append(b)
            ids[(tuple(th), tuple(ir.label))] = srv.submit(b, min(nev, b.dim), tol, seed, 300, precision)
        if cat not in keep:
            keep.append(cat)
    srv.run(0)
    out = {}
    for key, rid in ids.items():
        st, ev, stats, lat = srv.result(rid)
        if st != 0:
            raise _lib.SedError(st, f"sector {key} failed")
        out[key] = ev
    srv.close()
    return out
