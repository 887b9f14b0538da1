"""Flux-insertion observables over the sector solver: spectral flow, the many-body
charge pump and the flux-torus Chern number.

Host drivers of the reference's ``observables/spectrum_flow.jl``,
``observables/charge_pump.jl`` and ``observables/many_body_chern_number.jl``, under the
same names, argument meanings and error behaviour.  Every sector solve, every Fock-space
expansion and every projected position operator runs on the GPU through the HIP
library (``ed``); this module only selects the low-energy manifold at each flux and does
the small dense algebra on top (an n x n projected operator, n x n overlap
determinants, phase unwrapping), n being the manifold size.

Flux arguments are in units of 2 pi; the twisted model and the flux-aware translation
group come from ``lattice.twisted`` / ``lattice.model_group``.  Checkpoint files and
figures of the reference drivers are not written.
"""
from __future__ import annotations

import itertools
import math
from fractions import Fraction
from types import SimpleNamespace
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import ed
from .lattice import Lattice, Model, identity_group, model_group, translation_irreps, twisted

__all__ = [
    "many_body_position_phases", "flux_spectrum_flow", "flux_charge_pump", "many_body_chern_number",
    "tracked_phase_branches", "unit_det_link", "plaquette_phases",
]


# --------------------------------------------------------------------------
# Small dense helpers
# --------------------------------------------------------------------------
def many_body_position_phases(lattice: Lattice, direction: int, include_sublattice: bool = True) -> np.ndarray:
    """Site factors of Resta's position operator U = exp(2 pi i / L_d sum_j x_{j,d} n_j)
    (``observables/charge_pump.jl:13-29``): phase_j = cis(2 pi x_{j,d} / L_d) with x the
    cell coordinate along ``direction`` (1-based) plus, optionally, the sublattice offset."""
    if not 1 <= direction <= lattice.dim:
        raise ValueError(f"direction must be in 1:{lattice.dim}.")
    d = direction - 1
    Ld = float(lattice.sample_size[d])
    out = np.empty(lattice.n_site, dtype=np.complex128)
    for i, (cell, isub) in enumerate(lattice.site_list):
        x = float(cell[d])
        if include_sublattice:
            x += lattice.sub_crys[isub - 1][d]
        out[i] = complex(math.cos(2.0 * math.pi * x / Ld), math.sin(2.0 * math.pi * x / Ld))
    return out


def _best_phase_permutation(raw: Sequence[float], prev: Sequence[float]) -> Tuple[int, ...]:
    """Assignment of the new phases to the previous branches with the least summed squared
    jump modulo 1 (``observables/charge_pump.jl:134-167``); the first permutation in
    lexicographic order wins ties."""
    n = len(prev)
    best, best_cost = tuple(range(n)), math.inf
    for p in itertools.permutations(range(n)):
        cost = 0.0
        for b in range(n):
            dlt = raw[p[b]] - prev[b]
            dlt -= round(dlt)
            cost += dlt * dlt
            if cost >= best_cost:
                break
        else:
            best, best_cost = p, cost
    return best


def tracked_phase_branches(raw_phases: np.ndarray) -> np.ndarray:
    """Unwrap phase branches (rows = flux points, columns = branches, phases in units of
    2 pi) by best matching between neighbouring fluxes (``charge_pump.jl:114-131``)."""
    raw = np.asarray(raw_phases, dtype=np.float64)
    n_theta, nb = raw.shape
    out = np.empty_like(raw)
    out[0] = raw[0, np.argsort(raw[0], kind="stable")]
    prev = out[0].copy()
    for it in range(1, n_theta):
        perm = _best_phase_permutation(raw[it], prev)
        for b in range(nb):
            phi = raw[it, perm[b]]
            out[it, b] = phi + round(prev[b] - phi)
        prev = out[it].copy()
    return out


def _fock_overlap(a, b) -> complex:
    """<a|b> of two Fock expansions given as (ascending masks, amplitudes)."""
    ma, aa = a
    mb, ab = b
    _, ia, ib = np.intersect1d(ma, mb, assume_unique=True, return_indices=True)
    return complex(np.vdot(aa[ia], ab[ib]))


def unit_det_link(states_a, states_b, det_atol: float = 1e-10) -> Tuple[complex, float]:
    """U(1) link det(M)/|det(M)| of the overlap matrix M_ij = <a_i|b_j> between two
    multiplets, and |det M| (``many_body_chern_number.jl:51-81``).  Raises when
    |det M| <= det_atol."""
    n = len(states_a)
    if len(states_b) != n:
        raise RuntimeError("Chern link compares subspaces of different dimensions.")
    M = np.array([[_fock_overlap(states_a[i], states_b[j]) for j in range(n)] for i in range(n)])
    z = complex(np.linalg.det(M))
    if abs(z) <= det_atol:
        raise RuntimeError(f"Singular Berry link (|det|={abs(z)}); refine the grid or select an isolated manifold.")
    return z / abs(z), abs(z)


def plaquette_phases(Ux: np.ndarray, Uy: np.ndarray) -> np.ndarray:
    """Lattice Berry curvature arg(Ux(k) Uy(k+x) conj(Ux(k+y)) conj(Uy(k))) on the
    periodic flux grid (Fukui-Hatsugai-Suzuki; ``many_body_chern_number.jl:232-238``)."""
    nx, ny = Ux.shape
    F = np.empty((nx, ny))
    for ix in range(nx):
        for iy in range(ny):
            ixp, iyp = (ix + 1) % nx, (iy + 1) % ny
            F[ix, iy] = np.angle(Ux[ix, iy] * Uy[ixp, iy] * np.conj(Ux[ix, iyp]) * np.conj(Uy[ix, iy]))
    return F


# --------------------------------------------------------------------------
# Sector solves at one flux point
# --------------------------------------------------------------------------
def _n_filled(model: Model, filling_fraction) -> int:
    if filling_fraction is None:
        return model.n_filled
    nf = Fraction(filling_fraction).limit_denominator(1 << 20) * model.n_site
    if nf.denominator != 1:
        raise ValueError(f"filling_fraction {filling_fraction} does not give an integer particle number.")
    return int(nf)


def _labels(model: Model, sector_labels):
    L1, L2 = model.lattice.sample_size
    if isinstance(sector_labels, str):
        if sector_labels == "all":
            return [(i, j) for i in range(L1) for j in range(L2)], False
        if sector_labels == "identity":
            return ["identity"], True
        raise ValueError(f"unknown sector_labels {sector_labels!r}")
    if isinstance(sector_labels, tuple) and len(sector_labels) == 2 and all(isinstance(v, int) for v in sector_labels):
        return [tuple(sector_labels)], False
    return [tuple(int(v) for v in lab) for lab in sector_labels], False


class _FluxPoint:
    """Every momentum sector solved at one flux: ``sectors[i] = (label, basis, vals, vecs)``."""

    def __init__(self, model: Model, flux: Sequence[float], n_filled: int, nev: int, is_identity: bool,
                 ctx: ed.Context, catalog: Optional[ed.SymmetryOrbitCatalog] = None):
        flux = (float(flux[0]), float(flux[1]))
        self.model = twisted(model, flux)
        G = identity_group(model.n_site) if is_identity else model_group(self.model, flux)
        if catalog is not None:
            catalog.update_phases(G)
            self.catalog = catalog
        else:
            self.catalog = ed.build_symmetry_orbit_catalog(n_site=model.n_site, n_filled=n_filled, symmetry_group=G,
                                                           particle_statistics=model.statistics, ctx=ctx)
        self.sectors = []
        for ir in translation_irreps(G, model.lattice):
            label = "identity" if is_identity else tuple(ir.label)
            basis = ed.build_symmetry_sector_basis(self.catalog, ir)
            if basis.dim == 0:
                basis.close()
                self.sectors.append((label, None, np.zeros(0), np.zeros((0, 0), dtype=np.complex128)))
                continue
            H, n = ed.hamiltonian_linear_operator(basis, self.model.bilinear_terms, self.model.density_terms)
            vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=nev)
            self.sectors.append((label, basis, vals, vecs))

    def select(self, nstates: int, what: str):
        """Global lowest ``nstates`` states over all sectors (stable order: sector, then
        level) and the gap above them; raises when no state lies above or the gap closes."""
        cand = sorted(((float(v[l]), i, l) for i, (_, _, v, _) in enumerate(self.sectors) for l in range(len(v))),
                      key=lambda c: c[0])
        if len(cand) <= nstates:
            raise RuntimeError(f"Need a state above the {what} to measure its gap.")
        gap = cand[nstates][0] - cand[nstates - 1][0]
        return gap, [(i, l) for _, i, l in cand[:nstates]]

    def close(self):
        for _, basis, _, _ in self.sectors:
            if basis is not None:
                basis.close()
        self.sectors = []


# --------------------------------------------------------------------------
# Drivers
# --------------------------------------------------------------------------
def flux_spectrum_flow(model: Model, sector_labels="all", *, filling_fraction=None, flux_direction: int = 1,
                       twisted_phases_over_2pi_list: Optional[Sequence[float]] = None, nev: int = 3,
                       ctx: Optional[ed.Context] = None, max_slots: int = 8):
    """Spectral flow E(theta) (``observables/spectrum_flow.jl:27-104``): every momentum
    sector solved at every flux of the list (one continuous-batching server run; one
    orbit catalog refreshed per flux).  Returns a namespace with ``energies[itheta,
    isector, level]`` (NaN where a sector has fewer levels) for ``sector_labels``, and
    ``global_energies[itheta, :]`` = the lowest ``nev`` energies over all sectors."""
    fluxes = list(np.linspace(0.0, 1.0, 9)) if twisted_phases_over_2pi_list is None else \
        [float(t) for t in twisted_phases_over_2pi_list]
    labels, is_identity = _labels(model, sector_labels)
    if not 1 <= flux_direction <= model.lattice.dim:
        raise ValueError(f"flux_direction must be in 1:{model.lattice.dim}.")
    m = twisted(model, (0.0, 0.0))
    m.n_filled = _n_filled(model, filling_fraction)
    thetas = []
    for t in fluxes:
        f = [0.0, 0.0]
        f[flux_direction - 1] = t
        thetas.append(tuple(f))
    if is_identity:
        res = _identity_scan(m, thetas, nev, ctx or ed.default_context())
    else:
        res = ed.ed_scan_flux(m, thetas, nev=nev, ctx=ctx, max_slots=max_slots)
    energies = np.full((len(fluxes), len(labels), nev), np.nan)
    glob = []
    for it, th in enumerate(thetas):
        allv = np.sort(np.concatenate([v for (t, _), v in res.items() if t == th] or [np.zeros(0)]))
        glob.append(allv[:min(nev, len(allv))])
        for il, lab in enumerate(labels):
            v = res.get((th, (0, 0) if is_identity else lab))
            if v is not None:
                nv = min(len(v), nev)
                energies[it, il, :nv] = v[:nv]
    width = min(len(g) for g in glob) if glob else 0
    return SimpleNamespace(twisted_phases_over_2pi_list=fluxes, energies=energies, sector_labels=labels,
                           flux_direction=flux_direction, nev=nev, is_identity=is_identity,
                           global_energies=np.array([g[:width] for g in glob]))


def _identity_scan(m: Model, thetas, nev: int, ctx: ed.Context):
    out = {}
    for th in thetas:
        pt = _FluxPoint(m, th, m.n_filled, nev, True, ctx)
        for _, _, v, _ in pt.sectors:
            out[(th, (0, 0))] = v
        pt.close()
        pt.catalog.close()
    return out


def _manifold_size(labels, is_identity, nev_per_sector, manifold_states, manifold_size) -> int:
    """State count of the manifold as ``_normalize_manifold_state_specs``
    (``charge_pump.jl:169-209``) infers it, with its validation errors."""
    if nev_per_sector <= 0:
        raise ValueError("nev_per_sector must be positive.")
    if manifold_states is None:
        n = len(labels) * nev_per_sector
    else:
        if is_identity:
            raise ValueError("Explicit manifold_states are only supported for symmetry-sector ED.")
        seen = set()
        for st in manifold_states:
            sector = tuple(int(v) for v in (st["sector"] if isinstance(st, dict) else st[0]))
            level = int(st["level"] if isinstance(st, dict) else st[1])
            if level <= 0:
                raise ValueError(f"Manifold levels must be positive; got {level} in sector {sector}.")
            if sector not in labels:
                raise ValueError(f"Manifold state ({sector}, level {level}) is absent from sector_labels={labels}.")
            if (sector, level) in seen:
                raise ValueError("manifold_states contains duplicate (sector, level) entries.")
            seen.add((sector, level))
        if not seen:
            raise ValueError("manifold_states must not be empty.")
        n = len(seen)
    n = n if manifold_size is None else int(manifold_size)
    if n <= 0:
        raise ValueError("manifold_size must be positive.")
    return n


def flux_charge_pump(model: Model, sector_labels, *, filling_fraction=None, flux_direction: int = 1,
                     polarization_direction: Optional[int] = None,
                     twisted_phases_over_2pi_list: Optional[Sequence[float]] = None, nev_per_sector: int = 1,
                     manifold_states=None, manifold_size: Optional[int] = None, polarization_atol: float = 1e-10,
                     include_sublattice: bool = True, ctx: Optional[ed.Context] = None):
    """Many-body charge pump (``observables/charge_pump.jl:211-388``).

    At each flux along ``flux_direction`` every momentum sector is solved for
    ``nstates + 1`` levels, the globally lowest ``nstates`` states are selected (error when
    the gap above them is <= 1e-9), Resta's position operator along
    ``polarization_direction`` is projected into them (GPU position-operator blocks
    between the sectors involved), its eigenvalues sorted by phase; the phase branches are
    unwrapped along the flux path.  ``pumped_charges`` = branch polarization at the last
    flux minus the first."""
    if isinstance(sector_labels, str) and sector_labels == "all" and manifold_size is None:
        raise ValueError("Specify manifold_size with :all.")
    labels, is_identity = _labels(model, sector_labels)
    dim = model.lattice.dim
    if polarization_direction is None:
        polarization_direction = flux_direction if dim == 1 else (flux_direction % dim) + 1
    if not 1 <= flux_direction <= dim:
        raise ValueError(f"flux_direction must be in 1:{dim}.")
    if not 1 <= polarization_direction <= dim:
        raise ValueError(f"polarization_direction must be in 1:{dim}.")
    nstates = _manifold_size(labels, is_identity, nev_per_sector, manifold_states, manifold_size)
    fluxes = list(np.linspace(0.0, 1.0, 9)) if twisted_phases_over_2pi_list is None else \
        [float(t) for t in twisted_phases_over_2pi_list]
    ctx = ctx or ed.default_context()
    n_filled = _n_filled(model, filling_fraction)
    all_labels = ["identity"] if is_identity else \
        [(i, j) for i in range(model.lattice.sample_size[0]) for j in range(model.lattice.sample_size[1])]
    energies = np.full((len(fluxes), len(all_labels), nstates), np.nan)
    polarization_eigenvalues = np.empty((len(fluxes), nstates), dtype=np.complex128)
    gaps, min_svs, selected_states = [], [], []
    site_phases = many_body_position_phases(model.lattice, polarization_direction, include_sublattice)
    catalog = None
    specs = []
    try:
        for it, t in enumerate(fluxes):
            f = [0.0, 0.0]
            f[flux_direction - 1] = t
            pt = _FluxPoint(model, f, n_filled, nstates + 1, is_identity, ctx, catalog)
            catalog = pt.catalog
            try:
                gap, sel = pt.select(nstates, "manifold")
                if not gap > 1e-9:
                    raise RuntimeError(f"Pump manifold is not isolated at flux index {it + 1} (gap={gap}).")
                gaps.append(gap)
                specs = [{"sector": pt.sectors[i][0], "level": l + 1} for i, l in sel]
                selected_states.append(specs)
                for i in sorted({i for i, _ in sel}):
                    nv = max(l for j, l in sel if j == i) + 1
                    energies[it, all_labels.index(pt.sectors[i][0]), :nv] = pt.sectors[i][2][:nv]
                blocks = {}
                P = np.zeros((nstates, nstates), dtype=np.complex128)
                for r, (ir_, lr) in enumerate(sel):
                    for c, (ic, lc) in enumerate(sel):
                        key = (ir_, ic)
                        if key not in blocks:
                            blocks[key] = ed.position_operator_matrix(pt.sectors[ir_][1], pt.sectors[ic][1],
                                                                      site_phases)
                        P[r, c] = np.vdot(pt.sectors[ir_][3][:, lr], blocks[key] @ pt.sectors[ic][3][:, lc])
                min_sv = float(np.linalg.svd(P, compute_uv=False).min())
                min_svs.append(min_sv)
                if not min_sv > polarization_atol:
                    raise RuntimeError(f"Projected Resta operator is singular at flux index {it + 1}; "
                                       "polarization is undefined.")
                ev = np.linalg.eigvals(P)
                polarization_eigenvalues[it] = ev[np.argsort(np.angle(ev), kind="stable")]
            finally:
                pt.close()
    finally:
        if catalog is not None:
            catalog.close()
    raw = np.angle(polarization_eigenvalues) / (2.0 * math.pi)
    polarizations = tracked_phase_branches(raw)
    traj = polarizations - polarizations[:1, :]
    return SimpleNamespace(
        twisted_phases_over_2pi_list=fluxes, energies=energies, sector_labels=all_labels,
        flux_direction=flux_direction, polarization_direction=polarization_direction, nev_per_sector=nstates,
        manifold_states=specs, manifold_size=nstates, selected_states=selected_states, gaps=np.array(gaps),
        min_gap=float(min(gaps)), min_position_singular_values=np.array(min_svs),
        polarization_eigenvalues=polarization_eigenvalues, polarizations=polarizations,
        pumped_charge_trajectories=traj, pumped_charges=traj[-1].copy(), include_sublattice=include_sublattice,
        is_identity=is_identity)


def many_body_chern_number(model: Model, sector_labels, *, filling_fraction=None, flux_grid_size=(5, 5),
                           nev_per_sector: int = 1, manifold_size: Optional[int] = None, det_atol: float = 1e-10,
                           ctx: Optional[ed.Context] = None):
    """Chern number of the globally lowest ``manifold_size`` states over the twisted-boundary
    flux torus (``observables/many_body_chern_number.jl:150-258``): at every grid flux
    (theta_x, theta_y) = (ix/nx, iy/ny) all sectors are solved, the manifold's states are
    expanded into the Fock basis on the GPU, the U(1) links are determinant phases of the
    overlap matrices between neighbouring grid points, and ``chern_number`` = sum of the
    plaquette phases / 2 pi."""
    if model.lattice.dim != 2:
        raise ValueError("many_body_chern_number currently expects a 2D lattice.")
    if isinstance(flux_grid_size, int):
        nx = ny = int(flux_grid_size)
    else:
        nx, ny = int(flux_grid_size[0]), int(flux_grid_size[1])
    if nx < 2 or ny < 2:
        raise ValueError("flux_grid_size must be at least 2 in both directions.")
    if isinstance(sector_labels, str) and sector_labels == "all" and manifold_size is None:
        raise ValueError("Specify manifold_size with :all.")
    labels, is_identity = _labels(model, sector_labels)
    nstates = len(labels) * nev_per_sector if manifold_size is None else int(manifold_size)
    if nstates <= 0:
        raise ValueError("manifold_size must be positive.")
    ctx = ctx or ed.default_context()
    n_filled = _n_filled(model, filling_fraction)
    flux_x = np.arange(nx) / nx
    flux_y = np.arange(ny) / ny
    states = [[None] * ny for _ in range(nx)]
    energies = [[None] * ny for _ in range(nx)]
    selected = [[None] * ny for _ in range(nx)]
    gaps = np.zeros((nx, ny))
    catalog = None
    try:
        for ix in range(nx):
            for iy in range(ny):
                pt = _FluxPoint(model, (flux_x[ix], flux_y[iy]), n_filled, nstates + 1, is_identity, ctx, catalog)
                catalog = pt.catalog
                try:
                    gap, sel = pt.select(nstates, "selected manifold")
                    if not gap > 1e-9:
                        raise RuntimeError(f"Low-energy manifold is not isolated at flux "
                                           f"{[flux_x[ix], flux_y[iy]]} (gap={gap}).")
                    gaps[ix, iy] = gap
                    states[ix][iy] = [ed.fock_amplitudes(pt.sectors[i][3][:, l], pt.sectors[i][1]) for i, l in sel]
                    energies[ix][iy] = np.array([pt.sectors[i][2][l] for i, l in sel])
                    selected[ix][iy] = [(pt.sectors[i][0], l + 1) for i, l in sel]
                finally:
                    pt.close()
    finally:
        if catalog is not None:
            catalog.close()
    Ux = np.empty((nx, ny), dtype=np.complex128)
    Uy = np.empty((nx, ny), dtype=np.complex128)
    dx = np.empty((nx, ny))
    dy = np.empty((nx, ny))
    for ix in range(nx):
        for iy in range(ny):
            Ux[ix, iy], dx[ix, iy] = unit_det_link(states[ix][iy], states[(ix + 1) % nx][iy], det_atol)
            Uy[ix, iy], dy[ix, iy] = unit_det_link(states[ix][iy], states[ix][(iy + 1) % ny], det_atol)
    F = plaquette_phases(Ux, Uy)
    c = float(F.sum() / (2.0 * math.pi))
    return SimpleNamespace(
        chern_number=c, manifold_size=nstates, gaps=gaps, min_gap=float(gaps.min()), selected_states=selected,
        rounded_chern=int(round(c)), berry_curvature=F, flux_x=flux_x, flux_y=flux_y, sector_labels=labels,
        nev_per_sector=nev_per_sector, flux_grid_size=(nx, ny), energies=energies,
        min_link_det=float(min(dx.min(), dy.min())), link_det_abs_x=dx, link_det_abs_y=dy)
