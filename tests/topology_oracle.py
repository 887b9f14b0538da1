"""The flux-insertion observables (charge pump, flux-torus Chern number) computed with
the oracle: oracle catalogs and sectors, dense eigendecompositions of the oracle's
sector blocks, the oracle's Fock expansion and projected position operator.  The small
dense steps on top (branch unwrapping, determinant links, plaquettes) are the product's
helpers from ``topology``; the reference's stored notebook outputs pin them together with
the oracle (tests/test_topology.py)."""
import math
from types import SimpleNamespace

import numpy as np

from realspace_exactdiagonalization_amd import lattice as L
from realspace_exactdiagonalization_amd import topology as T


def flux_point(O, model, flux, nev):
    """Every momentum sector at one flux: [(label, sector, vals, vecs)] (dense eigh)."""
    m = L.twisted(model, flux)
    G = L.model_group(m, flux)
    cat = O.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    out = []
    for ir in L.translation_irreps(G, m.lattice):
        sec = O.Sector(cat, ir.values)
        if sec.dim == 0:
            out.append((tuple(ir.label), sec, np.zeros(0), np.zeros((0, 0), dtype=np.complex128)))
            continue
        sec.set_terms(*L.term_arrays(m))
        w, v = np.linalg.eigh(sec.dense())
        out.append((tuple(ir.label), sec, w[:nev], v[:, :nev]))
    return cat, out


def select(sectors, nstates):
    cand = sorted(((float(v[l]), i, l) for i, (_, _, v, _) in enumerate(sectors) for l in range(len(v))),
                  key=lambda c: c[0])
    return cand[nstates][0] - cand[nstates - 1][0], [(i, l) for _, i, l in cand[:nstates]]


def position_block(sec_to, sec_from, phi):
    rows, vals = sec_to.position_operator(sec_from, phi)
    U = np.zeros((sec_to.dim, sec_from.dim), dtype=np.complex128)
    for col, (r, v) in enumerate(zip(rows, vals)):
        if r > 0 and abs(v) > 1e-13:
            U[r - 1, col] = v
    return U


def charge_pump(O, model, fluxes, nstates, flux_direction=1, polarization_direction=2):
    phi = T.many_body_position_phases(model.lattice, polarization_direction)
    gaps, evs, selected, energies = [], [], [], []
    for t in fluxes:
        f = [0.0, 0.0]
        f[flux_direction - 1] = t
        cat, secs = flux_point(O, model, tuple(f), nstates + 1)
        gap, sel = select(secs, nstates)
        gaps.append(gap)
        selected.append([(secs[i][0], l + 1) for i, l in sel])
        energies.append({secs[i][0]: secs[i][2] for i in range(len(secs))})
        P = np.zeros((nstates, nstates), dtype=np.complex128)
        for r, (ir_, lr) in enumerate(sel):
            for c, (ic, lc) in enumerate(sel):
                P[r, c] = np.vdot(secs[ir_][3][:, lr], position_block(secs[ir_][1], secs[ic][1], phi) @ secs[ic][3][:, lc])
        ev = np.linalg.eigvals(P)
        evs.append(ev[np.argsort(np.angle(ev), kind="stable")])
    pol = T.tracked_phase_branches(np.angle(np.array(evs)) / (2 * math.pi))
    return SimpleNamespace(gaps=np.array(gaps), min_gap=min(gaps), polarization_eigenvalues=np.array(evs),
                           pumped_charges=pol[-1] - pol[0], selected_states=selected, energies=energies)


def chern_number(O, model, grid, nstates):
    nx, ny = grid
    states, gaps = {}, np.zeros((nx, ny))
    for ix in range(nx):
        for iy in range(ny):
            cat, secs = flux_point(O, model, (ix / nx, iy / ny), nstates + 1)
            gaps[ix, iy], sel = select(secs, nstates)
            # rank-indexed dense expansions; ranks serve as the (ascending) keys
            states[ix, iy] = []
            for i, l in sel:
                a = secs[i][1].fock_expand(secs[i][3][:, l])
                states[ix, iy].append((np.arange(len(a), dtype=np.uint64), a))
    Ux = np.empty((nx, ny), dtype=np.complex128)
    Uy = np.empty((nx, ny), dtype=np.complex128)
    dets = []
    for ix in range(nx):
        for iy in range(ny):
            Ux[ix, iy], d1 = T.unit_det_link(states[ix, iy], states[(ix + 1) % nx, iy])
            Uy[ix, iy], d2 = T.unit_det_link(states[ix, iy], states[ix, (iy + 1) % ny])
            dets += [d1, d2]
    F = T.plaquette_phases(Ux, Uy)
    return SimpleNamespace(chern_number=F.sum() / (2 * math.pi), min_gap=gaps.min(), min_link_det=min(dets),
                           berry_curvature=F)
