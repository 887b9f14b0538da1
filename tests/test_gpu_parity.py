"""Parity of the HIP path (through the C ABI of libsed.so) with the CPU oracle and the
reference's stored results.

Bars: bit-exact for integer tables (rep index, group index, sign, representatives,
stabilizer orders); rep amplitudes exact at zero twist and within 1e-13 otherwise
(the reference forms them with @fastmath products); matvec within 1e-12 relative
(FP64) and 2e-6 relative (FP32); eigenvalues within 1e-10 relative.
"""
import math
import os

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import lattice as L

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


CASES = {
    "hald2x3": (lambda: L.haldane_honeycomb(2, 3, n_filled=3), (0.0, 0.0)),
    "hub2x3": (lambda: L.hubbard_square(2, 3, U=4.0, n_filled=6), (0.0, 0.0)),
    "heis14": (lambda: L.heisenberg_chain(14), (0.0, 0.0)),
    "checker3x2_f": (lambda: L.checkerboard(3, 2, n_filled=4, V1=1.0, V2=0.5), (0.0, 0.0)),
    "checker3x2_f_flux": (lambda: L.checkerboard(3, 2, n_filled=4, V1=1.0, V2=0.5), (0.23, 0.17)),
    "checker3x2_b_flux": (lambda: L.checkerboard(3, 2, n_filled=4, statistics=L.BOSONIC), (0.23, 0.17)),
    "hald3x2_flux": (lambda: L.haldane_honeycomb(3, 2, n_filled=3, V1=0.3), (0.4, 0.0)),
    "checker3x4_f": (lambda: L.checkerboard(3, 4, n_filled=8), (0.0, 0.0)),
    # more than 32 sites: 64-bit masks (the instantiations the checkerboard 4x6 workload runs)
    "checker4x6_k4": (lambda: L.checkerboard(4, 6, n_filled=4), (0.0, 0.0)),
    "checker4x6_k3_flux": (lambda: L.checkerboard(4, 6, n_filled=3), (0.3, 0.7)),
    "checker3x6_k6_flux": (lambda: L.checkerboard(3, 6, n_filled=6), (0.11, 0.4)),
    "hald4x5_b_k4": (lambda: L.haldane_honeycomb(4, 5, n_filled=4), (0.0, 0.0)),
}
# sectors of ~10^5 states: the matvec test takes the first few momenta only
IRREP_LIMIT = {"checker3x6_k6_flux": 4}


def assert_lowest_levels(vals, ref, dim, what, rel=1e-10):
    """vals (the solver's lowest eigenvalues) against the oracle's lowest levels ref (a few
    more than len(vals)), within rel relative.  Where the oracle's spectrum has no exactly
    degenerate level among the compared ones, vals are its lowest len(vals) eigenvalues.
    Dimensions above 500 go through single-vector Lanczos, which (as the reference's
    KrylovKit eigsolve from one start vector) spans one direction of an exactly degenerate
    eigenspace: there each value is an oracle level, the lowest is the ground level, and
    every oracle level below the largest value appears."""
    scale = max(1.0, np.abs(ref).max())
    tol = rel * scale
    nev = len(vals)
    assert np.all(np.diff(vals) >= -tol), (what, vals)
    degenerate = np.any(np.diff(ref[:nev + 1]) <= 1e-8 * scale)
    if dim <= 500 or not degenerate:
        assert np.abs(vals - ref[:nev]).max() <= tol, (what, vals, ref)
        return
    levels = [ref[0]]
    for v in ref[1:]:
        if v - levels[-1] > 1e-8 * scale:
            levels.append(v)
    levels = np.array(levels)
    assert abs(vals[0] - ref[0]) <= tol, (what, vals, ref)
    for v in vals:
        assert np.abs(levels - v).min() <= tol, (what, v, levels)
    for lv in levels[levels < vals[-1] - tol]:
        assert np.abs(vals - lv).min() <= tol, (what, lv, vals)
    print(f"[eigs] {what}: exactly degenerate oracle levels {ref[:nev + 1]}, solver {vals}", flush=True)


def make(name):
    f, th = CASES[name]
    m = L.twisted(f(), th)
    return m, L.model_group(m, th)


@pytest.mark.parametrize("name", list(CASES))
def test_catalog_tables_match_reference_layout(ed, ctx, oracle_mod, name):
    m, G = make(name)
    print(f"[catalog] {name} N={m.n_site} k={m.n_filled} |G|={G.order}", flush=True)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    assert gc.n_total == oc.n_total and gc.n_orbits == oc.n_orbits
    orr, ors, ora = oc.tables()
    grr, grs, gra = gc.tables()
    assert np.array_equal(orr, grr)
    assert np.array_equal(ors, grs)
    exact = all(abs(p - 1) == 0 for p in G.phases.reshape(-1))
    if exact:
        assert np.array_equal(ora, gra)
    else:
        assert np.abs(ora - gra).max() < 1e-13
    oreps, oso, ogi, ogp = oc.orbits()
    greps, gso, ggi, ggp = gc.orbits()
    assert np.array_equal(oreps, greps) and np.array_equal(oso, gso) and np.array_equal(ogi, ggi)
    assert np.abs(ogp - ggp).max() < 1e-13


@pytest.mark.parametrize("name", list(CASES))
def test_sector_basis_and_matvec(ed, ctx, oracle_mod, name):
    m, G = make(name)
    print(f"[sector/matvec] {name}", flush=True)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    rng = np.random.default_rng(7)
    for ir in L.translation_irreps(G, m.lattice)[:IRREP_LIMIT.get(name)]:
        os_ = oracle_mod.Sector(oc, ir.values)
        gb = ed.build_symmetry_sector_basis(gc, ir)
        assert gb.dim == os_.dim
        if os_.dim == 0:
            continue
        oreps, ostab = os_.basis()
        assert np.array_equal(oreps, gb.representative_mask_list)
        assert np.array_equal(ostab, gb.stabilizer_order_list)
        os_.set_terms(*L.term_arrays(m))
        H = ed.MatrixFreeHamiltonian(gb, m.bilinear_terms, m.density_terms)
        x = rng.standard_normal(gb.dim) + 1j * rng.standard_normal(gb.dim)
        y_ref = os_.apply_table(x, 1)
        y = H(x)
        scale = np.abs(y_ref).max() + 1e-300
        assert np.abs(y - y_ref).max() <= 1e-12 * scale
        H.precision = ed.PREC_FP32
        y32 = H(x)
        assert np.abs(y32 - y_ref).max() <= 2e-6 * scale * max(1, math.sqrt(len(m.bilinear_terms) / 10))
        # flux_regression.jl:32-34 probe vector x = 1:n + i
        H.precision = ed.PREC_FP64
        xr = np.arange(1, gb.dim + 1) + 1j
        assert np.abs(H(xr) - os_.apply_table(xr, 1)).max() <= 1e-12 * np.abs(os_.apply_table(xr, 1)).max()


def test_matvec_deterministic(ed, ctx):
    m, G = make("checker3x4_f")
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(G, m.lattice)[0])
    H = ed.MatrixFreeHamiltonian(gb, m.bilinear_terms, m.density_terms)
    x = np.random.default_rng(3).standard_normal(gb.dim) + 0.5j
    a, b = H(x), H(x)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["hald2x3", "hub2x3", "heis14", "checker3x2_f_flux", "hald3x2_flux",
                                  "checker4x6_k4", "checker4x6_k3_flux", "checker3x6_k6_flux"])
@pytest.mark.parametrize("precision", [0, 2])
def test_eigenvalues_vs_oracle_all_sectors(ed, ctx, oracle_mod, name, precision):
    m, G = make(name)
    print(f"[eigs] {name} precision={precision}", flush=True)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    for ir in L.translation_irreps(G, m.lattice)[:IRREP_LIMIT.get(name)]:
        os_ = oracle_mod.Sector(oc, ir.values)
        if os_.dim == 0:
            continue
        os_.set_terms(*L.term_arrays(m))
        nev = min(3, os_.dim)
        ref = oracle_mod.lowest_eigenvalues(os_, min(nev + 3, os_.dim))
        gb = ed.build_symmetry_sector_basis(gc, ir)
        H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
        vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=nev, precision=precision)
        assert_lowest_levels(vals, ref, n, (name, ir.label))
        # returned vectors are eigenvectors of the reference-basis operator
        for j in range(nev):
            r = H(vecs[:, j]) - vals[j] * vecs[:, j]
            assert np.linalg.norm(r) < 1e-6


@pytest.mark.parametrize("twist", [(0.0, 0.0), (0.0, 0.3)])
def test_dense_path_eigenvectors_in_degenerate_cluster(ed, ctx, oracle_mod, twist):
    """Dense path (dimension <= 500, src/symmetry_resolved_ed.jl:1004-1017): U = 8
    Hubbard 2x3 at k = (0, 1), dimension 150, whose levels 2-6 form a five-fold cluster
    (-2.5606 at zero twist, -2.3580 at twist 0.3 along y).  The six lowest eigenvalues
    equal the oracle's (1e-12), the returned vectors are orthonormal (1e-12) with
    residuals |H v - lambda v| at rounding level (1e-11 |H|), and the span of the vectors
    of each cluster equals the oracle's eigenspace (projector difference 1e-9)."""
    m = L.twisted(L.reference_benchmark_model("Hubbard_Fermion", "2×3"), twist)
    G = L.model_group(m, twist)
    print(f"[dense] hubbard 2x3 k=(0,1) twist={twist}", flush=True)
    ir = [r for r in L.translation_irreps(G, m.lattice) if tuple(r.label) == (0, 1)][0]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    Hd = os_.dense()
    Hd = 0.5 * (Hd + Hd.conj().T)
    w, U = np.linalg.eigh(Hd)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    assert gb.dim == os_.dim == 150
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    nev = 6
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=nev)
    hn = np.abs(w).max()
    assert np.abs(vals - w[:nev]).max() <= 1e-12 * hn
    assert np.abs(vecs.conj().T @ vecs - np.eye(nev)).max() <= 1e-12
    for j in range(nev):
        assert np.linalg.norm(H(vecs[:, j]) - vals[j] * vecs[:, j]) <= 1e-11 * hn
    # clusters of the oracle's spectrum among the six, each compared as a subspace
    j = 0
    while j < nev:
        c = j
        while c + 1 < nev and w[c + 1] - w[c] < 1e-8 * hn:
            c += 1
        if c + 1 < len(w) and w[c + 1] - w[c] < 1e-8 * hn:   # cluster cut by nev: skip it
            break
        P = vecs[:, j:c + 1] @ vecs[:, j:c + 1].conj().T
        Q = U[:, j:c + 1] @ U[:, j:c + 1].conj().T
        assert np.abs(P - Q).max() <= 1e-9, (j, c)
        j = c + 1
    assert w[5] - w[1] < 1e-8 * hn and w[6] - w[5] > 1e-3   # the five-fold level, whole
    gb.close()
    gc.close()


GOLDEN = [("Heisenberg", "N=20"), ("Heisenberg", "N=24"), ("Haldane_Boson", "3×4"), ("Haldane_Boson", "2×7"),
          ("Hubbard_Fermion", "2×5"), ("Hubbard_Fermion", "3×4")]


@pytest.mark.parametrize("model,label", GOLDEN)
def test_lowest_energy_vs_reference_csv(ed, ctx, golden, model, label):
    from tests.test_oracle import csv_model
    print(f"[golden] {model} {label}", flush=True)
    rows = [r for r in golden["benchmark_raw"] if r["model"] == model and r["label"] == label]
    m = csv_model(model, label)
    data = ed.build_ed_data(m, ctx=ctx)
    assert data.orbit_catalog.n_orbits == rows[0]["n_orbits"]
    for prec in (0, 2):
        vals, _ = ed.ed_scan_at_irrep_matrixfree(data.irrep_list[0].label, data, nev=1, precision=prec,
                                                 return_vectors=False)
        assert data.sector_dims[0] == rows[0]["sector_dim"]
        for r in rows:
            assert abs(vals[0] - r["energy"]) <= 1e-10 * abs(r["energy"]), (prec, vals[0], r["energy"])


def test_site_densities(ed, ctx, oracle_mod):
    """Sector densities vs the oracle eigenvector (orbit expansion), plus I4: translation
    invariance per sublattice and sum = particle number."""
    m, G = make("checker3x2_f_flux")
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    ir = L.translation_irreps(G, m.lattice)[0]
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=2)
    d = ed.sector_density(H, 0)
    assert abs(d.sum() - m.n_filled) < 1e-10
    for s in (1, 2):
        sub = [d[i - 1] for i in range(1, m.n_site + 1) if m.lattice.site_list[i - 1][1] == s]
        assert np.ptp(sub) < 1e-10
    # oracle: <n_i> = sum_r |c_r|^2 / |G| * #{g : i in g(rep_r)}
    reps = gb.representative_mask_list
    c = vecs[:, 0]
    ref = np.zeros(m.n_site)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    for r, mask in enumerate(reps):
        for g in range(1, G.order + 1):
            img, _ = oc.apply_operation(int(mask), g)
            for i in range(m.n_site):
                if (img >> i) & 1:
                    ref[i] += abs(c[r]) ** 2 / G.order
    assert np.abs(ref - d).max() < 1e-12


def test_server_results_alone_and_in_mix(ed, ctx):
    """Each solve gives the same eigenvalues alone and when batched with others."""
    m0 = L.haldane_honeycomb(3, 4)
    bases = []
    keep = []
    for th in [(0.0, 0.0), (0.25, 0.0), (0.5, 0.0)]:
        m = L.twisted(m0, th)
        G = L.model_group(m, th)
        cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                              particle_statistics=m.statistics, ctx=ctx)
        keep.append(cat)
        for ir in L.translation_irreps(G, m.lattice)[:4]:
            b = ed.build_symmetry_sector_basis(cat, ir)
            b.set_terms(m.bilinear_terms, m.density_terms)
            bases.append(b)
    alone = []
    for b in bases:
        H = ed.MatrixFreeHamiltonian.__new__(ed.MatrixFreeHamiltonian)
        H.basis, H.n, H.precision = b, b.dim, ed.PREC_FP64
        alone.append(ed.diagonalize_block_matrixfree(H, b.dim, nev=3, return_vectors=False)[0])
    srv = ed.SectorServer(ctx, max_slots=4)
    ids = [srv.submit(b, 3, precision=p) for b in bases for p in (ed.PREC_FP64,)]
    srv.run(0)
    for rid, ref in zip(ids, alone):
        st, ev, stats, lat = srv.result(rid)
        assert st == 0
        assert np.abs(ev - ref).max() <= 1e-10 * np.abs(ref).max()
        assert lat >= 0
    ids32 = [srv.submit(b, 3, precision=ed.PREC_FP32_REFINE) for b in bases]
    srv.run(0)
    for rid, ref in zip(ids32, alone):
        st, ev, stats, lat = srv.result(rid)
        assert st == 0 and np.abs(ev - ref).max() <= 1e-10 * np.abs(ref).max()
    srv.close()


@pytest.mark.parametrize("name,irrep", [("checker3x4_f", 5), ("checker4x6_k4", 7), ("checker3x6_k6_flux", 1)])
def test_every_launch_shape_matches_oracle(ed, ctx, oracle_mod, name, irrep):
    """Each H|psi> launch shape (lanes per row x hops per batch, site-by-site and flat hop
    loops, three workgroup sizes) and XCD run length gives the oracle's matvec, bitwise
    identical between two runs of the same shape; on 32-bit and 64-bit masks (48 and 36
    sites, fermions, with and without twist)."""
    m, G = make(name)
    ir = L.translation_irreps(G, m.lattice)[irrep]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H = ed.MatrixFreeHamiltonian(gb, m.bilinear_terms, m.density_terms)
    x = np.random.default_rng(11).standard_normal(gb.dim) + 1j * np.random.default_rng(12).standard_normal(gb.dim)
    ref = os_.apply_table(x, 1)
    n_shapes = len(gb.shape(ed.PREC_FP64)["tune_ms"])
    assert n_shapes >= 4
    print(f"[shape] {name} irrep={irrep} dim={gb.dim}", flush=True)
    for prec, tol in ((ed.PREC_FP64, 1e-12), (ed.PREC_FP32, 5e-6)):
        for sh, chunk in [(s, 1) for s in range(n_shapes)] + [(2, 0), (2, 3), (1, 32), (n_shapes - 1, 3)]:
            print(f"[shape] {name} prec={prec} shape={sh} xcd_chunk={chunk}", flush=True)
            gb.set_shape(prec, sh, chunk)
            H.precision = prec
            y1, y2 = H(x), H(x)
            assert np.array_equal(y1, y2)
            assert np.abs(y1 - ref).max() <= tol * np.abs(ref).max()


def test_lanczos_graph_replay_matches_eager(ed, ctx, monkeypatch):
    """The captured Lanczos step (hipGraph) gives bitwise the eigenvalues of eager launches,
    over several solves in a row that reuse the graph with eager work in between."""
    m, G = make("checker3x4_f")
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    irs = L.translation_irreps(G, m.lattice)
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SED_LANCZOS_GRAPH", mode)
        out = []
        for ir in (irs[0], irs[3], irs[0]):
            gb = ed.build_symmetry_sector_basis(gc, ir)
            H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
            for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
                out.append(ed.diagonalize_block_matrixfree(H, n, nev=3, precision=prec, return_vectors=False)[0])
                H(np.ones(n, dtype=np.complex128))   # eager work between replays
        res[mode] = out
    for a, b in zip(res["0"], res["1"]):
        assert np.array_equal(a, b)


def test_step_graph_not_reused_for_a_new_sector_of_equal_shape(ed, oracle_mod):
    """Regression: sectors of equal dimension built one after another, each destroyed
    (with its catalog) before the next is built, must not replay the previous sector's
    captured Lanczos step.  Each solve matches the oracle's eigenvalues."""
    c = ed.Context(0)
    try:
        for th in (0.0, 0.37, 0.0, 0.61):
            m = L.twisted(L.heisenberg_chain(16), (th, 0.0))
            G = L.model_group(m, (th, 0.0))
            print(f"[graph-reuse] xxz16 theta={th}", flush=True)
            ir = L.translation_irreps(G, m.lattice)[0]
            oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, False)
            os_ = oracle_mod.Sector(oc, ir.values)
            os_.set_terms(*L.term_arrays(m))
            ref = oracle_mod.lowest_eigenvalues(os_, 3)
            gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                                 particle_statistics=m.statistics, ctx=c)
            gb = ed.build_symmetry_sector_basis(gc, ir)
            assert gb.dim == 810
            H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
            vals = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
            assert np.abs(vals - ref).max() <= 1e-10 * np.abs(ref).max(), (th, vals, ref)
            if th == 0.37:
                gc.close()   # catalog first: released with its last sector
                gb.close()
            else:
                gb.close()
                gc.close()
        # context first, then its sector and catalog
        gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                             particle_statistics=m.statistics, ctx=c)
        gb = ed.build_symmetry_sector_basis(gc, ir)
        c.close()
        gb.close()
        gc.close()
    finally:
        c.close()


def test_packed_engine_matches_lane_path(ed, ctx, monkeypatch):
    """Small sectors solved in the packed persistent launches give the eigenvalues of the
    per-sector lane path (FP64 and FP32+refine), for bosons and fermions."""
    cases = [(L.haldane_honeycomb(3, 4), (0.3, 0.0)), (L.checkerboard(3, 4, n_filled=6), (0.0, 0.21))]
    bases, keep = [], []
    for base, th in cases:
        m = L.twisted(base, th)
        G = L.model_group(m, th)
        cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                              particle_statistics=m.statistics, ctx=ctx)
        keep.append(cat)
        for ir in L.translation_irreps(G, m.lattice)[:5]:
            b = ed.build_symmetry_sector_basis(cat, ir)
            if b.dim <= 500:
                continue
            b.set_terms(m.bilinear_terms, m.density_terms)
            bases.append(b)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SED_PACKED", mode)
        srv = ed.SectorServer(ctx, max_slots=4)
        ids = [(srv.submit(b, 3, precision=p), p) for b in bases for p in (ed.PREC_FP64, ed.PREC_FP32_REFINE)]
        srv.run(0)
        res = []
        for rid, p in ids:
            st, ev, stats, lat = srv.result(rid)
            assert st == 0, (mode, rid)
            res.append(ev)
        srv.close()
        out[mode] = res
    for a, b in zip(out["0"], out["1"]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def test_packed_engine_partial_reorthogonalisation(ed, ctx, oracle_mod, monkeypatch):
    """The packed persistent launches (SED_PACKED=1) with partial re-orthogonalisation give
    the eigenvalues of a whole-basis pass at every step (SED_PRO=0) to 1e-11 relative and
    the oracle's to 1e-10, FP64 and FP32+refine, with at most half of the steps
    orthogonalised against the whole basis (every step without it).  Twisted Haldane 3x4
    bosons and checkerboard 3x4 fermions, sectors of dimension 1,000-12,000."""
    cases = [(L.haldane_honeycomb(3, 4), (0.3, 0.0)), (L.checkerboard(3, 4, n_filled=6), (0.0, 0.21))]
    bases, refs, keep = [], [], []
    for base, th in cases:
        m = L.twisted(base, th)
        G = L.model_group(m, th)
        cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                              particle_statistics=m.statistics, ctx=ctx)
        oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
        keep.append(cat)
        for ir in L.translation_irreps(G, m.lattice)[:3]:
            b = ed.build_symmetry_sector_basis(cat, ir)
            if b.dim <= 500:
                continue
            b.set_terms(m.bilinear_terms, m.density_terms)
            os_ = oracle_mod.Sector(oc, ir.values)
            os_.set_terms(*L.term_arrays(m))
            bases.append(b)
            refs.append(oracle_mod.lowest_eigenvalues(os_, 3))
    monkeypatch.setenv("SED_PACKED", "1")
    out = {}
    for mode in ("pro", "full"):
        if mode == "full":
            monkeypatch.setenv("SED_PRO", "0")
        else:
            monkeypatch.delenv("SED_PRO", raising=False)
        srv = ed.SectorServer(ctx, max_slots=4)
        ids = [(srv.submit(b, 3, precision=p), p) for b in bases for p in (ed.PREC_FP64, ed.PREC_FP32_REFINE)]
        srv.run(0)
        res = []
        for rid, p in ids:
            st, ev, stats, lat = srv.result(rid)
            assert st == 0, (mode, rid)
            steps = stats["iters_fp32"] + stats["iters_fp64"]
            print(f"[packed-pro] {mode} prec={p} steps={steps} full={stats['full_orth_steps']}", flush=True)
            if mode == "full":
                assert stats["full_orth_steps"] == steps
            else:
                assert 0 < stats["full_orth_steps"] <= 0.5 * steps
            res.append(ev)
        srv.close()
        out[mode] = res
    for i, (a, b) in enumerate(zip(out["pro"], out["full"])):
        ref = refs[i // 2]
        assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()
        assert np.abs(a - ref).max() <= 1e-10 * np.abs(ref).max()


def test_mfma_reorthogonalisation_matches_vector_loops(ed, ctx, oracle_mod, monkeypatch):
    """The Lanczos re-orthogonalisation and restart rotation on the FP64 matrix cores
    (forced with SED_REORTH=mfma: this sector is below the size from which they are the
    default) give the eigenvalues of the row-wise VALU loops (SED_REORTH=valu) to 1e-12
    relative, both match the oracle to 1e-10, and the MFMA path is bitwise reproducible.
    The sector (checkerboard 3x4, 8 fermions, dim 61,333) needs several thick restarts and
    basis sizes on both sides of every 16-vector tile boundary."""
    m, G = make("checker3x4_f")
    ir = L.translation_irreps(G, m.lattice)[0]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    ref = oracle_mod.lowest_eigenvalues(os_, 4)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    print(f"[reorth] checker3x4 dim={n}", flush=True)
    res = {}
    for mode in ("mfma", "valu", "mfma"):
        monkeypatch.setenv("SED_REORTH", mode)
        for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
            st = {}
            vals = ed.diagonalize_block_matrixfree(H, n, nev=4, precision=prec, return_vectors=False, stats=st)[0]
            assert st["restarts"] >= 2
            assert np.abs(vals - ref).max() <= 1e-10 * np.abs(ref).max(), (mode, prec, vals, ref)
            key = (mode, prec)
            if key in res:
                assert np.array_equal(res[key], vals)   # bitwise reproducible
            res[key] = vals
    for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
        a, b = res[("mfma", prec)], res[("valu", prec)]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    gb.close()
    gc.close()


PRO_CASES = {
    # name: (model, twist, irrep index, nev)
    "checker3x4_f": (lambda: L.checkerboard(3, 4, n_filled=8), (0.0, 0.0), 0, 4),
    "hald3x4_flux_nev8": (lambda: L.haldane_honeycomb(3, 4, n_filled=6), (0.3, 0.0), 5, 8),
    "hald3x4_flux_nev10": (lambda: L.haldane_honeycomb(3, 4, n_filled=6), (0.3, 0.0), 4, 10),
    "hald3x4_flux_nev20": (lambda: L.haldane_honeycomb(3, 4, n_filled=6), (0.3, 0.0), 2, 20),
    "xxz20": (lambda: L.heisenberg_chain(20), (0.0, 0.0), 0, 3),
}


@pytest.mark.parametrize("mode", ["default", "pro_forced", "full"])
def test_degenerate_spectrum_nev20(ed, ctx, oracle_mod, monkeypatch, mode):
    """Spin-degenerate spectrum with the largest basis: U = 8 Hubbard 2x4 at twist
    (0.21, 0.37), k = 0 (dimension 1638), nev = 20.  "default": above the partial
    re-orthogonalisation limit (nev <= 5) every step is orthogonalised against the whole
    basis; "pro_forced": partial re-orthogonalisation forced on (SED_PRO_MAX_NEV=64), whose
    pairs miss tol here, so the FP64 phase's residual check must catch that and repeat the
    solve with whole-basis passes; "full": SED_PRO=0.  Single-vector Lanczos may miss
    copies of a degenerate level (as the reference's KrylovKit can), so the bar is what
    must hold regardless: every returned value is an eigenvalue of the oracle's matrix
    (1e-9), none is returned more often than its multiplicity (no ghost copies), and the
    returned vectors are orthonormal (1e-8) with residuals |H x - lambda x| <= 1e-9."""
    if mode == "full":
        monkeypatch.setenv("SED_PRO", "0")
    if mode == "pro_forced":
        monkeypatch.setenv("SED_PRO_MAX_NEV", "64")
    th = (0.21, 0.37)
    m = L.twisted(L.hubbard_square(2, 4, U=8.0), th)
    G = L.model_group(m, th)
    ir = L.translation_irreps(G, m.lattice)[0]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    Hd = os_.dense()
    w = np.linalg.eigvalsh(0.5 * (Hd + Hd.conj().T))
    hn = np.abs(w).max()
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    assert n == 1638
    st = {}
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=20, stats=st)
    print(f"[degenerate nev20] {mode}: steps {st['iters_fp64']} full {st['full_orth_steps']} "
          f"restarts {st['restarts']} matvecs {st['matvecs']} residual {st['residual_max']:.1e}", flush=True)
    for lam in vals:
        assert np.abs(w - lam).min() <= 1e-9 * hn
        mult = int(np.sum(np.abs(w - lam) <= 1e-8 * hn))
        assert int(np.sum(np.abs(vals - lam) <= 1e-8 * hn)) <= mult
    assert np.abs(vecs.conj().T @ vecs - np.eye(20)).max() <= 1e-8
    for j in range(20):
        assert np.linalg.norm(H(vecs[:, j]) - vals[j] * vecs[:, j]) <= 1e-9
    assert st["residual_max"] <= 1e-10
    if mode == "pro_forced":   # the check found pairs above tol and the whole-basis repeat ran
        assert st["repeated"] == 1
    else:
        assert st["repeated"] == 0
    gb.close()
    gc.close()


@pytest.mark.parametrize("name,reorth", [("checker3x4_f", "mfma"), ("checker3x4_f", None),
                                         ("hald3x4_flux_nev8", None), ("hald3x4_flux_nev10", "mfma"),
                                         ("hald3x4_flux_nev20", "mfma"), ("xxz20", None)])
def test_partial_reorthogonalisation_matches_full(ed, ctx, oracle_mod, monkeypatch, name, reorth):
    """Partial re-orthogonalisation (the default: Simon's recurrence decides when a step is
    orthogonalised against the whole basis) against a pass over the whole basis at every
    step (SED_PRO=0): the same eigenvalues to 1e-11 relative in FP64 and FP32+refine, both
    within 1e-10 of the oracle, the same number of Lanczos steps within 10 %, at most half
    as many steps orthogonalised against the whole basis for nev <= 5 (larger nev always
    orthogonalises against the whole basis), and bitwise reproducible runs.
    Sectors with several thick restarts: checkerboard 3x4 fermions (dim 61,333; also with
    the matrix-core kernels forced, SED_REORTH=mfma, which large sectors use by default),
    the twisted Haldane 3x4 bosons with nev = 8 (dim ~11,200), nev = 10 (m = 44) and nev = 20
    (the largest Krylov basis, m = 64), the last two with the matrix-core kernels forced (the
    restart rotation's three instantiations: m <= 40, <= 48, <= 64), XXZ N=20 (dim 9,252)."""
    if reorth:
        monkeypatch.setenv("SED_REORTH", reorth)
    f, th, k_idx, nev = PRO_CASES[name]
    m = L.twisted(f(), th)
    G = L.model_group(m, th)
    ir = L.translation_irreps(G, m.lattice)[k_idx]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    ref = oracle_mod.lowest_eigenvalues(os_, nev)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    print(f"[pro] {name} dim={n} nev={nev}", flush=True)
    res = {}
    for mode in ("pro", "full", "pro"):
        if mode == "full":
            monkeypatch.setenv("SED_PRO", "0")
        else:
            monkeypatch.delenv("SED_PRO", raising=False)
        for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
            st = {}
            vals = ed.diagonalize_block_matrixfree(H, n, nev=nev, precision=prec, return_vectors=False, stats=st)[0]
            print(f"[pro] {name} {mode} prec={prec} steps={st['iters_fp32']}+{st['iters_fp64']} "
                  f"full={st['full_orth_steps']} restarts={st['restarts']}", flush=True)
            assert np.abs(vals - ref).max() <= 1e-10 * np.abs(ref).max(), (mode, prec, vals, ref)
            key = (mode, prec)
            if key in res:
                assert np.array_equal(res[key][0], vals)   # bitwise reproducible
            res[key] = (vals, st)
    for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
        (a, sa), (b, sb) = res[("pro", prec)], res[("full", prec)]
        assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()
        steps_a = sa["iters_fp32"] + sa["iters_fp64"]
        steps_b = sb["iters_fp32"] + sb["iters_fp64"]
        assert abs(steps_a - steps_b) <= 0.1 * steps_b
        assert sb["full_orth_steps"] == steps_b
        if nev <= 5:
            assert sa["full_orth_steps"] <= 0.5 * steps_a
        else:   # more wanted pairs than partial re-orthogonalisation is used for
            assert sa["full_orth_steps"] == steps_a
    gb.close()
    gc.close()


OBS_GPU_CASES = ["checker3x2_f_flux", "checker3x2_b_flux", "hub2x3", "hald3x2_flux", "checker3x4_f"]


@pytest.mark.parametrize("name", OBS_GPU_CASES)
def test_density_correlation_and_one_body_match_oracle(ed, ctx, oracle_mod, name):
    """<n_i n_j>, <n_i> and <a_i^dag a_j> of a sector eigenvector on the GPU against the
    oracle's restatement of the reference's orbit-basis loops (the oracle itself is
    pinned against literal operators in the full Fock space) on the same state, to
    1e-12; plus Hermiticity, trace = particle number, sum_j <n_i n_j> = k <n_i>."""
    m, G = make(name)
    irs = L.translation_irreps(G, m.lattice)
    ir = irs[1 % len(irs)]
    print(f"[observables] {name} N={m.n_site} k={m.n_filled} |G|={G.order}", flush=True)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=2)
    c = vecs[:, 0]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    os_ = oracle_mod.Sector(oc, ir.values, basis=(gb.representative_mask_list, gb.stabilizer_order_list))
    C_ref, d_ref = os_.density_correlation(c)
    C, d = ed.density_correlation_matrix(H, 0, subtract_disconnected=False)
    assert np.abs(d - d_ref).max() < 1e-12 and np.abs(C - C_ref).max() < 1e-12
    assert np.abs(C.sum(axis=1) - m.n_filled * d).max() < 1e-10
    Cc, _ = ed.density_correlation_matrix(H, 0)
    assert np.abs(Cc - (C_ref - np.outer(d_ref, d_ref))).max() < 1e-12
    rho = ed.off_diagonal_long_range_order(H, 0)
    rho_ref = os_.one_body(c)
    assert np.abs(rho - rho_ref).max() < 1e-12, np.abs(rho - rho_ref).max()
    assert np.abs(rho - rho.conj().T).max() < 1e-12
    assert abs(np.trace(rho) - m.n_filled) < 1e-10
    gb.close()
    gc.close()


@pytest.mark.parametrize("tpp", [-0.8, -0.58, -0.2])
def test_odlro_demo(ed, ctx, oracle_mod, tpp):
    """The ODLRO demo of test/bosonic_fci.jl:277-335 on the GPU: Haldane bosons 2x3, 3
    bosons, k = (0,0) ground state; the one-body density matrix from the GPU, transformed
    to rho(k) on the reference's 61 x 61 grid, passes the demo's assertions (finite,
    >= -1e-10, peak at Gamma for the superfluid at t'' = -0.2) and equals the map of the
    oracle's one-body matrix of the same eigenvector (1e-12)."""
    from tests.test_oracle import odlro_checks, odlro_k_map
    m = L.haldane_honeycomb(2, 3, tpp=tpp, n_filled=3)
    G = L.model_group(m)
    ir = L.translation_irreps(G, m.lattice)[0]
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=1)
    k, M = odlro_k_map(m.lattice, ed.off_diagonal_long_range_order(H, 0))
    odlro_checks(k, M, tpp)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, False)
    os_ = oracle_mod.Sector(oc, ir.values, basis=(gb.representative_mask_list, gb.stabilizer_order_list))
    _, M_ref = odlro_k_map(m.lattice, os_.one_body(vecs[:, 0]))
    assert np.abs(M - M_ref).max() < 1e-12
    gb.close()
    gc.close()


@pytest.mark.parametrize("case", range(4))
def test_structure_factor_demo(ed, ctx, oracle_mod, case):
    """The static structure factor of test/bosonic_fci.jl:359-397 (Haldane bosons 2x3 at
    the three t'' of the demo) and of twisted checkerboard fermions: S(q) from the GPU's
    connected density correlations on the reference's 61 x 61 grid is finite and
    non-negative (a covariance), vanishes at Gamma (fixed particle number) to 1e-12, and
    equals the map of the oracle's correlations of the same eigenvector (1e-12)."""
    from tests.test_oracle import SQ_CASES, k_map, sq_checks, sq_model
    m, th = sq_model(*SQ_CASES[case])
    G = L.model_group(m, th)
    ir = L.translation_irreps(G, m.lattice)[0]
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=1)
    Cc, d = ed.density_correlation_matrix(H, 0)
    k, S = k_map(m.lattice, Cc)
    sq_checks(k, S)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    os_ = oracle_mod.Sector(oc, ir.values, basis=(gb.representative_mask_list, gb.stabilizer_order_list))
    C_ref, d_ref = os_.density_correlation(vecs[:, 0])
    _, S_ref = k_map(m.lattice, C_ref - np.outer(d_ref, d_ref))
    assert np.abs(S - S_ref).max() < 1e-12
    gb.close()
    gc.close()


@pytest.mark.parametrize("name", ["hald2x3", "hub2x3", "checker3x2_f_flux", "checker3x2_b_flux", "hald3x2_flux"])
def test_matrix_mode_block_matches_oracle(ed, ctx, oracle_mod, golden, name):
    """Matrix mode (build_ed_Hamiltonian_symmetry_block) from the GPU tables against the
    oracle's sector matrix (projection table + diagonal, duplicates summed) for two
    momenta, to 1e-13; Hermitian; equal to the matrix-free operator on a random vector;
    for Haldane 2x3 at k=(0,0) the stored nnz of the reference's block (756)."""
    m, G = make(name)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    irs = L.translation_irreps(G, m.lattice)
    rng = np.random.default_rng(3)
    for ir in (irs[0], irs[len(irs) // 2]):
        gb = ed.build_symmetry_sector_basis(gc, ir)
        print(f"[matrix mode] {name} k={ir.label} dim={gb.dim}", flush=True)
        A = ed.build_ed_Hamiltonian_symmetry_block(gb, m.bilinear_terms, m.density_terms)
        os_ = oracle_mod.Sector(oc, ir.values)
        os_.set_terms(*L.term_arrays(m))
        B = os_.sparse().tocsc()
        B.eliminate_zeros()
        scale = max(abs(B).max(), 1.0)
        assert abs(A - B).max() <= 1e-13 * scale
        # entries that cancel to zero in one summation order can be ~1e-17 in another
        # (twisted phases): equal patterns above the tolerance, exact nnz without twist
        big = lambda M: int((abs(M.tocoo().data) > 1e-13 * scale).sum())
        assert big(A) == big(B)
        if all(abs(p - 1) == 0 for p in G.phases.reshape(-1)):
            assert A.nnz == B.nnz
        assert abs(A - A.conj().T).max() <= 1e-13 * scale
        H = ed.MatrixFreeHamiltonian.__new__(ed.MatrixFreeHamiltonian)
        H.basis, H.n, H.precision = gb, gb.dim, ed.PREC_FP64
        x = rng.standard_normal(gb.dim) + 1j * rng.standard_normal(gb.dim)
        assert np.abs(A @ x - H(x)).max() <= 1e-12 * np.abs(A @ x).max()
        if name == "hald2x3" and ir is irs[0]:
            assert A.shape[0] == golden["design_notebook"]["haldane_2x3"]["sector_dim"]
            assert A.nnz == golden["design_notebook"]["haldane_2x3"]["nnz_H"] == 756
        gb.close()
    gc.close()


@pytest.mark.parametrize("name", ["checker3x2_f_flux", "checker3x2_b_flux", "hub2x3", "hald3x2_flux"])
def test_fock_expansion_and_position_operator_match_oracle(ed, ctx, oracle_mod, name):
    """Sector -> Fock expansion and the projected position operator on the GPU against
    the oracle's restatements (pinned against the Fock-space embedding) to 1e-12:
    same masks in ascending order, same amplitudes; same rows and values per column."""
    m, G = make(name)
    irs = L.translation_irreps(G, m.lattice)
    print(f"[fock/position] {name}", flush=True)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    gb0, gb1 = ed.build_symmetry_sector_basis(gc, irs[1 % len(irs)]), ed.build_symmetry_sector_basis(gc, irs[0])
    os0 = oracle_mod.Sector(oc, irs[1 % len(irs)].values)
    os1 = oracle_mod.Sector(oc, irs[0].values)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(gb0.dim) + 1j * rng.standard_normal(gb0.dim)
    masks, amps = ed.fock_amplitudes(c, gb0)
    ref = os0.fock_expand(c)
    nz = np.nonzero(ref)[0]
    assert len(masks) == len(nz) and np.all(np.diff(masks.astype(np.float64)) > 0)
    assert all(oracle_mod.rank_combination(int(x)) == int(k) for x, k in zip(masks, nz))
    assert np.abs(amps - ref[nz]).max() < 1e-12
    d = ed.expand_sector_state_to_fock_amplitudes(c, gb0)
    assert len(d) == len(nz) and abs(sum(abs(v) ** 2 for v in d.values()) - 1) < 1e-12
    phi = np.exp(2j * np.pi * rng.random(m.n_site))
    P = ed.position_operator_matrix(gb1, gb0, phi)
    rows, vals = os1.position_operator(os0, phi)
    dense = np.zeros((gb1.dim, gb0.dim), dtype=np.complex128)
    for col, (r, v) in enumerate(zip(rows, vals)):
        if r > 0 and abs(v) > 1e-13:
            dense[r - 1, col] = v
    assert np.abs(P.toarray() - dense).max() < 1e-12
    gb0.close()
    gb1.close()
    gc.close()


@pytest.mark.parametrize("name,th0,th1", [("checker3x2_f_flux", (0.0, 0.0), (0.23, 0.17)),
                                          ("hald3x2_flux", (0.1, 0.0), (0.4, 0.0)),
                                          ("checker3x2_b_flux", (0.3, 0.3), (0.0, 0.0))])
def test_flux_refresh_of_a_device_catalog(ed, ctx, oracle_mod, name, th0, th1):
    """sed_catalog_update_phases (update_orbit_stabilizer_phases!) on a catalog built at
    theta0 reproduces the oracle catalog built at theta1 (tables bit-exact, amplitudes and
    stabilizer phases to 1e-13); a sector built after the refresh solves theta1, one built
    before it still solves theta0 (1e-10 against the oracle)."""
    f = CASES[name][0]
    m0, m1 = L.twisted(f(), th0), L.twisted(f(), th1)
    G0, G1 = L.model_group(m0, th0), L.model_group(m1, th1)
    ferm = m0.statistics == L.FERMIONIC
    print(f"[flux refresh] {name} {th0} -> {th1}", flush=True)
    gc = ed.build_symmetry_orbit_catalog(n_site=m0.n_site, n_filled=m0.n_filled, symmetry_group=G0,
                                         particle_statistics=m0.statistics, ctx=ctx)
    ir0 = L.translation_irreps(G0, m0.lattice)[1]
    old = ed.build_symmetry_sector_basis(gc, ir0)
    Hold, nold = ed.hamiltonian_linear_operator(old, m0.bilinear_terms, m0.density_terms)
    ed.update_orbit_stabilizer_phases(gc, G1)
    oc1 = oracle_mod.Catalog(m1.n_site, m1.n_filled, G1, ferm)
    a, b = gc.tables(), oc1.tables()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.abs(a[2] - b[2]).max() < 1e-13
    ga, ob = gc.orbits(), oc1.orbits()
    assert np.array_equal(ga[2], ob[2]) and np.abs(ga[3] - ob[3]).max() < 1e-13
    for mm, G, ir, basis in ((m1, G1, L.translation_irreps(G1, m1.lattice)[1], None), (m0, G0, ir0, old)):
        if basis is None:
            basis = ed.build_symmetry_sector_basis(gc, ir)
            H, n = ed.hamiltonian_linear_operator(basis, mm.bilinear_terms, mm.density_terms)
        else:
            H, n = Hold, nold
        oc = oracle_mod.Catalog(mm.n_site, mm.n_filled, G, ferm)
        os_ = oracle_mod.Sector(oc, ir.values)
        os_.set_terms(*L.term_arrays(mm))
        ref = oracle_mod.lowest_eigenvalues(os_, 3)
        vals = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
        assert np.abs(vals - ref).max() <= 1e-10 * np.abs(ref).max()
        basis.close()
    gc.close()


def test_flux_scan_reusing_one_catalog_matches_rebuilds(ed, ctx):
    """ed_scan_flux with one catalog refreshed per twist gives the eigenvalues of a catalog
    rebuilt per twist (the reference's loop), for every (twist, momentum)."""
    base = L.haldane_honeycomb(3, 2, n_filled=3, V1=0.3)
    thetas = [(i / 5, 0.1 * i) for i in range(5)]
    a = ed.ed_scan_flux(base, thetas, nev=2, ctx=ctx, reuse_catalog=True)
    b = ed.ed_scan_flux(base, thetas, nev=2, ctx=ctx, reuse_catalog=False)
    assert a.keys() == b.keys() and len(a) == 5 * 6
    for k in a:
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * np.abs(b[k]).max(), k


@pytest.mark.parametrize("name", ["hub2x3", "checker3x2_f_flux", "hald3x2_flux"])
def test_device_matvec_all_precisions(ed, ctx, oracle_mod, name):
    """sed_matvec_device on device buffers (hipMalloc through the HIP runtime libsed itself
    links; complex double for FP64, complex float for FP32), twisted and untwisted
    sectors, against the oracle's matvec: 1e-12 (FP64) and 5e-6 (FP32) relative."""
    import ctypes as C
    from realspace_exactdiagonalization_amd import _lib
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    m, G = make(name)
    ir = L.translation_irreps(G, m.lattice)[1]
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(m))
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, ir)
    gb.set_terms(m.bilinear_terms, m.density_terms)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(gb.dim) + 1j * rng.standard_normal(gb.dim)
    ref = os_.apply_table(x, 1)
    print(f"[device matvec] {name} dim={gb.dim}", flush=True)
    for prec, dt, tol in ((ed.PREC_FP64, np.complex128, 1e-12), (ed.PREC_FP32, np.complex64, 5e-6)):
        xh = np.ascontiguousarray(x.astype(dt))
        yh = np.zeros_like(xh)
        xd, yd = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(xd), xh.nbytes) == 0 and hip.hipMalloc(C.byref(yd), yh.nbytes) == 0
        try:
            assert hip.hipMemcpy(xd, xh.ctypes.data, xh.nbytes, 1) == 0         # host to device
            assert hip.hipDeviceSynchronize() == 0
            _lib.check(_lib.lib().sed_matvec_device(gb._h, xd, yd, prec))
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(yh.ctypes.data, yd, yh.nbytes, 2) == 0         # device to host
        finally:
            hip.hipFree(xd)
            hip.hipFree(yd)
        y = yh.astype(np.complex128)
        assert np.abs(y - ref).max() <= tol * np.abs(ref).max(), (prec, np.abs(y - ref).max())
    gb.close()
    gc.close()


def test_ed_scan_sectors_filter_resume_and_flux_mode(ed, ctx, oracle_mod):
    """ed_scan (ed_scan!, src/symmetry_resolved_ed.jl:1199-1356): every momentum sector of a
    twisted fermion model solved, each equal to the oracle's lowest eigenvalues (1e-10); a
    scanned_sectors filter solves only those sectors; a repeated call with the same nev
    recomputes nothing (resume semantics); matrix mode gives the matrix-free eigenvalues;
    flux mode returns per-twist data whose eigenvalues equal the batched ed_scan_flux."""
    base = L.checkerboard(3, 2, n_filled=4, V1=1.0, V2=0.5)
    th = (0.23, 0.17)
    m = L.twisted(base, th)
    G = L.model_group(m, th)
    data = ed.build_ed_data(m, symmetry_group=G, ctx=ctx)
    assert ed.ed_scan(data, nev=3, mode="matrixfree") is data
    oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, True)
    for idx, ir in enumerate(data.irrep_list):
        os_ = oracle_mod.Sector(oc, ir.values)
        vals = data.ed_scan_res[idx][0]
        assert data.sector_dims[idx] == os_.dim
        if os_.dim:
            os_.set_terms(*L.term_arrays(m))
            ref = oracle_mod.lowest_eigenvalues(os_, min(3, os_.dim))
            assert np.abs(vals - ref).max() <= 1e-10 * np.abs(ref).max()
    d2 = ed.build_ed_data(m, symmetry_group=G, ctx=ctx)
    ed.ed_scan(d2, nev=2, mode="matrix", scanned_sectors=[(0, 0), (2, 1)])
    labels = {tuple(d2.irrep_list[i].label) for i in d2.ed_scan_res}
    assert labels == {(0, 0), (2, 1)}
    kept = {i: d2.ed_scan_res[i] for i in d2.ed_scan_res}
    ed.ed_scan(d2, nev=2, scanned_sectors=[(0, 0), (2, 1)])
    assert all(d2.ed_scan_res[i] is kept[i] for i in kept)
    for i in kept:
        assert np.abs(kept[i][0] - data.ed_scan_res[i][0][:2]).max() <= 1e-10 * np.abs(kept[i][0]).max()
    with pytest.raises(ValueError):
        ed.ed_scan(d2, nev=2, mode="arpack")
    thetas = [0.0, 0.25, 0.5]
    b0 = ed.build_ed_data(base, ctx=ctx)
    per = ed.ed_scan(b0, nev=2, mode="matrixfree", twisted_phases_over_2pi_list=thetas, ctx=ctx)
    ref = ed.ed_scan_flux(base, [(t, 0.0) for t in thetas], nev=2, ctx=ctx)
    assert len(per) == len(thetas)
    for t, d in zip(thetas, per):
        assert d.model.theta == (t, 0.0)
        for idx, ir in enumerate(d.irrep_list):
            if d.sector_dims[idx] == 0:
                continue
            a = d.ed_scan_res[idx][0]
            r = ref[((t, 0.0), tuple(ir.label))]
            assert np.abs(a - r[:len(a)]).max() <= 1e-10 * np.abs(r).max()


def test_checkpointed_gpu_scan_resumes(ed, ctx, tmp_path):
    """A GPU flux scan with per-twist checkpoint files (ed_scan! flux mode with
    checkpoints, src/symmetry_resolved_ed.jl:1320-1355): a first run limited to two momentum
    sectors, as a job stopped partway would leave it, then the full run over the same
    directory resumes each twist (resume_ed_scan!, :1266-1277) and ends with the
    eigenvalues of an uninterrupted scan; the stored eigenvectors are the GPU's."""
    base = L.checkerboard(3, 2, n_filled=4, V1=1.0, V2=0.5)
    thetas = [0.0, 0.3]
    ck = str(tmp_path / "ck")
    full = ed.ed_scan(ed.build_ed_data(base, ctx=ctx), nev=2, mode="matrixfree",
                      twisted_phases_over_2pi_list=thetas, ctx=ctx)
    part = ed.ed_scan(ed.build_ed_data(base, ctx=ctx), nev=2, mode="matrixfree", scanned_sectors=[(0, 0), (1, 1)],
                      twisted_phases_over_2pi_list=thetas, ctx=ctx, checkpoint_dir=ck)
    assert all(sorted(d.ed_scan_res) == [0, 3] for d in part)
    out = ed.ed_scan(ed.build_ed_data(base, ctx=ctx), nev=2, mode="matrixfree",
                     twisted_phases_over_2pi_list=thetas, ctx=ctx, checkpoint_dir=ck)
    for d, f, p in zip(out, full, part):
        assert d.checkpoint_path == p.checkpoint_path and os.path.isfile(d.checkpoint_path)
        for i in range(len(d.irrep_list)):
            a, b = d.ed_scan_res[i][0], f.ed_scan_res[i][0]
            assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max())
        for i in (0, 3):   # carried over from the partial run, not recomputed
            assert np.array_equal(d.ed_scan_res[i][1], p.ed_scan_res[i][1])
