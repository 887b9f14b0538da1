"""Groups whose site phases are not of gauge form, through the C ABI, against the oracle.

The wheel (tests/wheel_model.py): C_N rotations of a ring around a centre site that
every rotation fixes, with the phase cis(2 pi m j / N) on the centre.  The oracle is
pinned on these groups by brute force (tests/test_oracle.py: the operators form a
representation commuting with H and the union of the oracle's sector spectra is the full
Fock-space spectrum).  Here the GPU path, which takes every hop's amplitude from the
catalog's per-rank table, must give the oracle's tables (integer tables bit-exact,
amplitudes and stabilizer phases to 1e-13), sector bases (bit-exact), H x (1e-12 FP64,
2e-6 FP32), eigenvalues (1e-10; dense path and Lanczos), matrix-mode CSC (1e-13),
observables and Fock expansions (1e-12), a phase refresh equal to a rebuild, and the
same eigenvalues through the server.
"""
import numpy as np
import pytest

from realspace_exactdiagonalization_amd import _lib
from realspace_exactdiagonalization_amd import lattice as L
from tests.test_gpu_parity import assert_lowest_levels
from tests.wheel_model import wheel

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


# (ring sites N, centre angular momentum m, particles k, statistics): C(N+1, k)/N rows per
# sector: 35/6 and 462/10 (dense path), 24310/16 (Lanczos).  k = N/2 + 1 lets the centre
# and every second ring site form an orbit fixed by the even rotations, whose stabilizer
# phases cis(2 pi m j / N) are complex.
CASES = {
    "wheel6_m1_b": (6, 1, 4, L.BOSONIC),
    "wheel6_m2_f": (6, 2, 4, L.FERMIONIC),
    "wheel10_m3_f": (10, 3, 6, L.FERMIONIC),
    "wheel16_m5_f": (16, 5, 9, L.FERMIONIC),
    "wheel16_m1_b": (16, 1, 9, L.BOSONIC),
}


def make(name):
    N, m, k, st = CASES[name]
    return wheel(N, m, k, st)


def catalogs(ed, ctx, oracle_mod, model, G):
    gc = ed.build_symmetry_orbit_catalog(n_site=model.n_site, n_filled=model.n_filled, symmetry_group=G,
                                         particle_statistics=model.statistics, ctx=ctx)
    oc = oracle_mod.Catalog(model.n_site, model.n_filled, G, model.statistics == L.FERMIONIC)
    return gc, oc


@pytest.mark.parametrize("name", list(CASES))
def test_catalog_tables_match_oracle(ed, ctx, oracle_mod, name):
    model, G, _ = make(name)
    print(f"[general catalog] {name} N={model.n_site} k={model.n_filled}", flush=True)
    gc, oc = catalogs(ed, ctx, oracle_mod, model, G)
    assert gc.n_total == oc.n_total and gc.n_orbits == oc.n_orbits
    orr, ors, ora = oc.tables()
    grr, grs, gra = gc.tables()
    assert np.array_equal(orr, grr) and np.array_equal(ors, grs)
    assert np.abs(ora - gra).max() < 1e-13
    oreps, oso, ogi, ogp = oc.orbits()
    greps, gso, ggi, ggp = gc.orbits()
    assert np.array_equal(oreps, greps) and np.array_equal(oso, gso) and np.array_equal(ogi, ggi)
    assert np.abs(ogp - ggp).max() < 1e-13
    # the case is one the sign-only tables cannot hold: complex stabilizer phases
    assert np.abs(ggp[ggi > 0].imag).max() > 0.5
    gc.close()


@pytest.mark.parametrize("name", list(CASES))
def test_sector_bases_matvec_and_eigenvalues_match_oracle(ed, ctx, oracle_mod, name):
    model, G, irreps = make(name)
    print(f"[general sectors] {name}", flush=True)
    gc, oc = catalogs(ed, ctx, oracle_mod, model, G)
    rng = np.random.default_rng(11)
    terms = L.term_arrays(model)
    total = 0
    for ir in irreps:
        os_ = oracle_mod.Sector(oc, ir.values)
        gb = ed.build_symmetry_sector_basis(gc, ir)
        assert gb.dim == os_.dim, ir.label
        total += gb.dim
        if gb.dim == 0:
            gb.close()
            continue
        oreps, ostab = os_.basis()
        assert np.array_equal(oreps, gb.representative_mask_list)
        assert np.array_equal(ostab, gb.stabilizer_order_list)
        os_.set_terms(*terms)
        H, n = ed.hamiltonian_linear_operator(gb, model.bilinear_terms, model.density_terms)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y_ref = os_.apply_table(x, 1)
        scale = np.abs(y_ref).max()
        assert np.abs(H(x) - y_ref).max() <= 1e-12 * scale, ir.label
        H.precision = ed.PREC_FP32
        assert np.abs(H(x) - y_ref).max() <= 2e-6 * scale, ir.label
        H.precision = ed.PREC_FP64
        nev = min(3, n)
        ref = oracle_mod.lowest_eigenvalues(os_, min(nev + 3, n))
        for prec in (ed.PREC_FP64, ed.PREC_FP32_REFINE):
            vals = ed.diagonalize_block_matrixfree(H, n, nev=nev, precision=prec, return_vectors=False)[0]
            assert_lowest_levels(np.asarray(vals), ref, n, f"{name} k={ir.label} prec={prec}")
        gb.close()
    assert total == gc.n_total
    gc.close()


def test_launch_shapes_and_refusals(ed, ctx, oracle_mod):
    """A sector of the wheel runs one launch shape (shape 1; the others are refused with
    SED_EUNSUPPORTED and the autotuner skips them); site phases that are not a
    representation (one rotation's phase on a ring site changed, its inverse's matching
    phase with it), or that would need hop amplitudes for 2^31 ranks or more, are refused at
    catalog build."""
    model, G, irreps = make("wheel16_m5_f")
    gc, oc = catalogs(ed, ctx, oracle_mod, model, G)
    gb = ed.build_symmetry_sector_basis(gc, irreps[3])
    H, n = ed.hamiltonian_linear_operator(gb, model.bilinear_terms, model.density_terms)
    os_ = oracle_mod.Sector(oc, irreps[3].values)
    os_.set_terms(*L.term_arrays(model))
    x = np.arange(1, n + 1) + 1j
    y_ref = os_.apply_table(x, 1)
    n_shapes = len(gb.shape(ed.PREC_FP64)["tune_ms"])
    ran = 0
    for prec in (ed.PREC_FP64, ed.PREC_FP32):
        H.precision = prec
        for sh in range(n_shapes):
            gb.set_shape(prec, sh, 1)
            try:
                y = H(x)
            except _lib.SedError as e:
                assert e.code == _lib.SED_EUNSUPPORTED and sh != 1, (prec, sh, e)
                continue
            assert sh == 1
            assert np.abs(y - y_ref).max() <= (1e-12 if prec == ed.PREC_FP64 else 2e-6) * np.abs(y_ref).max()
            ran += 1
        gb.set_shape(prec, -1)
    assert ran == 2
    gb.close()
    # a ring site's phase under rotation 3 times z, and its image's under the inverse
    # rotation 13 times conj(z): inverses still match, compositions no longer do
    bad = L.SymmetryGroup(G.name, G.perm.copy(), G.phases.copy(), G.labels, G.identity_idx)
    z = np.exp(0.1j)
    i0 = 0
    bad.phases[3, i0] *= z
    bad.phases[13, G.perm[3, i0] - 1] *= np.conj(z)
    with pytest.raises(_lib.SedError) as e:
        ed.build_symmetry_orbit_catalog(n_site=model.n_site, n_filled=model.n_filled, symmetry_group=bad,
                                        particle_statistics=model.statistics, ctx=ctx)
    assert e.value.code == _lib.SED_EUNSUPPORTED and "representation" in str(e.value)
    # hop amplitudes are not kept for catalogs of 2^31 ranks or more (C(34, 17) here)
    big, Gb, _ = wheel(33, 1, 17, L.BOSONIC)
    with pytest.raises(_lib.SedError) as e:
        ed.build_symmetry_orbit_catalog(n_site=big.n_site, n_filled=big.n_filled, symmetry_group=Gb,
                                        particle_statistics=big.statistics, ctx=ctx)
    assert e.value.code == _lib.SED_EUNSUPPORTED and "2^31" in str(e.value)
    gc.close()


@pytest.mark.parametrize("name", ["wheel6_m2_f", "wheel10_m3_f", "wheel16_m5_f"])
def test_matrix_mode_observables_and_fock_expansion_match_oracle(ed, ctx, oracle_mod, name):
    model, G, irreps = make(name)
    print(f"[general observables] {name}", flush=True)
    gc, oc = catalogs(ed, ctx, oracle_mod, model, G)
    ir = irreps[1]
    gb = ed.build_symmetry_sector_basis(gc, ir)
    os_ = oracle_mod.Sector(oc, ir.values)
    os_.set_terms(*L.term_arrays(model))
    A = ed.build_ed_Hamiltonian_symmetry_block(gb, model.bilinear_terms, model.density_terms)
    B = os_.sparse().tocsc()
    scale = max(abs(B).max(), 1.0)
    assert abs(A - B).max() <= 1e-13 * scale
    assert abs(A - A.conj().T).max() <= 1e-13 * scale
    H, n = ed.hamiltonian_linear_operator(gb, model.bilinear_terms, model.density_terms)
    vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=2)
    c = vecs[:, 0]
    osb = oracle_mod.Sector(oc, ir.values, basis=(gb.representative_mask_list, gb.stabilizer_order_list))
    C_ref, d_ref = osb.density_correlation(c)
    C, d = ed.density_correlation_matrix(H, 0, subtract_disconnected=False)
    assert np.abs(d - d_ref).max() < 1e-12 and np.abs(C - C_ref).max() < 1e-12
    rho = ed.off_diagonal_long_range_order(H, 0)
    assert np.abs(rho - osb.one_body(c)).max() < 1e-12
    assert abs(np.trace(rho) - model.n_filled) < 1e-10
    masks, amps = ed.fock_amplitudes(c, gb)
    ref = os_.fock_expand(c)
    nz = np.nonzero(ref)[0]
    assert len(masks) == len(nz)
    assert np.abs(amps - ref[nz]).max() < 1e-12
    gb.close()
    gc.close()


def test_phase_refresh_equals_rebuild_and_server(ed, ctx, oracle_mod):
    """update_orbit_stabilizer_phases from m = 1 to m = 3 on the device catalog: tables and
    stabilizer phases equal the oracle catalog built at m = 3, a sector built afterwards
    solves m = 3 and one built before still solves m = 1 (its own amplitude table); the
    server solves a wheel sector (lane route: the packed engine does not take per-rank
    amplitudes) to the same eigenvalues."""
    N, k, st = 16, 9, L.FERMIONIC
    m1, G1, irs = wheel(N, 1, k, st)
    m3, G3, _ = wheel(N, 3, k, st)
    gc = ed.build_symmetry_orbit_catalog(n_site=m1.n_site, n_filled=k, symmetry_group=G1, particle_statistics=st,
                                         ctx=ctx)
    old = ed.build_symmetry_sector_basis(gc, irs[2])
    Hold, nold = ed.hamiltonian_linear_operator(old, m1.bilinear_terms, m1.density_terms)
    ed.update_orbit_stabilizer_phases(gc, G3)
    oc3 = oracle_mod.Catalog(m3.n_site, k, G3, True)
    a, b = gc.tables(), oc3.tables()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.abs(a[2] - b[2]).max() < 1e-13
    ga, ob = gc.orbits(), oc3.orbits()
    assert np.array_equal(ga[2], ob[2]) and np.abs(ga[3] - ob[3]).max() < 1e-13
    srv = ed.SectorServer(ctx)
    for mm, G, basis in ((m3, G3, None), (m1, G1, old)):
        if basis is None:
            basis = ed.build_symmetry_sector_basis(gc, irs[2])
            H, n = ed.hamiltonian_linear_operator(basis, mm.bilinear_terms, mm.density_terms)
        else:
            H, n = Hold, nold
        oc = oracle_mod.Catalog(mm.n_site, k, G, True)
        os_ = oracle_mod.Sector(oc, irs[2].values)
        assert os_.dim == n
        os_.set_terms(*L.term_arrays(mm))
        ref = oracle_mod.lowest_eigenvalues(os_, 6)
        vals = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
        assert_lowest_levels(np.asarray(vals), ref, n, f"refresh m={mm.name}")
        rid = srv.submit(basis, 3)
        srv.run(0)
        status, ev, _, _ = srv.result(rid)
        assert status == 0 and np.abs(ev - vals).max() <= 1e-10 * np.abs(vals).max()
    srv.close()
    old.close()
    gc.close()
