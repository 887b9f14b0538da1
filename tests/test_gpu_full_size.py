"""The benchmark workloads at their full sizes, through the C ABI.

Where the oracle cannot follow (sectors of 4M and 15.7M states), parity rests on
size-independent properties: sector dimensions against the character formula (closed
form C1) and their sum against C(N, k); the matvec's Hermiticity and linearity; FP32
against FP64; FP32+refine against FP64 eigenvalues.  Config[0] (XXZ N=16, dim 810) is
small enough for the oracle's dense spectrum.
"""
import math

import numpy as np
import pytest

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
