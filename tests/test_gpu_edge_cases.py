"""Edge cases of the HIP path through the C ABI: empty and full filling (the catalog's
n_filled in {0, N} branch, src/symmetry_resolved_ed.jl:410-431), empty sectors, the
largest lattices (64 sites: the top bit of a 64-bit mask), and error reporting
(status codes and messages, the library usable after a failed call)."""
import math

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import lattice as L
from realspace_exactdiagonalization_amd import _lib

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


FILLING_CASES = [
    (lambda k: L.checkerboard(3, 2, n_filled=k, V1=1.0, V2=0.5), 0),
    (lambda k: L.checkerboard(3, 2, n_filled=k, V1=1.0, V2=0.5), 12),
    (lambda k: L.haldane_honeycomb(2, 3, n_filled=k, V1=0.3), 0),
    (lambda k: L.haldane_honeycomb(2, 3, n_filled=k, V1=0.3), 12),
    (lambda k: L.hubbard_square(2, 3, U=4.0, n_filled=k), 12),
]


@pytest.mark.parametrize("case", range(len(FILLING_CASES)))
@pytest.mark.parametrize("theta", [(0.0, 0.0), (0.3, 0.2)])
def test_empty_and_full_filling(ed, ctx, oracle_mod, case, theta):
    """n_filled = 0 or N: one orbit; tables equal the oracle's, exactly one momentum sector
    has dimension 1 (the others 0), its energy is the oracle's, every site density is
    n_filled / N, and solving an empty sector returns no eigenvalues."""
    f, k = FILLING_CASES[case]
    m = L.twisted(f(k), theta)
    G = L.model_group(m, theta)
    ferm = m.statistics == L.FERMIONIC
    print(f"[filling] case {case} k={k} theta={theta}", flush=True)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=k, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    oc = oracle_mod.Catalog(m.n_site, k, G, ferm)
    assert gc.n_total == oc.n_total == 1 and gc.n_orbits == oc.n_orbits == 1
    a, b = gc.tables(), oc.tables()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.abs(a[2] - b[2]).max() < 1e-13
    dims = []
    for ir in L.translation_irreps(G, m.lattice):
        gb = ed.build_symmetry_sector_basis(gc, ir)
        os_ = oracle_mod.Sector(oc, ir.values)
        assert gb.dim == os_.dim
        dims.append(gb.dim)
        H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
        vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=2)
        if n == 0:
            assert vals.shape == (0,) and vecs.shape == (0, 0)
        else:
            os_.set_terms(*L.term_arrays(m))
            ref = np.linalg.eigvalsh(os_.dense())
            assert vals.shape == (1,) and abs(vals[0] - ref[0]) <= 1e-12 * max(1.0, abs(ref[0]))
            assert np.abs(ed.sector_density(H, 0) - k / m.n_site).max() < 1e-12
        gb.close()
    assert sorted(dims)[-1] == 1 and sum(dims) == 1
    gc.close()


@pytest.mark.parametrize("k", [2, 3])
def test_64_site_chain_against_free_fermions(ed, ctx, k):
    """64-site XX chain (hard-core bosons, top mask bit in use): the lowest energy over all
    64 momentum sectors equals free fermions with antiperiodic (even k) or periodic (odd
    k) boundary, sum of the k lowest cos(2 pi (n + shift) / 64), to 1e-10.  k = 2 runs the
    dense small-sector path, k = 3 (sectors of ~651 states) the Lanczos path."""
    N = 64
    m = L.heisenberg_chain(N, n_filled=k)
    m.density_terms = []
    G = L.model_group(m)
    gc = ed.build_symmetry_orbit_catalog(n_site=N, n_filled=k, symmetry_group=G, particle_statistics=m.statistics,
                                         ctx=ctx)
    assert gc.n_total == math.comb(N, k)
    # representatives are orbit minima (low bits); images with site 64 occupied are the
    # highest ranks, e.g. the last one (the top k sites): its table entry points at an orbit
    rr, rs, ra = gc.tables()
    assert rr[-1] > 0 and rs[-1] > 0 and abs(ra[-1]) == 1.0
    best, total = np.inf, 0
    for ir in L.translation_irreps(G, m.lattice):
        gb = ed.build_symmetry_sector_basis(gc, ir)
        total += gb.dim
        if gb.dim:
            H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
            best = min(best, ed.diagonalize_block_matrixfree(H, n, nev=1, return_vectors=False)[0][0])
        gb.close()
    shift = 0.0 if k % 2 == 1 else 0.5
    eps = sorted(math.cos(2 * math.pi * (n + shift) / N) for n in range(N))
    assert abs(best - sum(eps[:k])) < 1e-10
    # summed over all momenta the sector dimensions give the whole space, C(64, k)
    assert total == math.comb(N, k)
    gc.close()


def _err(fn):
    with pytest.raises(_lib.SedError) as e:
        fn()
    return e.value.code


def test_error_codes_and_recovery(ed, ctx):
    """Invalid input gives the documented status code and a message; the library keeps
    working afterwards (a valid solve succeeds after every failure)."""
    import ctypes as C
    lib = _lib.lib()
    m = L.checkerboard(3, 2, n_filled=4)
    G = L.model_group(m)

    def valid_solve():
        gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=4, symmetry_group=G,
                                             particle_statistics=m.statistics, ctx=ctx)
        gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(G, m.lattice)[0])
        H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
        v = ed.diagonalize_block_matrixfree(H, n, nev=2, return_vectors=False)[0]
        gb.close()
        gc.close()
        return v

    ref = valid_solve()
    # a permutation that is not one
    bad = L.SymmetryGroup(G.name, G.perm.copy(), G.phases.copy(), G.labels, G.identity_idx)
    bad.perm[1, 0] = bad.perm[1, 1]
    assert _err(lambda: ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=4, symmetry_group=bad,
                                                        particle_statistics=m.statistics, ctx=ctx)) == _lib.SED_EINVAL
    assert "permutation" in lib.sed_last_error().decode()
    # site phases that are not of gauge form f(i)/f(perm(i))
    bad = L.SymmetryGroup(G.name, G.perm.copy(), G.phases.copy(), G.labels, G.identity_idx)
    bad.phases[1, :] = 1j
    bad.phases[G.order - 1, :] = -1j
    code = _err(lambda: ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=4, symmetry_group=bad,
                                                        particle_statistics=m.statistics, ctx=ctx))
    assert code in (_lib.SED_EUNSUPPORTED, _lib.SED_EINVAL)
    # C(N, k) beyond 2^32 - 2 ranks: C(36, 18) = 9,075,135,300
    big = L.heisenberg_chain(36)
    assert _err(lambda: ed.build_symmetry_orbit_catalog(n_site=36, n_filled=18, symmetry_group=L.model_group(big),
                                                        particle_statistics=big.statistics, ctx=ctx)) \
        == _lib.SED_EUNSUPPORTED
    assert "2^32" in lib.sed_last_error().decode()
    # more than 2^31 orbits: the identity group on 34 sites at half filling (C(34, 17) orbits)
    assert _err(lambda: ed.build_symmetry_orbit_catalog(n_site=34, n_filled=17, symmetry_group=L.identity_group(34),
                                                        particle_statistics=big.statistics, ctx=ctx)) \
        == _lib.SED_EUNSUPPORTED
    assert "orbits exceed 2^31" in lib.sed_last_error().decode()
    # terms: out-of-range site, i_from == i_to
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=4, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(G, m.lattice)[0])
    x = np.ones(gb.dim, dtype=np.complex128)
    Hn = ed.MatrixFreeHamiltonian.__new__(ed.MatrixFreeHamiltonian)
    Hn.basis, Hn.n, Hn.precision = gb, gb.dim, ed.PREC_FP64
    assert _err(lambda: Hn(x)) == _lib.SED_ESTATE                       # matvec before set_terms
    assert _err(lambda: ed.sector_density(Hn, 0)) == _lib.SED_ESTATE    # no solve yet
    assert _err(lambda: gb.set_terms([(1, 13, 1.0)], [])) == _lib.SED_EINVAL
    assert _err(lambda: gb.set_terms([(2, 2, 1.0)], [])) == _lib.SED_EINVAL
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    H.precision = 7                                                     # bad precision
    assert _err(lambda: H(x)) == _lib.SED_EINVAL
    gb.close()
    gc.close()
    # a Lanczos-path sector (XXZ N=16, k=0, dim 810): nev beyond the Krylov basis, and
    # no convergence within one restart at an unreachable tolerance
    xm = L.heisenberg_chain(16)
    xg = L.model_group(xm)
    gc = ed.build_symmetry_orbit_catalog(n_site=16, n_filled=8, symmetry_group=xg, particle_statistics=xm.statistics,
                                         ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(xg, xm.lattice)[0])
    H, n = ed.hamiltonian_linear_operator(gb, xm.bilinear_terms, xm.density_terms)
    assert n == 810
    assert _err(lambda: ed.diagonalize_block_matrixfree(H, n, nev=200, return_vectors=False)) == _lib.SED_EINVAL
    assert _err(lambda: ed.diagonalize_block_matrixfree(H, n, nev=3, tol=1e-30, maxiter=1,
                                                        return_vectors=False)) == _lib.SED_ENOCONV
    gb.close()
    gc.close()
    assert np.array_equal(valid_solve(), ref)


def _device_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_calls_release_their_device_buffers(ed, ctx):
    """Regression: the synchronous entry points free the device buffers they allocate per
    call (matvec in both precisions, solve with eigenvectors, site densities, pair
    densities, Fock expansion).  Heisenberg N=24, k=0 (dimension 112,720): ten rounds of
    those calls after two warm-up rounds leave the device's free memory within 24 MiB
    (one round used to keep about 11 MB)."""
    m = L.heisenberg_chain(24)
    G = L.model_group(m)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(G, m.lattice)[0])
    assert gb.dim == 112720
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    x = np.random.default_rng(5).standard_normal(n) + 0.25j

    def one_round():
        H.precision = ed.PREC_FP64
        y64 = H(x)
        H.precision = ed.PREC_FP32
        y32 = H(x)
        H.precision = ed.PREC_FP64
        assert np.abs(y32 - y64).max() <= 1e-5 * np.abs(y64).max()
        vals, vecs = ed.diagonalize_block_matrixfree(H, n, nev=2)
        d = ed.sector_density(H, 0)
        assert abs(d.sum() - m.n_filled) < 1e-9
        ed.density_correlation_matrix(H, 0)
        ed.fock_amplitudes(vecs[:, 0], gb)
        return vals

    ref = one_round()
    one_round()
    free0 = _device_free_bytes()
    for _ in range(10):
        assert np.abs(one_round() - ref).max() < 1e-10
    free1 = _device_free_bytes()
    print(f"[memory] free before {free0 / 2**20:.1f} MiB, after {free1 / 2**20:.1f} MiB", flush=True)
    assert free0 - free1 < 24 * 2**20
    gb.close()
    gc.close()


@pytest.mark.parametrize("name,dim", [("xxz16", 810), ("haldane3x4", 11240)])
def test_event_timed_solve_equals_graph_solve(ed, ctx, capfd, monkeypatch, name, dim):
    """SED_SOLVE_EVENTS=1 launches the steps eagerly with HIP events between their parts
    and prints one timing line per Lanczos loop; the eigenvalues equal the graph-replayed
    solve's to 1e-12, on a sector whose step tail runs in one workgroup (810 rows) and on
    one that runs the multi-workgroup sequence (11,240 rows)."""
    m = L.workload(name)
    G = L.model_group(m)
    gc = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    gb = ed.build_symmetry_sector_basis(gc, L.translation_irreps(G, m.lattice)[0])
    assert gb.dim == dim
    H, n = ed.hamiltonian_linear_operator(gb, m.bilinear_terms, m.density_terms)
    ref = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
    capfd.readouterr()
    monkeypatch.setenv("SED_SOLVE_EVENTS", "1")
    vals = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
    monkeypatch.delenv("SED_SOLVE_EVENTS")
    err = capfd.readouterr().err
    lines = [l for l in err.splitlines() if l.startswith("[sed-solve-events] fp64")]
    assert len(lines) == 1 and f"dim={dim}" in lines[0], err[-2000:]
    fields = dict(kv.split("=") for kv in lines[0].split()[2:])
    assert float(fields["matvec_ms"]) > 0 and float(fields["wall_ms"]) > 0
    assert np.abs(np.asarray(vals) - np.asarray(ref)).max() <= 1e-12 * np.abs(ref).max()
    gb.close()
    gc.close()
