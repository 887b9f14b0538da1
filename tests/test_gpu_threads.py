"""The C ABI's threading contract: distinct handles on distinct host threads are
independent.  Four host threads, each with its own context, catalog and sector of a
different model, build the tables, apply H and solve for the lowest levels at the same
time (ctypes releases the interpreter lock for the duration of every library call, so the
calls overlap in the library and on the device); each thread goes through its model three
times.  Every sector and dimension equals the
oracle's, H x equals the oracle's _apply_table! (1e-12 relative), the eigenvalues equal the
oracle's lowest levels (1e-10 relative, FP64 and FP32+refine), and all of it equals the same
calls made one after another on one thread: H x bitwise and the eigenvalues bitwise (launch
shapes are fixed first, since the autotuner's timings would see the other threads' work)."""
import threading

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import lattice as L
from tests.test_gpu_parity import assert_lowest_levels

pytestmark = pytest.mark.gpu

# name: (model, twist, momentum index)
MODELS = {
    "xxz18": (lambda: L.heisenberg_chain(18), (0.0, 0.0), 0),
    "hald3x4_flux": (lambda: L.haldane_honeycomb(3, 4, n_filled=6), (0.3, 0.1), 1),
    "checker3x4_f": (lambda: L.checkerboard(3, 4, n_filled=8), (0.0, 0.0), 0),
    "hub2x4_flux": (lambda: L.hubbard_square(4, 2, U=4.0, n_filled=6), (0.2, 0.0), 1),
}


@pytest.fixture(scope="module")
def ed():
    from realspace_exactdiagonalization_amd import ed as E
    return E


@pytest.fixture(scope="module")
def cases(oracle_mod):
    """Per model: the twisted model, its group and irrep, the oracle's dimension, H x of a
    seeded vector and lowest levels."""
    out = {}
    for name, (f, th, k) in MODELS.items():
        m = L.twisted(f(), th)
        G = L.model_group(m, th)
        ir = L.translation_irreps(G, m.lattice)[k]
        oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
        os_ = oracle_mod.Sector(oc, ir.values)
        os_.set_terms(*L.term_arrays(m))
        rng = np.random.default_rng(len(name))
        x = rng.standard_normal(os_.dim) + 1j * rng.standard_normal(os_.dim)
        out[name] = {"m": m, "G": G, "ir": ir, "dim": os_.dim, "x": x, "hx": os_.apply_table(x, 1),
                     "levels": oracle_mod.lowest_eigenvalues(os_, 6)}
        print(f"[threads] {name} dim={os_.dim}", flush=True)
    return out


def work(ed, c, out):
    """One model on its own context: catalog, sector, H x, FP64 and FP32+refine solves."""
    ctx = ed.Context(0)
    m = c["m"]
    cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=c["G"],
                                          particle_statistics=m.statistics, ctx=ctx)
    b = ed.build_symmetry_sector_basis(cat, c["ir"])
    H, n = ed.hamiltonian_linear_operator(b, m.bilinear_terms, m.density_terms)
    for prec in (ed.PREC_FP64, ed.PREC_FP32):
        b.set_shape(prec, 1, 1)      # fixed launch shape: no timing-dependent choice
    out["dim"] = n
    out["hx"] = H(c["x"])
    out["ev64"] = ed.diagonalize_block_matrixfree(H, n, nev=3, return_vectors=False)[0]
    out["ev32r"] = ed.diagonalize_block_matrixfree(H, n, nev=3, precision=ed.PREC_FP32_REFINE,
                                                   return_vectors=False)[0]
    b.close()
    cat.close()
    ctx.close()


def test_distinct_handles_on_distinct_threads(ed, cases):
    serial = {}
    for name, c in cases.items():
        serial[name] = {}
        work(ed, c, serial[name])
    conc = {name: [{} for _ in range(3)] for name in cases}
    errors = []

    def run(name):
        try:
            for out in conc[name]:
                work(ed, cases[name], out)
        except Exception as e:   # reported below, on the test's thread
            errors.append((name, repr(e)))

    threads = [threading.Thread(target=run, args=(name,)) for name in cases]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for name, c in cases.items():
        s = serial[name]
        for p in conc[name]:
            print(f"[threads] {name}: ev64 {p['ev64']} serial {s['ev64']}", flush=True)
            assert p["dim"] == s["dim"] == c["dim"]
            scale = np.abs(c["hx"]).max()
            assert np.abs(p["hx"] - c["hx"]).max() <= 1e-12 * scale, name
            assert np.array_equal(p["hx"], s["hx"]), name
            for key in ("ev64", "ev32r"):
                assert_lowest_levels(p[key], c["levels"], c["dim"], (name, key))
                assert np.array_equal(p[key], s[key]), (name, key, p[key], s[key])
