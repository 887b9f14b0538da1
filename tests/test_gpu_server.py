"""The continuous-batching server on the serving-stream request families, through the C ABI.

* A seeded stream of XXZ chains (N = 20, 22), 2 x 6 Hubbard ladders at quarter filling and
  checkerboard lattices (3 x 4, 3 x 5) at 1/3 filling of the lower band, random momentum and
  twist, is fed to a server by a thread while earlier requests run (continuous batching):
  every result equals the oracle's lowest levels (1e-10 relative), in FP64 and FP32+refine,
  and equals the same request solved alone (1e-10 relative).
* Admission by HBM footprint: a server with a small budget never charges more than the
  budget, the device's used memory stays within it (sampled while the stream runs), every
  result is still right, and a request larger than the budget fails with SED_ENOMEM.
"""
import ctypes as C
import threading
import time

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import _lib
from realspace_exactdiagonalization_amd import lattice as L
from tests.test_gpu_parity import assert_lowest_levels

pytestmark = pytest.mark.gpu

SIZES = {"xxz": [20, 22], "ladder": [6], "checker": [(3, 4), (3, 5)]}


@pytest.fixture(scope="module")
def ed():
    from realspace_exactdiagonalization_amd import ed as E
    return E


@pytest.fixture(scope="module")
def ctx(ed):
    c = ed.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def stream(ed, ctx, oracle_mod):
    """Sectors of 12 seeded stream requests (built on the GPU) with the oracle's levels."""
    reqs = L.stream_requests(12, 77, SIZES)
    out = []
    for r in reqs:
        m, G, ir = L.stream_sector(r)
        oc = oracle_mod.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
        os_ = oracle_mod.Sector(oc, ir.values)
        if os_.dim == 0:
            continue
        os_.set_terms(*L.term_arrays(m))
        nev = min(3, os_.dim)
        ref = oracle_mod.lowest_eigenvalues(os_, min(nev + 3, os_.dim))
        cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                              particle_statistics=m.statistics, ctx=ctx)
        b = ed.build_symmetry_sector_basis(cat, ir)
        assert b.dim == os_.dim
        b.set_terms(m.bilinear_terms, m.density_terms)
        print(f"[stream] {r['family']} {r['size']} k={ir.label} theta={r['theta']} dim={b.dim}", flush=True)
        out.append({"req": r, "cat": cat, "b": b, "nev": nev, "ref": ref})
    assert {s["req"]["family"] for s in out} == set(SIZES)
    yield out
    for s in out:
        s["b"].close()
        s["cat"].close()


def solve_alone(ed, b, nev, precision):
    H = ed.MatrixFreeHamiltonian.__new__(ed.MatrixFreeHamiltonian)
    H.basis, H.n, H.precision = b, b.dim, ed.PREC_FP64
    return ed.diagonalize_block_matrixfree(H, b.dim, nev=nev, precision=precision, return_vectors=False)[0]


def device_used():
    """Bytes in use on the current device (hipMemGetInfo: total - free)."""
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


def feed(srv, stream, precision, gap_s, ids):
    """Submit every request, gap_s apart, from a thread while the server runs."""
    def run():
        for i, s in enumerate(stream):
            ids[i] = srv.submit(s["b"], s["nev"], precision=precision)
            time.sleep(gap_s)
    th = threading.Thread(target=run)
    th.start()
    return th


@pytest.mark.parametrize("precision", [0, 2])
def test_stream_results_match_oracle_alone_and_in_the_mix(ed, ctx, stream, precision):
    alone = [solve_alone(ed, s["b"], s["nev"], precision) for s in stream]
    for s, a in zip(stream, alone):
        assert_lowest_levels(a, s["ref"], s["b"].dim, ("alone", s["req"]["family"], s["req"]["size"]))
    srv = ed.SectorServer(ctx, max_slots=4)
    ids = [None] * len(stream)
    th = feed(srv, stream, precision, 0.002, ids)
    th.join()
    srv.run(0)
    for s, a, rid in zip(stream, alone, ids):
        st, ev, stats, lat = srv.result(rid)
        assert st == 0, (s["req"], st)
        what = ("mix", s["req"]["family"], s["req"]["size"], precision)
        assert_lowest_levels(ev, s["ref"], s["b"].dim, what)
        assert np.abs(ev - a).max() <= 1e-10 * np.abs(a).max(), what
        assert lat >= 0
    srv.close()


def test_admission_keeps_within_the_hbm_budget(ed, ctx, stream):
    """A 160 MiB budget (a quarter of it the packed engine's arena) over 8 lanes: the
    server's charge never exceeds the budget, device memory in use (sampled every 2 ms
    while the stream runs) grows by at most the budget plus 64 MiB of allocation
    granularity, the results are right, and an XXZ N = 26 sector (footprint ~0.3 GB)
    fails at submission with SED_ENOMEM while the rest still complete."""
    budget = 160 << 20
    base = device_used()
    big_m = L.heisenberg_chain(26)
    big_G = L.model_group(big_m)
    big_cat = ed.build_symmetry_orbit_catalog(n_site=26, n_filled=13, symmetry_group=big_G,
                                              particle_statistics=big_m.statistics, ctx=ctx)
    big = ed.build_symmetry_sector_basis(big_cat, L.translation_irreps(big_G, big_m.lattice)[0])
    big.set_terms(big_m.bilinear_terms, big_m.density_terms)
    base = device_used()
    for precision in (ed.PREC_FP32_REFINE, ed.PREC_FP64):
        srv = ed.SectorServer(ctx, hbm_budget_bytes=budget, max_slots=8)
        mem0 = srv.memory()
        assert mem0["budget"] == budget and 0 < mem0["fixed"] < budget // 2
        rid_big = srv.submit(big, 3, precision=precision)
        ids = [None] * len(stream)
        th = feed(srv, stream, precision, 0.0, ids)
        peak_dev = 0
        while True:
            peak_dev = max(peak_dev, device_used() - base)
            th.join(0)
            if not th.is_alive() and srv.run(-1) >= len(stream) + 1:
                break
            time.sleep(0.002)
        srv.run(0)
        mem = srv.memory()
        print(f"[admission] precision={precision} budget={budget >> 20} MiB fixed={mem['fixed'] >> 20} MiB "
              f"peak charge={mem['peak'] >> 20} MiB peak device use={peak_dev >> 20} MiB", flush=True)
        assert mem["peak"] <= budget
        assert peak_dev <= budget + (64 << 20)
        st, ev, _, _ = srv.result(rid_big)
        assert st == _lib.SED_ENOMEM
        for s, rid in zip(stream, ids):
            st, ev, _, _ = srv.result(rid)
            assert st == 0, (s["req"], st)
            assert_lowest_levels(ev, s["ref"], s["b"].dim, ("budget", s["req"]["family"], s["req"]["size"]))
        srv.close()
    big.close()
    big_cat.close()


def test_same_sector_never_runs_twice_at_once(ed, ctx, stream):
    """Requests on one sector object, submitted back to back while every lane is busy (so
    small sectors take the packed engine), run one after another and all give the same
    eigenvalues."""
    small = min(stream, key=lambda s: s["b"].dim if s["b"].dim > 500 else 1 << 40)
    others = [s for s in stream if s is not small]
    srv = ed.SectorServer(ctx, max_slots=2)
    busy = [srv.submit(s["b"], s["nev"]) for s in others[:4]]
    ids = [srv.submit(small["b"], small["nev"]) for _ in range(4)]
    srv.run(0)
    res = [srv.result(r) for r in ids]
    for st, ev, stats, lat in res:
        assert st == 0
        assert np.abs(ev - res[0][1]).max() <= 1e-12 * np.abs(res[0][1]).max()
    for r in busy:
        assert srv.result(r)[0] == 0
    srv.close()
