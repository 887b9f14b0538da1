"""sed_matvec_device_noalloc (the matvec whose scratch is a fixed per-GPU device array)
against sed_matvec_device, through the C ABI, on two calls in a row, the second on a
smaller sector, in FP64 and FP32, with and without twisted boundaries; an over-size sector
is refused with SED_EINVAL.  Without a twist both run the same kernels on the same launch
shape and agree bitwise; with a twist the gauge product and the prescale are one kernel
here and two there (rounding of the fused products may differ), so the bar is 1e-13
(FP64) and 1e-6 (FP32) relative to the largest |y|.  The oracle check of
sed_matvec_device itself lives in the parity tests."""
import ctypes as C

import numpy as np
import pytest

from realspace_exactdiagonalization_amd import _lib
from realspace_exactdiagonalization_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ed():
    from realspace_exactdiagonalization_amd import ed as E
    return E


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def sector(ed, ctx, base, theta, k):
    m = L.twisted(base, theta)
    G = L.model_group(m, theta)
    cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                          particle_statistics=m.statistics, ctx=ctx)
    b = ed.build_symmetry_sector_basis(cat, L.translation_irreps(G, m.lattice)[k])
    b.set_terms(m.bilinear_terms, m.density_terms)
    return cat, b


def run(hip, fn, b, xh, prec):
    """fn(sector, x_dev, y_dev, precision) on device copies of xh; returns y on the host."""
    yh = np.zeros_like(xh)
    xd, yd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(xd), xh.nbytes) == 0 and hip.hipMalloc(C.byref(yd), yh.nbytes) == 0
    try:
        assert hip.hipMemcpy(xd, xh.ctypes.data, xh.nbytes, 1) == 0
        assert hip.hipDeviceSynchronize() == 0
        _lib.check(fn(b._h, xd, yd, prec))
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(yh.ctypes.data, yd, yh.nbytes, 2) == 0
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
    return yh


def check(y, ref, prec, b):
    if np.all(b.catalog.symmetry_group.phases == 1):
        assert np.array_equal(y, ref)
        return
    tol = 1e-13 if prec == 0 else 1e-6
    assert np.abs(y - ref).max() <= tol * np.abs(ref).max()


def test_noalloc_matvec_equals_matvec_device(ed, hip):
    ctx = ed.Context(0)
    lib = _lib.lib()
    # first call: checkerboard 3x4, 8 fermions (dim ~61k); second: twisted Haldane 3x4 bosons (~11k)
    cases = [sector(ed, ctx, L.checkerboard(3, 4, n_filled=8), (0.0, 0.0), 0),
             sector(ed, ctx, L.haldane_honeycomb(3, 4), (0.3, 0.1), 2),
             sector(ed, ctx, L.checkerboard(3, 4, n_filled=6), (0.0, 0.21), 1)]
    dims = [b.dim for _, b in cases]
    print(f"[noalloc] dims {dims}", flush=True)
    assert dims[0] > dims[1] and dims[1] > 0 and dims[2] > 0
    rng = np.random.default_rng(9)
    for prec, dt in ((ed.PREC_FP64, np.complex128), (ed.PREC_FP32, np.complex64)):
        for cat, b in cases:
            x = np.ascontiguousarray((rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)).astype(dt))
            ref = run(hip, lib.sed_matvec_device, b, x, prec)      # chooses the launch shape
            y1 = run(hip, lib.sed_matvec_device_noalloc, b, x, prec)
            y2 = run(hip, lib.sed_matvec_device_noalloc, b, x, prec)
            print(f"[noalloc] prec={prec} dim={b.dim} max|diff|={np.abs(y1 - ref).max():.1e}", flush=True)
            check(y1, ref, prec, b)
            assert np.array_equal(y1, y2)
    # the calls in a row: large sector, then the smaller one, without matvec_device between
    xs = [np.ascontiguousarray(rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)) for _, b in cases]
    refs = [run(hip, lib.sed_matvec_device, b, x, ed.PREC_FP64) for (_, b), x in zip(cases, xs)]
    outs = [run(hip, lib.sed_matvec_device_noalloc, b, x, ed.PREC_FP64) for (_, b), x in zip(cases, xs)]
    for a, r, (_, b) in zip(outs, refs, cases):
        check(a, r, ed.PREC_FP64, b)
    for cat, b in cases:
        b.close()
        cat.close()
    ctx.close()


def test_noalloc_matvec_refuses_oversize_sector(ed, hip):
    """Checkerboard 4x6 (dim 15,724,214): 252 MB of FP64 scratch, over the 64 MiB array."""
    ctx = ed.Context(0)
    m = L.workload("checker4x6")
    G = L.model_group(m)
    cat = ed.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                          particle_statistics=m.statistics, ctx=ctx)
    b = ed.build_symmetry_sector_basis(cat, L.translation_irreps(G, m.lattice)[0])
    b.set_terms(m.bilinear_terms, m.density_terms)
    one = C.c_void_p(1 << 20)   # never dereferenced: the size check comes first
    rc = _lib.lib().sed_matvec_device_noalloc(b._h, one, one, ed.PREC_FP64)
    assert rc == _lib.SED_EINVAL and b"scratch" in _lib.lib().sed_last_error()
    b.close()
    cat.close()
    ctx.close()
