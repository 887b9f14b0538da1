#!/usr/bin/env python3
"""Sector H|psi> throughput on one MI355X.

Workload (default): spinless fermions on the checkerboard lattice 4x6 (48 sites),
8 fermions, t=1 with phase pi/4, t'=+-0.3, t''=-0.2, V1=1, Z4xZ6 translations,
momentum sector k=(0,0) (dim 15,724,214).  One step = one sector matvec
y = H x with complex-FP64 vectors over resident device buffers: the reference's
arithmetic (ComplexF64) and the precision of the library's default solve.  The
complex-FP32 matvec (FP32 accumulation, the first phase of the FP32+refine solve) is
timed the same way and reported beside it (`fp32`).

Timed region: the H|psi> kernel alone, launched back to back on the library stream
over resident buffers, bracketed by HIP events on that stream.  Its input is the
prescaled vector xhat = sqrt(stab) * x; the prescale pass (D * 2s bytes, 1.3 % of a
matvec's bytes here) is not timed: inside a Lanczos step it is fused into the kernel
that normalises the new basis vector.

value = roofline.achieved = algorithmic bytes of the kernel in this library's data
layout / its mean duration.  Algorithmic bytes per matvec (s = 8 for c32, 16 for c64;
D, M_g, M counted exactly on the device):
  D * (8 rep mask + 1 stabilizer order + s x[row] + s y[row])
  + M_g * 4   (one 4-byte sector-table entry per gated hop; 8 bytes on sectors whose
               rows do not fit the 4-byte entry, e.g. --workload xxz34)
  + M   * s   (one gather of xhat per in-sector hop)
The workload's own per-unit figures (an 8-byte table entry and a 4-byte norm per row)
give workload_bytes_per_launch, reported beside it; they count bytes this layout does not
move, so their rate is not an HBM fraction.  roofline.traffic is the HBM traffic per
launch measured by rocprofv3 counters (FETCH_SIZE and WRITE_SIZE, copy-calibrated) in
the committed profile named in roofline.traffic_source (null for other workloads).
hbm_streaming gives the rates of two plain streaming kernels over 2 GiB measured in the
same run (torch's device copy, read + write, and sum, read only): what library streaming
kernels reach on this GPU, beside the published peak (null under rocprofv3, where this
process starts no child process).  matvecs_per_s = 1000 / ms_per_step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg, **_):
    """Progress lines go to stderr; stdout carries only the result JSON line."""
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(dim, gated, ins, s):
    """Algorithmic bytes of one sector matvec with the per-unit figures of the workload
    definition: per row the representative (8 B), its norm (4 B), x[row] and the y[row]
    write (2s); per gated hop one rep-table entry of 8 B; per in-sector hop one gather of
    x (s).  s = bytes per complex element."""
    return dim * (8 + 4 + 2 * s) + gated * 8 + ins * s


def table_entry_bytes(n_site, group_order, dim):
    """Bytes of one sector-table entry: row << shift | g << 1 | sign in 4 bytes, or 8 when
    the row does not fit beside shift = ceil(log2 |G|) + 1 bits (sectors above 32 sites)."""
    shift = max(0, (group_order - 1).bit_length()) + 1
    return 8 if n_site > 32 and dim > (1 << (32 - shift)) - 1 else 4


def layout_bytes(dim, gated, ins, s, entry=4):
    """The same count for this library's compressed layout: 1-byte stabilizer order and
    a 4-byte (8-byte on the largest sectors) table entry (row, group element and sign packed)."""
    return dim * (8 + 1 + 2 * s) + gated * entry + ins * s


COUNTER_PROFILE = os.path.join("profiles", "matvec_checker4x6_counters.md")
COUNTER_WORKLOAD = "checker4x6"


def measured_traffic(precision, workload):
    """(HBM bytes per launch of the sector matvec on this workload, source) from the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, scaled by the
    plain-copy calibration); (None, None) when the summary has no line for it or covers
    another workload."""
    path = os.path.join(ROOT, COUNTER_PROFILE)
    if workload != COUNTER_WORKLOAD or not os.path.exists(path):
        return None, None
    key = f"traffic_bytes_{precision}:"
    src = None
    val = None
    with open(path) as f:
        for line in f:
            if line.strip().startswith("build:"):
                src = line.split(":", 1)[1].strip()
            if line.strip().startswith(key):
                val = int(line.split(":", 1)[1].strip())
    if val is None:
        return None, None
    return val, f"{COUNTER_PROFILE} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated; {src or 'build unstated'})"


_COPY_PROBE = r"""
import json, sys, torch
gib, reps = int(sys.argv[1]), int(sys.argv[2])
n = (gib << 30) // 4
a = torch.ones(n, dtype=torch.float32, device="cuda:0")
b = torch.empty_like(a)
for _ in range(3):
    b.copy_(a)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range(reps):
    b.copy_(a)
e1.record()
e1.synchronize()
ms = e0.elapsed_time(e1) / reps
for _ in range(3):
    s = a.sum()
e0.record()
for _ in range(reps):
    s = a.sum()
e1.record()
e1.synchronize()
ms_sum = e0.elapsed_time(e1) / reps
print(json.dumps({"copy_GB/s": round(2 * 4 * n / (ms * 1e-3) / 1e9, 1), "copy_bytes": 2 * 4 * n,
                  "copy_ms": round(ms, 4), "sum_GB/s": round(4 * n / (ms_sum * 1e-3) / 1e9, 1),
                  "sum_bytes": 4 * n, "sum_ms": round(ms_sum, 4)}))
"""


def copy_bandwidth(gib=2, reps=10):
    """Rates (GB/s) of two plain streaming kernels over `gib` GiB of float32, timed with
    HIP events on torch's stream: a device-to-device copy (`copy_`, read + write) and a
    sum (read only).  Reported beside the published 8 TB/s peak as what library streaming
    kernels reach on this GPU.  Measured in a child process (torch brings its own HIP
    runtime, which cannot share this process with the library's)."""
    import subprocess
    if any(k.startswith("ROCPROF") for k in os.environ):
        # under rocprofv3 this process is set up for the GPU before it starts; it then
        # starts no other program
        return None
    r = subprocess.run([sys.executable, "-c", _COPY_PROBE, str(gib), str(reps)], capture_output=True, text=True,
                       timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"copy probe failed:\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def host_info():
    """CPU model, logical CPUs, CPUs this process may run on, RAM (GiB)."""
    model, ram = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    ram = round(int(line.split()[1]) / 2 ** 20, 1)
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity, "ram_gib": ram,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_solves(threads):
    """Per-solve CPU times of the oracle's restatement of the reference's matrix-free
    sector solve (ed_scan_at_irrep_matrixfree!, src/symmetry_resolved_ed.jl:1092-1128):
    sector basis, serial projection-table build (:680-728), Lanczos eigensolve (ARPACK
    through scipy over _apply_table!, :731-765, with `threads` per-thread buffers; the
    reference uses KrylovKit) at tol 1e-10; the catalog is built before the timed region,
    as the reference's benchmark does (benchmark/benchmark.jl:226-241).  The XXZ N=16 sector
    (k=0, dim 810, nev 4) and the flux scan's Haldane 3x4 boson sector at zero twist
    (k=(0,0), dim 11,240), the latter at nev 1 beside the reference's published 0.872 s
    (benchmark/benchmark_data/benchmark_raw.csv:20-21, hardware unrecorded)."""
    import numpy as np
    from scipy.sparse.linalg import LinearOperator, eigsh
    from oracle import oracle as O
    from realspace_exactdiagonalization_amd import lattice as L
    out = {}
    for name, wl, nev, published in (("xxz16_k0", "xxz16", 4, None), ("haldane3x4_k00", "haldane3x4", 1, 0.872),
                                     ("haldane3x4_k00_nev3", "haldane3x4", 3, None)):
        m = L.workload(wl)
        G = L.model_group(m)
        ir = L.translation_irreps(G, m.lattice)[0]
        cat = O.Catalog(m.n_site, m.n_filled, G, m.statistics == L.FERMIONIC)
        t0 = time.time()
        sec = O.Sector(cat, ir.values)
        sec.set_terms(*L.term_arrays(m))
        n = sec.dim
        op = LinearOperator((n, n), matvec=lambda x: sec.apply_table(x, threads), dtype=np.complex128)
        v0 = np.random.default_rng(1234).standard_normal(n) + 0j
        vals = np.sort(eigsh(op, k=nev, which="SA", tol=1e-10, v0=v0, return_eigenvectors=False).real)
        dt = time.time() - t0
        out[name] = {"dim": n, "nev": nev, "s_per_solve": round(dt, 4), "evals": vals.tolist()}
        if published is not None:
            out[name]["reference_published_s"] = published
        log(f"[cpu] solve {name} dim {n} nev {nev}: {dt:.3f} s", flush=True)
    return out


def cpu_baseline(model, G, ir, cat, basis, seconds, threads, bytes_c64, gated, ins):
    """The oracle's restatement of the reference's matrix-free matvec, timed on this host
    on bounded samples of the sector, in two forms:

    - the path the reference's solver runs (`value`): build_matrixfree_projection_table
      (src/symmetry_resolved_ed.jl:680-728, serial) over runs of consecutive columns spread
      over the sector, then
      _apply_table! (:731-765: per-thread dense buffers zeroed per call, static split of
      the entries over the threads, serial reduction) over that partial table, timed from
      _apply_table!'s own timing of its parts: buffer zeroing, diagonal and reduction run
      over the whole sector as they are, and only the entry loop is scaled by M / nnz_sample
      (M = the sector's in-sector hops = its table entries); the serial table build is
      extrapolated to all columns as well (the reference builds it once per sector);
    - the on-the-fly apply_hamiltonian! (:813-878, threaded over columns), extrapolated
      the same way over columns (`on_the_fly`).

    Its lookup tables are the device catalog dumped in the reference's layout (the test
    suite checks them bit-exact against the oracle's own serial build, which would take
    minutes at C(48,8) = 377M states)."""
    import numpy as np
    from oracle import oracle as O
    from realspace_exactdiagonalization_amd import lattice as L
    t0 = time.time()
    rr, rs, ra = cat.tables()
    oc = O.WrappedCatalog(model.n_site, model.n_filled, G, model.statistics == L.FERMIONIC,
                          cat.representatives(), rr, rs, ra)
    osec = O.Sector(oc, ir.values, basis=(basis.representative_mask_list, basis.stabilizer_order_list))
    osec.set_terms(*L.term_arrays(model), build_table=False)
    prep = time.time() - t0
    D = osec.dim
    x = np.ones(D, dtype=np.complex128) * (1 + 1j)

    # _apply_table! over a serially built partial table: runs of RUN consecutive columns
    # spread evenly over the sector, so that its entries keep the full table's
    # column-to-column locality and scatter over the whole row range as the full table's do
    RUN = 2048
    c_cal = min(D, 4 * RUN)
    t0 = time.time()
    osec.build_table_cols(0, D, max(RUN, D // 4), min(RUN, D))
    per_col_build = max(time.time() - t0, 1e-9) / c_cal
    want = int(min(D, max(c_cal, 0.8 * seconds / per_col_build)))
    period = max(RUN, (D // max(want, 1)) * RUN) if want < D else 1
    block = min(RUN, D) if want < D else 1
    nnz = 0
    t0 = time.time()
    nnz = osec.build_table_cols(0, D, period, block)
    t_build = time.time() - t0
    cols_t = sum(min(block, D - c) for c in range(0, D, period)) if period > 1 else D

    osec.apply_table(x, threads)                      # allocates the per-thread buffers
    parts = [osec.apply_table_timed(x, threads)[1] for _ in range(3)]
    med = {k: float(np.median([p[k] for p in parts])) for k in parts[0]}
    osec.build_table_cols(0, 0)
    t_fixed_table = med["zero"] + med["diagonal"] + med["reduction"]
    t_table = t_fixed_table + med["entries"] * ins / max(nnz, 1)
    t_build_full = t_build * D / cols_t
    gbs = bytes_c64 / t_table / 1e9
    log(f"[cpu] prep {prep:.1f}s table of {cols_t} cols (runs of {block} every {period}, {nnz} entries) "
        f"built in {t_build:.2f}s serially; "
        f"_apply_table! zero {med['zero']:.3f}s diagonal {med['diagonal']:.3f}s entries {med['entries']:.3f}s "
        f"reduction {med['reduction']:.3f}s -> {t_table:.2f} s/matvec {gbs:.2f} GB/s "
        f"(whole table build {t_build_full:.0f}s serial)", flush=True)

    # on-the-fly apply_hamiltonian! over columns
    t0 = time.time()
    osec.apply_hamiltonian(x, threads, 0, 0)
    t_fixed = time.time() - t0
    cols = min(D, 20000)
    t0 = time.time()
    osec.apply_hamiltonian(x, threads, 0, cols)
    t_cal = time.time() - t0
    per_col = max(t_cal - t_fixed, 1e-9) / cols
    cols = int(min(D, max(cols, (0.3 * seconds - t_fixed) / per_col)))
    t0 = time.time()
    osec.apply_hamiltonian(x, threads, 0, cols)
    t_sample = time.time() - t0
    t_full = t_fixed + max(t_sample - t_fixed, 0.0) * D / cols
    gbs_otf = bytes_c64 / t_full / 1e9
    log(f"[cpu] on-the-fly fixed {t_fixed:.2f}s sample {cols} cols {t_sample:.2f}s -> "
        f"{t_full:.2f} s/matvec {gbs_otf:.2f} GB/s", flush=True)
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"_apply_table! (FP64) over the projection table of {cols_t} of {D} columns (runs of "
                      f"{block} consecutive columns every {period}) "
                      f"({nnz} of {ins} entries, {100.0 * nnz / max(ins, 1):.2f}%, built serially in "
                      f"{t_build:.1f}s); medians of 3 calls: buffer zeroing {med['zero']:.3f}s, diagonal "
                      f"{med['diagonal']:.3f}s, reduction {med['reduction']:.3f}s (all for the whole "
                      f"sector), entries {med['entries']:.4f}s scaled by {ins}/{nnz}; full matvec "
                      f"{t_table:.2f}s; rate on the same layout bytes as the GPU's c64 figure",
            "phases_s": {k: round(v, 4) for k, v in med.items()},
            "s_per_matvec": round(t_table, 3),
            "table_build_s_extrapolated": round(t_build_full, 1),
            "on_the_fly": {"value": round(gbs_otf, 3), "unit": "GB/s", "s_per_matvec": round(t_full, 3),
                           "sample": f"apply_hamiltonian! over columns [0,{cols}) of {D} "
                                     f"({100.0 * cols / D:.2f}%), {t_sample:.1f}s, of it {t_fixed:.2f}s "
                                     f"per-thread buffer zero + reduction"},
            "host": host_info()}


def flux_scan(E, L, ctx, n_twists=16, nev=3, lanes=8):
    """Flux scan: Haldane-boson 3x4 (24 sites, 6 bosons), all 12 momentum
    sectors x n_twists twists theta_x in [0, 1), lowest nev eigenvalues each.

    Three ways: the plain FP64 path one solve after another (sed_solve), and the
    continuous-batching server with FP64 and with FP32+FP64 refinement.  Returns
    solves/s, p50/p99 latency and the largest eigenvalue difference between the modes."""
    import numpy as np
    base = L.workload("haldane3x4")
    keep, sectors = [], []
    for i in range(n_twists):
        th = (i / n_twists, 0.0)
        m = L.twisted(base, th)
        G = L.model_group(m, th)
        cat = E.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                             particle_statistics=m.statistics, ctx=ctx)
        keep.append(cat)
        for ir in L.translation_irreps(G, m.lattice):
            b = E.build_symmetry_sector_basis(cat, ir)
            b.set_terms(m.bilinear_terms, m.density_terms)
            sectors.append(b)
    out = {"solves": len(sectors), "dims": [min(b.dim for b in sectors), max(b.dim for b in sectors)], "nev": nev}
    # plain FP64, sequential
    ev_plain, lat = [], []
    t0 = time.time()
    for b in sectors:
        H = E.MatrixFreeHamiltonian.__new__(E.MatrixFreeHamiltonian)
        H.basis, H.n, H.precision = b, b.dim, E.PREC_FP64
        t1 = time.time()
        ev_plain.append(E.diagonalize_block_matrixfree(H, b.dim, nev=nev, return_vectors=False)[0])
        lat.append(time.time() - t1)
    wall = time.time() - t0
    out["plain_fp64"] = {"solves_per_s": round(len(sectors) / wall, 2), "wall_s": round(wall, 3),
                         "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
                         "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2)}
    for name, prec in (("batched_fp64", E.PREC_FP64), ("batched_fp32_refine", E.PREC_FP32_REFINE)):
        srv = E.SectorServer(ctx, max_slots=lanes)
        t0 = time.time()
        ids = [srv.submit(b, nev, precision=prec) for b in sectors]
        srv.run(0)
        wall = time.time() - t0
        lat, err, iters = [], 0.0, []
        for rid, ref in zip(ids, ev_plain):
            st, ev, stats, l = srv.result(rid)
            if st != 0:
                raise RuntimeError(f"request {rid} failed with status {st}")
            lat.append(l)
            err = max(err, float(np.max(np.abs(ev - ref) / np.maximum(np.abs(ref), 1e-300))))
            iters.append(stats["iters_fp32"] + stats["iters_fp64"])
        srv.close()
        out[name] = {"solves_per_s": round(len(sectors) / wall, 2), "wall_s": round(wall, 3),
                     "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
                     "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
                     "max_rel_diff_vs_plain_fp64": err, "mean_lanczos_steps": round(float(np.mean(iters)), 1),
                     "lanes": lanes}
    out["speedup_batched_fp32_refine_vs_plain_fp64"] = round(
        out["batched_fp32_refine"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"], 3)
    for b in sectors:
        b.close()
    return out


def poisson_arrivals(n, rate, seed):
    """Arrival times (s) of a Poisson stream of n requests at `rate` per second, first at 0."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return t - t[0]


def serving_stream(E, L, ctx, n_req=32, seed=2024, nev=3, lanes=8, load=1.2):
    """Serving stream of mixed sector solves.  The sectors of all requests are built first (tables
    resident in HBM) and every model's matvec launch shape is tuned in both precisions (a
    one-time cost per model, reported apart, so that no timed path below pays it); then
    (1) the plain FP64 path solves them one after another, (2) the continuous-batching
    server (FP32 + FP64 refinement) takes all of them at once, which measures its
    saturated throughput, (2b) the same with FP64 vectors, and (3) the FP32 + FP64 server
    is fed a seeded Poisson stream at `load` times its saturated throughput.  The plain
    path's latencies for the same arrivals follow from its solve times (one FIFO server).
    Reports solves/s and p50/p99 latency."""
    import threading
    import numpy as np
    reqs = L.stream_requests(n_req, seed)
    t0 = time.time()
    keep, sectors, info = [], [], []
    for r in reqs:
        m, G, ir = L.stream_sector(r)
        cat = E.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                             particle_statistics=m.statistics, ctx=ctx)
        b = E.build_symmetry_sector_basis(cat, ir)
        b.set_terms(m.bilinear_terms, m.density_terms)
        keep.append(cat)
        if b.dim == 0:
            b.close()
            continue
        sectors.append(b)
        info.append((r["family"], r["size"], b.dim))
    t_build = time.time() - t0
    t0 = time.time()
    for b in sectors:
        H = E.MatrixFreeHamiltonian.__new__(E.MatrixFreeHamiltonian)
        H.basis, H.n, H.precision = b, b.dim, E.PREC_FP64
        for prec in (E.PREC_FP64, E.PREC_FP32):
            H.benchmark(prec, 1)
    t_tune = time.time() - t0
    n = len(sectors)
    dims = np.array([d for _, _, d in info])
    # (1) plain FP64, one solve after another
    ev_plain, t_solve = [], []
    for b in sectors:
        H = E.MatrixFreeHamiltonian.__new__(E.MatrixFreeHamiltonian)
        H.basis, H.n, H.precision = b, b.dim, E.PREC_FP64
        t1 = time.time()
        ev_plain.append(E.diagonalize_block_matrixfree(H, b.dim, nev=min(nev, b.dim), return_vectors=False)[0])
        t_solve.append(time.time() - t1)
    t_solve = np.array(t_solve)
    plain_rate = n / float(t_solve.sum())

    def collect(srv, ids):
        lat, err = [], 0.0
        for rid, ref in zip(ids, ev_plain):
            st, ev, stats, l = srv.result(rid)
            if st != 0:
                raise RuntimeError(f"stream request {rid} failed with status {st}")
            lat.append(l)
            ev = ev[:len(ref)]
            err = max(err, float(np.max(np.abs(ev - ref) / np.maximum(np.abs(ref), 1e-300))))
        return np.array(lat), err

    # (2) saturated throughput of the server
    srv = E.SectorServer(ctx, max_slots=lanes)
    t1 = time.time()
    ids = [srv.submit(b, min(nev, b.dim), precision=E.PREC_FP32_REFINE) for b in sectors]
    srv.run(0)
    sat_wall = time.time() - t1
    _, err_sat = collect(srv, ids)
    srv.close()
    sat_rate = n / sat_wall
    # (2b) the same with FP64 vectors
    srv = E.SectorServer(ctx, max_slots=lanes)
    t1 = time.time()
    ids = [srv.submit(b, min(nev, b.dim), precision=E.PREC_FP64) for b in sectors]
    srv.run(0)
    sat64_wall = time.time() - t1
    _, err_sat64 = collect(srv, ids)
    srv.close()
    sat64_rate = n / sat64_wall
    # (3) Poisson stream at load x the saturated rate
    arr = poisson_arrivals(n, load * sat_rate, seed + 1)
    srv = E.SectorServer(ctx, max_slots=lanes)
    ids = [None] * n

    def feed():
        t_ref = time.time()
        for i, b in enumerate(sectors):
            dt = t_ref + arr[i] - time.time()
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit(b, min(nev, b.dim), precision=E.PREC_FP32_REFINE)

    t1 = time.time()
    th = threading.Thread(target=feed)
    th.start()
    th.join()
    srv.run(0)
    wall = time.time() - t1
    lat, err = collect(srv, ids)
    srv.close()
    # plain FP64 on the same arrivals: single FIFO server
    done = np.zeros(n)
    tcur = 0.0
    for i in range(n):
        tcur = max(tcur, arr[i]) + t_solve[i]
        done[i] = tcur
    plain_lat = done - arr
    for b in sectors:
        b.close()
    q = lambda a, p: round(1e3 * float(np.percentile(a, p)), 2)
    return {
        "requests": n, "families": {f: int(sum(1 for x in info if x[0] == f)) for f in L.STREAM_SIZES},
        "dims": [int(dims.min()), int(np.median(dims)), int(dims.max())], "nev": nev, "seed": seed,
        "build_s": round(t_build, 2), "tune_s": round(t_tune, 2),
        "plain_fp64": {"solves_per_s": round(plain_rate, 3), "solve_s_total": round(float(t_solve.sum()), 3),
                       "stream_solves_per_s": round(n / float(done[-1]), 3),
                       "p50_ms": q(plain_lat, 50), "p99_ms": q(plain_lat, 99)},
        "server_fp32_refine": {"saturated_solves_per_s": round(sat_rate, 3),
                               "arrival_rate_per_s": round(load * sat_rate, 3),
                               "stream_solves_per_s": round(n / wall, 3),
                               "p50_ms": q(lat, 50), "p99_ms": q(lat, 99),
                               "max_rel_diff_vs_plain_fp64": max(err, err_sat), "lanes": lanes},
        "server_fp64": {"saturated_solves_per_s": round(sat64_rate, 3),
                        "max_rel_diff_vs_plain_fp64": err_sat64, "lanes": lanes},
        "speedup_saturated_vs_plain_fp64": round(sat_rate / plain_rate, 3),
        "speedup_saturated_fp64_vs_plain_fp64": round(sat64_rate / plain_rate, 3),
        "per_request_plain_s": [[f, str(sz), int(d), round(float(t), 3)] for (f, sz, d), t in zip(info, t_solve)],
    }


def solve_model(L, name):
    """A benchmark workload by name (see lattice.workload), or a serving-stream family:
    xxzN (XXZ chain, N sites, half filling), ladderL (2 x L Hubbard ladder, quarter filling)."""
    if name.startswith("xxz") and name[3:].isdigit() and name != "xxz16":
        return L.heisenberg_chain(int(name[3:]))
    if name.startswith("ladder") and name[6:].isdigit():
        return L.hubbard_square(int(name[6:]), 2, U=4.0, t=1.0, n_filled=int(name[6:]))
    return L.workload(name)


def reference_cases(args):
    """The reference's own single-sector benchmark on its own models
    (benchmark/benchmark.jl:226-320): per model, three repetitions of [build the orbit
    catalog (untimed), then time ed_scan_at_irrep_matrixfree on the first momentum sector
    with nev = 1], mean of repetitions 2-3 (the first warms caches; here it also autotunes).
    One JSON line per model, with the reference's stored matrix-free and matrix-mode times
    (benchmark/benchmark_data/benchmark_raw.csv, kept in tests/golden: CPU hardware and
    thread count unrecorded) and the energy against its stored value."""
    from realspace_exactdiagonalization_amd import ed as E
    from realspace_exactdiagonalization_amd import lattice as L
    with open(os.path.join(ROOT, "tests", "golden", "reference_values.json")) as f:
        rows = json.load(f)["benchmark_raw"]
    cases = []
    for r in rows:
        if (r["model"], r["label"]) not in cases:
            cases.append((r["model"], r["label"]))
    ctx = E.Context(0)
    for model, label in cases:
        ref = {r["mode"]: r for r in rows if r["model"] == model and r["label"] == label}
        m = L.reference_benchmark_model(model, label)
        times, e0, dim = [], None, 0
        for rep in range(3):
            data = E.build_ed_data(m, ctx=ctx)
            t0 = time.time()
            vals, _ = E.ed_scan_at_irrep_matrixfree(data.irrep_list[0].label, data, nev=1, return_vectors=False)
            times.append(time.time() - t0)
            e0, dim = float(vals[0]), data.sector_dims[0]
            data.orbit_catalog.close()
        t = sum(times[1:]) / 2
        mf, mx = ref["matrixfree"]["elapsed_s"], ref["matrix"]["elapsed_s"]
        print(json.dumps({"model": model, "label": label, "sector_dim": dim, "gpu_s": round(t, 5),
                          "reference_matrixfree_s": mf, "reference_matrix_s": mx,
                          "speedup_vs_reference_matrixfree": round(mf / t, 1),
                          "energy": e0, "rel_diff_vs_reference_energy":
                              abs(e0 - ref["matrixfree"]["energy"]) / abs(ref["matrixfree"]["energy"])}),
              flush=True)


def single_solves(args):
    """One sector solve per precision (plain FP64, then FP32 + FP64 refinement), each
    repeated; prints one JSON line per solve with its Lanczos step counts and time."""
    from realspace_exactdiagonalization_amd import ed as E
    from realspace_exactdiagonalization_amd import lattice as L
    m = solve_model(L, args.solve)
    G = L.model_group(m)
    ctx = E.Context(0)
    cat = E.build_symmetry_orbit_catalog(n_site=m.n_site, n_filled=m.n_filled, symmetry_group=G,
                                         particle_statistics=m.statistics, ctx=ctx)
    ir = L.translation_irreps(G, m.lattice)[args.irrep]
    b = E.build_symmetry_sector_basis(cat, ir)
    H, n = E.hamiltonian_linear_operator(b, m.bilinear_terms, m.density_terms)
    for prec, label in ((E.PREC_FP64, "fp64"), (E.PREC_FP32_REFINE, "fp32_refine")):
        for rep in range(args.solve_repeats):
            st = {}
            t0 = time.time()
            vals, _ = E.diagonalize_block_matrixfree(H, n, nev=args.nev, precision=prec, return_vectors=False,
                                                     stats=st)
            print(json.dumps({"solve": args.solve, "irrep": list(ir.label), "dim": n, "precision": label,
                              "repeat": rep, "wall_s": round(time.time() - t0, 4), "stats": st,
                              "evals": vals.tolist()}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="checker4x6")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flux", action="store_true", help="skip the batched flux-scan solves")
    ap.add_argument("--lanes", type=int, default=8, help="server lanes for the flux scan and the stream")
    ap.add_argument("--no-stream", action="store_true", help="skip the serving-stream benchmark")
    ap.add_argument("--stream-requests", type=int, default=32)
    ap.add_argument("--solve", default=None,
                    help="only time sector solves of this model (a workload name, xxzN or ladderL)")
    ap.add_argument("--irrep", type=int, default=0, help="momentum sector index for --solve")
    ap.add_argument("--nev", type=int, default=3, help="eigenpairs for --solve")
    ap.add_argument("--solve-repeats", type=int, default=2)
    ap.add_argument("--reference-cases", action="store_true",
                    help="only run the reference's single-sector benchmark models and protocol")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    if args.solve:
        single_solves(args)
        return
    if args.reference_cases:
        reference_cases(args)
        return

    if args.copy_calibration:
        import torch
        n = 1 << 28                      # 2^28 floats = 1 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                          "dispatches": 3}))
        return

    # before this process touches the GPU: the probe is a child process
    copy = copy_bandwidth()
    if copy:
        log(f"[bench] streaming kernels: copy {copy['copy_GB/s']:.0f} GB/s (read + write), "
            f"sum {copy['sum_GB/s']:.0f} GB/s (read)", flush=True)

    from realspace_exactdiagonalization_amd import ed as E
    from realspace_exactdiagonalization_amd import lattice as L

    t_setup = time.time()
    model = solve_model(L, args.workload)
    G = L.model_group(model)
    ir = L.translation_irreps(G, model.lattice)[0]
    ctx = E.Context(0)
    t0 = time.time()
    cat = E.build_symmetry_orbit_catalog(n_site=model.n_site, n_filled=model.n_filled, symmetry_group=G,
                                         particle_statistics=model.statistics, ctx=ctx)
    t_cat = time.time() - t0
    t0 = time.time()
    basis = E.build_symmetry_sector_basis(cat, ir)
    t_sec = time.time() - t0
    H = E.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_terms)
    gated, ins = basis.hop_counts()
    dim = basis.dim
    log(f"[bench] {args.workload}: C(N,k)={cat.n_total} orbits={cat.n_orbits} dim={dim} "
          f"gated_hops={gated} in_sector_hops={ins} catalog={t_cat:.2f}s sector={t_sec:.2f}s", flush=True)

    res = {}
    for prec, s, name in [(E.PREC_FP32, 8, "c32"), (E.PREC_FP64, 16, "c64")]:
        if args.warmup > 0:
            H.benchmark(prec, args.warmup)
        w0 = time.time()
        ms = H.benchmark(prec, args.steps)
        wall = time.time() - w0
        B = layout_bytes(dim, gated, ins, s, table_entry_bytes(model.n_site, G.order, dim))
        Bw = algorithmic_bytes(dim, gated, ins, s)
        traffic, src = measured_traffic(name, args.workload)
        res[name] = {"ms": ms, "bytes": B, "gbs": B / (ms * 1e-3) / 1e9, "wall_s": wall,
                     "workload_bytes": Bw, "workload_gbs": Bw / (ms * 1e-3) / 1e9, "shape": basis.shape(prec),
                     "traffic": traffic, "traffic_source": src}
        log(f"[bench] {name}: {ms:.3f} ms/matvec  {B/1e9:.2f} GB algorithmic  "
            f"{res[name]['gbs']:.0f} GB/s ({100*res[name]['gbs']/PEAK_HBM_GBS:.1f}% of 8 TB/s)  "
            f"shape {res[name]['shape']}", flush=True)

    flux = None
    if not args.no_flux:
        flux = flux_scan(E, L, ctx, lanes=args.lanes)
        log(f"[bench] flux scan {flux['solves']} solves: plain fp64 {flux['plain_fp64']['solves_per_s']}/s, "
              f"batched fp64 {flux['batched_fp64']['solves_per_s']}/s, batched fp32+refine "
              f"{flux['batched_fp32_refine']['solves_per_s']}/s (p99 {flux['batched_fp32_refine']['p99_ms']} ms)",
              flush=True)
    stream = None
    if not args.no_stream:
        stream = serving_stream(E, L, ctx, n_req=args.stream_requests, lanes=args.lanes)
        log(f"[bench] serving stream {stream['requests']} requests dims {stream['dims']}: plain fp64 "
              f"{stream['plain_fp64']['solves_per_s']}/s, server fp32+refine saturated "
              f"{stream['server_fp32_refine']['saturated_solves_per_s']}/s (fp64 "
              f"{stream['server_fp64']['saturated_solves_per_s']}/s), Poisson stream p50/p99 "
              f"{stream['server_fp32_refine']['p50_ms']}/{stream['server_fp32_refine']['p99_ms']} ms "
              f"(plain {stream['plain_fp64']['p50_ms']}/{stream['plain_fp64']['p99_ms']} ms)", flush=True)
    main_r = res["c64"]
    # the host cores this process may use (the harness's share on a shared box: OMP_NUM_THREADS)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(model, G, ir, cat, basis, args.cpu_seconds, threads, res["c64"]["bytes"], gated, ins)
        cpu["per_solve"] = cpu_solves(threads)

    def roof(r):
        return {"bound": "hbm", "achieved": round(r["gbs"], 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(r["gbs"] / PEAK_HBM_GBS, 4), "traffic": r["traffic"],
                "traffic_source": r["traffic_source"],
                "traffic_frac": None if r["traffic"] is None else round(r["traffic"] / (r["ms"] * 1e-3) / 1e9 /
                                                                      PEAK_HBM_GBS, 4)}
    out = {
        "metric": "sector matvec GB/s (% HBM peak)",
        "value": round(main_r["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms"], 4),
        "matvecs_per_s": round(1e3 / main_r["ms"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} k=(0,0) sector matvec", "sector_dim": dim,
                   "n_site": model.n_site, "n_filled": model.n_filled, "group_order": G.order,
                   "gated_hops": gated, "in_sector_hops": ins},
        "roofline": roof(main_r),
        "algorithmic_bytes_per_launch": main_r["bytes"],
        "workload_bytes_per_launch": {"c32": res["c32"]["workload_bytes"], "c64": res["c64"]["workload_bytes"],
                                      "c32_GB/s": round(res["c32"]["workload_gbs"], 2),
                                      "c64_GB/s": round(res["c64"]["workload_gbs"], 2)},
        "cpu_baseline": cpu,
        "hbm_streaming": copy,
        "fp32": {"ms_per_matvec": round(res["c32"]["ms"], 4), "matvecs_per_s": round(1e3 / res["c32"]["ms"], 1),
                 "algorithmic_bytes_per_launch": res["c32"]["bytes"],
                 "roofline": roof(res["c32"]), "launch_shape": res["c32"]["shape"]},
        "launch_shape": res["c64"]["shape"],
        "flux_scan_haldane3x4": flux,
        "serving_stream": stream,
        "setup_s": {"catalog": round(t_cat, 3), "sector": round(t_sec, 3), "total": round(time.time() - t_setup, 3)},
    }
    basis.close()
    cat.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
