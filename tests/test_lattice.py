"""Host-side builders: term counts and group structure of the benchmark workloads."""
import numpy as np

from realspace_exactdiagonalization_amd import lattice as L


def test_workload_term_counts():
    assert len(L.workload("checker4x6").bilinear_terms) == 576
    assert len(L.workload("checker4x6").density_terms) == 96
    assert len(L.workload("haldane3x4").bilinear_terms) == 288
    assert len(L.workload("hubbard4x4").bilinear_terms) == 128
    assert len(L.workload("hubbard4x4").density_terms) == 16
    assert len(L.workload("xxz16").bilinear_terms) == 32
    assert len(L.workload("xxz16").density_terms) == 32


def test_translation_group_order_and_inverse():
    m = L.workload("checker4x6")
    for order in ("dy_fastest", "dx_fastest"):
        G = L.translation_group(m.lattice, (0.3, 0.1), order=order)
        assert G.order == 24
        assert np.array_equal(G.perm[0], np.arange(1, 49))
        for g in range(G.order):
            assert sorted(G.perm[g]) == list(range(1, 49))
    G = L.translation_group(m.lattice, order="dy_fastest")
    assert G.labels[1] == (0, 1)
    G = L.translation_group(m.lattice, order="dx_fastest")
    assert G.labels[1] == (1, 0)


def test_twisted_terms_are_hermitian():
    m = L.twisted(L.workload("haldane3x4"), (0.37, 0.0))
    d = {}
    for a, b, t in m.bilinear_terms:
        d[(a, b)] = d.get((a, b), 0) + t
    for (a, b), t in d.items():
        assert abs(d[(b, a)] - np.conj(t)) < 1e-14


def test_bench_table_entry_bytes_follow_the_library_rule():
    """bench.py counts 8-byte sector-table entries exactly where the library switches to
    them: more than 32 sites and a row index that does not fit 32 - (ceil(log2|G|) + 1)
    bits; the counter profile's traffic is only attached to its own workload."""
    import bench
    assert bench.table_entry_bytes(48, 24, 15724214) == 4          # checkerboard 4x6
    assert bench.table_entry_bytes(32, 32, 18784170) == 4          # XXZ N=32
    assert bench.table_entry_bytes(34, 34, 68635478) == 8          # XXZ N=34: rows >= 2^25
    assert bench.table_entry_bytes(34, 34, (1 << 25) - 1) == 4
    assert bench.table_entry_bytes(34, 34, 1 << 25) == 8
    assert bench.table_entry_bytes(34, 1, (1 << 31) - 1) == 4      # identity group: 31 row bits
    assert bench.layout_bytes(10, 7, 5, 8, 8) - bench.layout_bytes(10, 7, 5, 8) == 7 * 4
    assert bench.measured_traffic("c32", "xxz34") == (None, None)
    val, src = bench.measured_traffic("c32", "checker4x6")
    assert val > 0 and "matvec_checker4x6_counters.md" in src
