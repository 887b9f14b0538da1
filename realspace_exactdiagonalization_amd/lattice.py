"""Lattices, second-quantized term lists, translation groups and momentum irreps.

These are the caller-side inputs of the sector solver: the site numbering, the
directed hopping list ``(i_from, i_to, t)``, the density list ``(i, j, V)``, the
group's site permutations with their U(1) site phases, and the characters of
each 1D irrep.  They restate the conventions the reference's callers use:

* vertex index (1-based) = 1 + s + n_sub * (x + L1 * y): sublattice fastest,
  then x, then y (the TightBinding ``site_list`` order the reference's
  examples rely on; the reference's ``doc/design.ipynb`` cell 26 fixes it:
  Haldane 2x3 has 38 orbits and its 4th representative is mask 21).
* a bond added with ``add_hopping_term!((Ra, sa), (Rb, sb)) => t;
  is_hermitian=true`` is expanded over every unit cell R into the directed
  term (site(R+Ra, sa) -> site(R+Rb, sb), t) and its conjugate, duplicates
  kept (their contributions add).
* a term ``(i_from, i_to, t)`` moves one particle from ``i_from`` to ``i_to``
  (``src/symmetry_resolved_ed.jl:709-710``), i.e. it is t * a^dag_{i_to} a_{i_from}.
* twisted boundaries: a hop from cell c_from to c_to = (c_from + delta) mod L
  winds w = (c_from + delta - c_to) / L times and carries exp(+i 2 pi theta.w);
  this sign is the one for which H(theta) commutes with the flux-aware
  translations of ``build_translation_group`` (``src/symmetry_resolved_ed.jl:1374-1405``).
* translation group element order: ``build_translation_group`` loops
  ``for dx in 0:L1-1, dy in 0:L2-1`` (dy fastest, ``:1393``); the benchmark's
  Hubbard builder uses ``Iterators.product`` (dx fastest,
  ``benchmark/benchmark.jl:207``).  Both are offered.
* irreps: chi_k(g) = cis(2 pi (k1 dx / L1 + k2 dy / L2)), irreps ordered k1
  outer, k2 inner (``src/symmetry_resolved_ed.jl:1415-1430``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

BOSONIC = 0
FERMIONIC = 1


def _cis(x: float) -> complex:
    return complex(math.cos(x), math.sin(x))


class Lattice:
    """Periodic Bravais lattice with ``n_sub`` sublattices per cell."""

    def __init__(self, sample_size: Sequence[int], n_sub: int,
                 pbc: Sequence[bool] = (True, True), name: str = "lattice",
                 sub_crys: Optional[Sequence[Sequence[float]]] = None):
        self.sample_size = (int(sample_size[0]), int(sample_size[1]))
        self.n_sub = int(n_sub)
        self.pbc = (bool(pbc[0]), bool(pbc[1]))
        self.name = name
        self.dim = 2
        # crystal coordinates of each sublattice inside the cell (TightBinding's sub_crys_list)
        sc = [(0.0, 0.0)] * self.n_sub if sub_crys is None else [(float(a), float(b)) for a, b in sub_crys]
        if len(sc) != self.n_sub:
            raise ValueError("sub_crys needs one offset per sublattice")
        self.sub_crys: List[Tuple[float, float]] = sc
        L1, L2 = self.sample_size
        self.n_site = L1 * L2 * self.n_sub
        # site_list[i-1] = ((x, y), isub) with isub 1-based, like TightBinding
        self.site_list: List[Tuple[Tuple[int, int], int]] = [
            ((x, y), s) for y in range(L2) for x in range(L1) for s in range(1, self.n_sub + 1)
        ]

    def index(self, cell: Sequence[int], isub: int) -> int:
        """1-based vertex index of (cell mod L, isub)."""
        L1, L2 = self.sample_size
        x = int(cell[0]) % L1
        y = int(cell[1]) % L2
        return 1 + (isub - 1) + self.n_sub * (x + L1 * y)

    def cell_of(self, i: int) -> Tuple[int, int]:
        return self.site_list[i - 1][0]


@dataclass(frozen=True)
class Bond:
    """One ``add_hopping_term!((Ra, sa), (Rb, sb)) => t; is_hermitian=true``."""
    Ra: Tuple[int, int]
    sa: int
    Rb: Tuple[int, int]
    sb: int
    t: complex


@dataclass
class Model:
    """Second-quantized model: term lists plus the data needed to re-twist them."""
    lattice: Lattice
    statistics: int
    bonds: List[Bond]
    density_terms: List[Tuple[int, int, complex]]
    params: Dict[str, float] = field(default_factory=dict)
    n_filled: int = 0
    group_order: str = "dy_fastest"
    theta: Tuple[float, float] = (0.0, 0.0)
    extra_bilinear: Optional[List[Tuple[int, int, complex]]] = None
    bilinear_terms: List[Tuple[int, int, complex]] = field(default_factory=list)

    @property
    def n_site(self) -> int:
        return self.lattice.n_site


def expand_bonds(lattice: Lattice, bonds: Sequence[Bond],
                 theta: Sequence[float] = (0.0, 0.0)) -> List[Tuple[int, int, complex]]:
    """Directed term list of ``bonds`` over all cells with twisted boundaries.

    Each bond gives, per cell R, the forward term and its Hermitian conjugate.
    A bond that leaves the sample along a non-periodic direction is dropped.
    """
    L = lattice.sample_size
    out: List[Tuple[int, int, complex]] = []
    for b in bonds:
        delta = (b.Rb[0] - b.Ra[0], b.Rb[1] - b.Ra[1])
        for y in range(L[1]):
            for x in range(L[0]):
                ua = (x + b.Ra[0], y + b.Ra[1])
                ub = (x + b.Rb[0], y + b.Rb[1])
                ok = True
                for d in range(2):
                    if not lattice.pbc[d] and (ua[d] < 0 or ua[d] >= L[d] or ub[d] < 0 or ub[d] >= L[d]):
                        ok = False
                if not ok:
                    continue
                cf = (ua[0] % L[0], ua[1] % L[1])
                ct = (ub[0] % L[0], ub[1] % L[1])
                w = ((cf[0] + delta[0] - ct[0]) // L[0], (cf[1] + delta[1] - ct[1]) // L[1])
                phase = _cis(2.0 * math.pi * (theta[0] * w[0] + theta[1] * w[1]))
                tt = complex(b.t) * phase
                i_from = lattice.index(cf, b.sa)
                i_to = lattice.index(ct, b.sb)
                out.append((i_from, i_to, tt))
                out.append((i_to, i_from, tt.conjugate()))
    return out


def twisted(model: Model, theta: Sequence[float]) -> Model:
    """The same model with hopping terms regenerated at twist ``theta`` (units of 2 pi)."""
    th = (float(theta[0]), float(theta[1]))
    m = replace(model, theta=th)
    m.bilinear_terms = _bilinear_for(model.lattice, model.bonds, th, model.extra_bilinear)
    return m


def _bilinear_for(lattice, bonds, theta, extra):
    terms = expand_bonds(lattice, bonds, theta)
    if extra is not None:
        terms = list(extra)
    return terms


# --------------------------------------------------------------------------
# Symmetry group + irreps
# --------------------------------------------------------------------------
@dataclass
class SymmetryGroup:
    """Finite group of site permutations with U(1) site phases.

    ``perm[g, i-1]`` is the 1-based target vertex of vertex i under element g and
    ``phases[g, i-1]`` its phase eta_g(i) (``src/symmetry_resolved_ed.jl:62-104``).
    """
    name: str
    perm: np.ndarray            # (nG, n_site) int32, 1-based
    phases: np.ndarray          # (nG, n_site) complex128
    labels: List[Tuple[int, int]]
    identity_idx: int = 1

    @property
    def order(self) -> int:
        return int(self.perm.shape[0])

    @property
    def n_site(self) -> int:
        return int(self.perm.shape[1])


def identity_group(n_site: int) -> SymmetryGroup:
    return SymmetryGroup("identity", np.arange(1, n_site + 1, dtype=np.int32)[None, :],
                         np.ones((1, n_site), dtype=np.complex128), [(0, 0)])


def translation_group(lattice: Lattice, theta: Sequence[float] = (0.0, 0.0),
                      order: str = "dy_fastest") -> SymmetryGroup:
    """Translations Z_L1 x Z_L2 with flux-aware site phases.

    phases[i] = cis(2 pi theta.(cell_i - shifted_i)/L) as in
    ``src/symmetry_resolved_ed.jl:1393-1402``.
    """
    L1, L2 = lattice.sample_size
    if order == "dy_fastest":
        shifts = [(dx, dy) for dx in range(L1) for dy in range(L2)]
    elif order == "dx_fastest":
        shifts = [(dx, dy) for dy in range(L2) for dx in range(L1)]
    else:
        raise ValueError(f"unknown group order {order!r}")
    n = lattice.n_site
    perm = np.zeros((len(shifts), n), dtype=np.int32)
    phases = np.ones((len(shifts), n), dtype=np.complex128)
    th = (float(theta[0]), float(theta[1]))
    any_flux = th[0] != 0.0 or th[1] != 0.0
    for g, (dx, dy) in enumerate(shifts):
        for i, (cell, isub) in enumerate(lattice.site_list):
            sc = ((cell[0] + dx) % L1, (cell[1] + dy) % L2)
            perm[g, i] = lattice.index(sc, isub)
            if any_flux:
                phases[g, i] = _cis(2.0 * math.pi * (th[0] * (cell[0] - sc[0]) / L1
                                                     + th[1] * (cell[1] - sc[1]) / L2))
    return SymmetryGroup("translations", perm, phases, shifts, identity_idx=1)


@dataclass
class Irrep:
    label: Tuple[int, int]
    values: np.ndarray          # (nG,) complex128 characters chi(g)


def translation_irreps(group: SymmetryGroup, lattice: Lattice) -> List[Irrep]:
    """All momentum irreps, k1 outer and k2 inner (``src/symmetry_resolved_ed.jl:1415-1430``)."""
    if group.name == "identity":
        return [Irrep((0, 0), np.ones(1, dtype=np.complex128))]
    L1, L2 = lattice.sample_size
    out = []
    for k1 in range(L1):
        for k2 in range(L2):
            chi = np.array([_cis(2.0 * math.pi * (k1 * dx / L1 + k2 * dy / L2))
                            for (dx, dy) in group.labels], dtype=np.complex128)
            out.append(Irrep((k1, k2), chi))
    return out


def irrep_index(irreps: Sequence[Irrep], label) -> int:
    for i, ir in enumerate(irreps):
        if tuple(ir.label) == tuple(label):
            return i
    raise KeyError(label)


# --------------------------------------------------------------------------
# Model builders
# --------------------------------------------------------------------------
def _finish(model: Model) -> Model:
    model.bilinear_terms = _bilinear_for(model.lattice, model.bonds, model.theta, model.extra_bilinear)
    return model


def heisenberg_chain(N: int, J: float = 1.0, n_filled: Optional[int] = None) -> Model:
    """S=1/2 Heisenberg/XXZ chain as hard-core bosons (``benchmark/benchmark.jl:66-111``).

    Hop J/2 on each bond both ways, J n_i n_{i+1} per bond, and -J/2 n_i on site.
    """
    lat = Lattice((N, 1), 1, pbc=(True, False), name="Heisenberg_Chain")
    bonds = [Bond((0, 0), 1, (1, 0), 1, complex(J / 2))]
    dens = []
    for x in range(N):
        dens.append((lat.index((x, 0), 1), lat.index(((x + 1) % N, 0), 1), complex(J)))
    for x in range(N):
        i = lat.index((x, 0), 1)
        dens.append((i, i, complex(-J / 2)))
    return _finish(Model(lat, BOSONIC, bonds, dens, {"J": J},
                         n_filled=N // 2 if n_filled is None else n_filled))


def haldane_honeycomb(L1: int, L2: int, t: float = 1.0, tp: float = 0.6, tpp: float = -0.58,
                      phi_over_2pi: float = 0.2, V1: float = 0.0, V2: float = 0.0,
                      n_filled: Optional[int] = None) -> Model:
    """Hard-core bosons on the Haldane honeycomb (``test/bosonic_fci.jl:22-112``,
    ``benchmark/benchmark.jl:114-146``)."""
    lat = Lattice((L1, L2), 2, name="Haldane_Honeycomb", sub_crys=[(0.0, 0.0), (1 / 3, 1 / 3)])
    ph = 2.0 * math.pi * phi_over_2pi
    e = lambda s: complex(math.cos(s * ph), math.sin(s * ph))
    B = Bond
    bonds = [
        B((0, 0), 1, (0, 0), 2, -t), B((0, 0), 1, (0, -1), 2, -t), B((0, 0), 1, (-1, 0), 2, -t),
        B((0, 0), 1, (1, 0), 1, -tp * e(1)), B((0, 0), 1, (0, 1), 1, -tp * e(-1)),
        B((0, 0), 1, (-1, 1), 1, -tp * e(1)),
        B((0, 0), 2, (1, 0), 2, -tp * e(-1)), B((0, 0), 2, (0, 1), 2, -tp * e(1)),
        B((0, 0), 2, (-1, 1), 2, -tp * e(-1)),
        B((0, 0), 2, (1, 1), 1, -tpp), B((1, 0), 1, (0, 1), 2, -tpp), B((1, 0), 2, (0, 1), 1, -tpp),
    ]
    dens: List[Tuple[int, int, complex]] = []
    seen = set()

    def add_pair(i, j, V):
        if i == j:
            return
        p = (min(i, j), max(i, j))
        if p in seen:
            return
        seen.add(p)
        dens.append((p[0], p[1], complex(V)))

    if V1 != 0.0:
        for (cell, isub) in lat.site_list:
            if isub != 1:
                continue
            i = lat.index(cell, 1)
            for sh in ((0, 0), (0, -1), (-1, 0)):
                add_pair(i, lat.index((cell[0] + sh[0], cell[1] + sh[1]), 2), V1)
    if V2 != 0.0:
        for (cell, isub) in lat.site_list:
            for sh in ((1, 0), (0, 1), (-1, 1)):
                add_pair(lat.index(cell, isub), lat.index((cell[0] + sh[0], cell[1] + sh[1]), isub), V2)
    nf = (L1 * L2) // 2 if n_filled is None else n_filled
    return _finish(Model(lat, BOSONIC, bonds, dens,
                         {"t": t, "tp": tp, "tpp": tpp, "phi_over_2pi": phi_over_2pi,
                          "V1": V1, "V2": V2}, n_filled=nf))


def hubbard_square(Lx: int, Ly: int, U: float = 8.0, t: float = 1.0,
                   n_filled: Optional[int] = None, group_order: str = "dx_fastest") -> Model:
    """Spinful Fermi-Hubbard, spin-orbitals interleaved (up = 2i-1, down = 2i)
    (``benchmark/benchmark.jl:149-219``, ``examples/fermion_hubbard_square.jl:37-150``)."""
    lat = Lattice((Lx, Ly), 2, name="Square_Hubbard")
    spinless = Lattice((Lx, Ly), 1)
    sbonds = [Bond((0, 0), 1, (1, 0), 1, complex(-t)), Bond((0, 0), 1, (0, 1), 1, complex(-t))]
    model = Model(lat, FERMIONIC, [], [], {"t": t, "U": U},
                  n_filled=Lx * Ly if n_filled is None else n_filled, group_order=group_order)
    # Spin doubling of the spinless bonds: each spinless bond (cell_a -> cell_b) becomes
    # an up bond and a down bond between the same cells.
    for b in sbonds:
        model.bonds.append(Bond(b.Ra, 1, b.Rb, 1, b.t))
        model.bonds.append(Bond(b.Ra, 2, b.Rb, 2, b.t))
    del spinless
    for y in range(Ly):
        for x in range(Lx):
            model.density_terms.append((lat.index((x, y), 1), lat.index((x, y), 2), complex(U)))
    return _finish(model)


def checkerboard(L1: int, L2: int, t: float = 1.0, phi_over_2pi: float = 1 / 8,
                 tp1: float = 0.3, tp2: float = -0.3, tpp: float = -0.2,
                 V1: float = 1.0, V2: float = 0.0, V3: float = 0.0, lam: float = 1.0,
                 n_filled: Optional[int] = None, statistics: int = FERMIONIC) -> Model:
    """Spinless fermions on the checkerboard lattice (``phase_exploration/src/model.jl:1-71``)."""
    lat = Lattice((L1, L2), 2, name="checkerboard", sub_crys=[(0.5, 0.0), (0.0, 0.5)])
    ph = 2.0 * math.pi * phi_over_2pi
    e = lambda s: complex(math.cos(s * ph), math.sin(s * ph))
    B = Bond
    bonds = [
        B((0, 0), 1, (0, 0), 2, -t * e(-1)), B((0, 0), 1, (1, 0), 2, -t * e(1)),
        B((0, 0), 2, (0, 1), 1, -t * e(-1)), B((1, 0), 2, (0, 1), 1, -t * e(1)),
        B((0, 0), 1, (1, 0), 1, complex(-tp1)), B((0, 0), 1, (0, 1), 1, complex(-tp2)),
        B((0, 0), 2, (1, 0), 2, complex(-tp2)), B((0, 0), 2, (0, 1), 2, complex(-tp1)),
        B((0, 0), 1, (1, 1), 1, complex(-tpp)), B((0, 0), 2, (1, 1), 2, complex(-tpp)),
        B((1, 0), 2, (0, 1), 2, complex(-tpp)), B((0, 1), 1, (1, 0), 1, complex(-tpp)),
    ]
    v1, v2, v3 = lam * V1, lam * V2, lam * V3
    dens: List[Tuple[int, int, complex]] = []
    for i_from, (cf, sf) in enumerate(lat.site_list, start=1):
        for i_to, (ct, st) in enumerate(lat.site_list, start=1):
            if i_from == i_to:
                continue
            sh = ((ct[0] - cf[0]) % L1, (ct[1] - cf[1]) % L2)
            if v1 != 0 and sf == 1 and st == 2 and sh in ((0, 0), (1, 0), (0, L2 - 1), (1, L2 - 1)):
                dens.append((i_from, i_to, complex(v1)))
            if v2 != 0 and sf == st and sh in ((1, 0), (0, 1)):
                dens.append((i_from, i_to, complex(v2)))
            if v3 != 0 and sf == st and sh in ((1, 1), (1, L2 - 1)):
                dens.append((i_from, i_to, complex(v3)))
    nf = (L1 * L2) // 3 if n_filled is None else n_filled
    return _finish(Model(lat, statistics, bonds, dens,
                         {"t": t, "phi_over_2pi": phi_over_2pi, "tp1": tp1, "tp2": tp2,
                          "tpp": tpp, "V1": V1, "V2": V2, "V3": V3, "lam": lam}, n_filled=nf))


def model_group(model: Model, theta: Optional[Sequence[float]] = None) -> SymmetryGroup:
    th = model.theta if theta is None else theta
    return translation_group(model.lattice, th, order=model.group_order)


def term_arrays(model: Model):
    """(i_from, i_to, t, di, dj, V) as int32 / complex128 arrays, 1-based sites."""
    bl = model.bilinear_terms
    dl = model.density_terms
    i_from = np.array([a for a, _, _ in bl], dtype=np.int32)
    i_to = np.array([b for _, b, _ in bl], dtype=np.int32)
    t = np.array([c for _, _, c in bl], dtype=np.complex128)
    di = np.array([a for a, _, _ in dl], dtype=np.int32)
    dj = np.array([b for _, b, _ in dl], dtype=np.int32)
    V = np.array([c for _, _, c in dl], dtype=np.complex128)
    return i_from, i_to, t, di, dj, V


def reference_benchmark_model(model: str, label: str) -> Model:
    """The models of the reference's benchmark table (``benchmark/benchmark.jl:66-219``,
    rows of ``benchmark/benchmark_data/benchmark_raw.csv``): ``("Heisenberg", "N=20")``,
    ``("Haldane_Boson", "3×4")`` (hard-core bosons at one per two cells' worth of sites, the
    FCI parameters), ``("Hubbard_Fermion", "2×5")`` (U = 8, half filling)."""
    if model == "Heisenberg":
        return heisenberg_chain(int(label.split("=")[1]))
    a, b = (int(x) for x in label.split("×"))
    if model == "Haldane_Boson":
        return haldane_honeycomb(a, b)
    if model == "Hubbard_Fermion":
        return hubbard_square(a, b, U=8.0)
    raise KeyError((model, label))


# --------------------------------------------------------------------------
# Named workloads (sizes of the path's benchmark configurations)
# --------------------------------------------------------------------------
def workload(name: str) -> Model:
    """Models of the five benchmark configurations, by name.

    xxz16      : XXZ chain N=16, 8 bosons, Z16 (k=0 sector dim 810)
    hubbard4x4 : Hubbard 4x4, 32 spin-orbitals, 10 fermions, U=4, Z4xZ4
    checker4x6 : checkerboard 4x6 (48 sites), 8 fermions, Z4xZ6
    haldane3x4 : Haldane honeycomb 3x4 (24 sites), 6 hard-core bosons, Z3xZ4
    """
    if name == "xxz16":
        return heisenberg_chain(16)
    if name == "hubbard4x4":
        return hubbard_square(4, 4, U=4.0, t=1.0, n_filled=10, group_order="dx_fastest")
    if name == "checker4x6":
        return checkerboard(4, 6, t=1.0, phi_over_2pi=1 / 8, tp1=0.3, tp2=-0.3, tpp=-0.2,
                            V1=1.0, V2=0.0, V3=0.0, lam=1.0, n_filled=8)
    if name == "haldane3x4":
        return haldane_honeycomb(3, 4, t=1.0, tp=0.6, tpp=-0.58, phi_over_2pi=0.2, n_filled=6)
    raise KeyError(name)


# --------------------------------------------------------------------------
# Serving-stream request families
# --------------------------------------------------------------------------
# XXZ chains at half filling, 2 x L Hubbard ladders at quarter filling (U = 4),
# checkerboard lattices at 1/3 filling of the lower band.
STREAM_SIZES = {"xxz": [20, 22, 24, 26, 28, 30, 32], "ladder": [6, 7, 8, 9],
                "checker": [(3, 4), (3, 5), (3, 6), (4, 4), (4, 5), (4, 6)]}


def stream_model(family: str, size) -> Model:
    """Model of one serving-stream request family at one size."""
    if family == "xxz":
        return heisenberg_chain(size)
    if family == "ladder":
        return hubbard_square(size, 2, U=4.0, t=1.0, n_filled=size)
    if family == "checker":
        return checkerboard(size[0], size[1])
    raise KeyError(family)


def stream_requests(n: int, seed: int, sizes=None):
    """Seeded request list: family and size uniform, momentum draw u in [0, 1) (irrep index
    floor(u * |G|)), twist uniform in [0, 1)^2 (the chain uses only its x component)."""
    sizes = STREAM_SIZES if sizes is None else sizes
    rng = np.random.default_rng(seed)
    fams = list(sizes)
    out = []
    for _ in range(n):
        f = fams[int(rng.integers(len(fams)))]
        size = sizes[f][int(rng.integers(len(sizes[f])))]
        u = float(rng.random())
        th = (float(rng.random()), float(rng.random()))
        if f == "xxz":
            th = (th[0], 0.0)
        out.append({"family": f, "size": size, "u": u, "theta": th})
    return out


def stream_sector(request):
    """(model at its twist, group, irrep) of one stream request."""
    m = twisted(stream_model(request["family"], request["size"]), request["theta"])
    G = model_group(m, request["theta"])
    irs = translation_irreps(G, m.lattice)
    return m, G, irs[min(len(irs) - 1, int(request["u"] * len(irs)))]
