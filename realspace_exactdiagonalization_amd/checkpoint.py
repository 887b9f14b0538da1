"""Checkpoints of sector scans: atomic per-scan files, a solver-revision gate, the
parameter-digest file names of flux scans, and the compatibility test used on resume.

Reference: ``save_checkpoint`` / ``load_checkpoint`` (``src/symmetry_resolved_ed.jl:1137-1160``),
``ed_scan_checkpoint_filename`` (``:1173-1183``), ``_same_flux_problem`` (``:1185-1196``), the
resume logic of ``_ed_scan_sectors!`` / ``resume_ed_scan!`` (``:1199-1277``).

The reference writes JLD2 (HDF5).  No HDF5 writer exists in this image, so a checkpoint here
is a NumPy ``.npz`` archive (loaded with ``allow_pickle=False``): a JSON header (solver
revision, the model, filling, group, irrep labels, sector dimensions) in the ``meta`` entry
and each finished sector's eigenvalues and eigenvectors as ``vals_<i>`` / ``vecs_<i>``
(0-based irrep index).  The orbit catalog is not stored: rebuilding it on the GPU takes a
fraction of a second.  The file name keeps the reference's format with the extension
``.npz`` in place of ``.jld2``.
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

SOLVER_REVISION = 2

_FNV_OFFSET = 0xCBF29CE484222325
_FNV_PRIME = 0x00000100000001B3
_M64 = (1 << 64) - 1


def fnv1a64(data: bytes) -> int:
    """64-bit FNV-1a hash (Fowler-Noll-Vo): h = (h xor byte) * prime per byte, from the
    offset basis, modulo 2^64 (the loop of ``:1178-1180``)."""
    h = _FNV_OFFSET
    for b in data:
        h = ((h ^ b) * _FNV_PRIME) & _M64
    return h


def julia_float(x: float) -> str:
    """A Float64 as Julia's ``repr`` prints it: the shortest round-trip digits, with a
    ``.0`` on integral values, in scientific form (``1.0e-5``, ``1.2345e6``) below 1e-4 and
    from 1e6 on in magnitude."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x != 0.0 and (abs(x) < 1e-4 or abs(x) >= 1e6):
        s = np.format_float_scientific(x, unique=True, trim="0", exp_digits=1)
        mant, ex = s.split("e")
        return f"{mant}e{int(ex)}"
    r = repr(x)
    if "e" in r:   # Python's own scientific range starts at 1e16
        r = np.format_float_positional(x, unique=True, trim="0")
    if "." not in r:
        r += ".0"
    return r


def julia_repr(v) -> str:
    """``repr`` of the values a model's parameter dict holds (strings, Bool, Int,
    Float64, ComplexF64) and of (key, value) pairs and vectors of them."""
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"').replace("$", "\\$") + '"'
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return julia_float(v)
    if isinstance(v, (complex, np.complexfloating)):
        im = v.imag
        sign = "-" if (im < 0 or (im == 0 and math.copysign(1.0, im) < 0)) else "+"
        return f"{julia_float(v.real)} {sign} {julia_float(abs(im))}im"
    if isinstance(v, tuple) and len(v) == 2 and isinstance(v[0], str):
        return f"{julia_repr(v[0])} => {julia_repr(v[1])}"
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(julia_repr(x) for x in v) + "]"
    raise TypeError(f"no Julia repr for {type(v).__name__}")


def _pairs_repr(params: Dict) -> str:
    """``repr(sort(collect(params); by=x->string(first(x))))``: a vector of key => value
    pairs sorted by key; a vector whose values mix types prints its element type first."""
    items = sorted(params.items(), key=lambda kv: str(kv[0]))
    body = julia_repr([(str(k), v) for k, v in items])
    kinds = {type(v) for _, v in items}
    if len(kinds) > 1:
        num = all(k in (float, int, bool) for k in kinds)
        body = f"Pair{{String, {'Real' if num else 'Any'}}}" + body
    return body


def ed_scan_checkpoint_filename(model, twisted_phases_over_2pi: Sequence[float], filling_fraction: Fraction,
                                identity: bool = False) -> str:
    """``ed_scan_checkpoint_filename`` (``:1173-1183``): ``solver_v2_<model>_<sample size>
    _nu=<p>_<q>_flux=<theta>_p=<FNV-1a-64 of the sorted parameter pairs, hex>``, the
    ``identity_`` prefix for identity-group scans (``:1334``), extension ``.npz``."""
    h = fnv1a64(_pairs_repr(model.params).encode("utf-8"))
    ff = Fraction(filling_fraction)
    size = julia_repr(list(model.lattice.sample_size))
    flux = julia_repr([float(t) for t in twisted_phases_over_2pi])
    name = (f"solver_v2_{model.lattice.name}_{size}_nu={ff.numerator}_{ff.denominator}_flux={flux}"
            f"_p={h:x}.npz")
    return ("identity_" + name) if identity else name


# ---------------------------------------------------------------- model (de)serialisation
def _c(z) -> List[float]:
    z = complex(z)
    return [z.real, z.imag]


def model_to_json(model) -> dict:
    lat = model.lattice
    return {
        "lattice": {"sample_size": list(lat.sample_size), "n_sub": lat.n_sub, "pbc": list(lat.pbc),
                    "name": lat.name, "sub_crys": [list(p) for p in lat.sub_crys]},
        "statistics": int(model.statistics),
        "bonds": [[list(b.Ra), b.sa, list(b.Rb), b.sb, _c(b.t)] for b in model.bonds],
        "density_terms": [[int(i), int(j), _c(v)] for i, j, v in model.density_terms],
        "params": {str(k): ({"re": complex(v).real, "im": complex(v).imag} if isinstance(v, complex) else v)
                   for k, v in model.params.items()},
        "n_filled": int(model.n_filled),
        "group_order": model.group_order,
        "theta": [float(model.theta[0]), float(model.theta[1])],
        "extra_bilinear": None if model.extra_bilinear is None else
        [[int(i), int(j), _c(t)] for i, j, t in model.extra_bilinear],
        "bilinear_terms": [[int(i), int(j), _c(t)] for i, j, t in model.bilinear_terms],
    }


def model_from_json(d: dict):
    from .lattice import Bond, Lattice, Model
    L = d["lattice"]
    lat = Lattice(L["sample_size"], L["n_sub"], pbc=L["pbc"], name=L["name"], sub_crys=L["sub_crys"])
    cz = lambda p: complex(p[0], p[1])
    m = Model(lat, int(d["statistics"]),
              [Bond(tuple(a), int(sa), tuple(b), int(sb), cz(t)) for a, sa, b, sb, t in d["bonds"]],
              [(int(i), int(j), cz(v)) for i, j, v in d["density_terms"]],
              params={k: (complex(v["re"], v["im"]) if isinstance(v, dict) else v) for k, v in d["params"].items()},
              n_filled=int(d["n_filled"]),
              group_order=d["group_order"], theta=tuple(d["theta"]),
              extra_bilinear=None if d["extra_bilinear"] is None else
              [(int(i), int(j), cz(t)) for i, j, t in d["extra_bilinear"]])
    m.bilinear_terms = [(int(i), int(j), cz(t)) for i, j, t in d["bilinear_terms"]]
    return m


def same_flux_problem(a, b) -> bool:
    """``_same_flux_problem`` (``:1185-1196``): same parameters, sample size, twist and
    statistics, and term lists equal after sorting, amplitudes within 1e-12."""
    if not (a.params == b.params and tuple(a.lattice.sample_size) == tuple(b.lattice.sample_size)
            and tuple(float(t) for t in a.theta) == tuple(float(t) for t in b.theta)
            and a.statistics == b.statistics):
        return False
    key = lambda t: (t[0], t[1], complex(t[2]).real, complex(t[2]).imag)
    for aa, bb in ((a.bilinear_terms, b.bilinear_terms), (a.density_terms, b.density_terms)):
        if len(aa) != len(bb):
            return False
        for x, y in zip(sorted(aa, key=key), sorted(bb, key=key)):
            if x[0] != y[0] or x[1] != y[1]:
                return False
            dx = abs(complex(x[2]) - complex(y[2]))
            if dx > 1e-12 + 1e-12 * max(abs(complex(x[2])), abs(complex(y[2]))):
                return False
    return True


# ---------------------------------------------------------------- files
class ScanState:
    """What a checkpoint holds: the model, filling, group name, irrep labels, sector
    dimensions and the finished sectors' (eigenvalues, eigenvectors)."""

    def __init__(self, model, n_filled: int, filling_fraction: Fraction, group_name: str,
                 irrep_labels: List[Tuple[int, int]], sector_dims: List[int],
                 ed_scan_res: Dict[int, Tuple[np.ndarray, np.ndarray]]):
        self.model = model
        self.n_filled = n_filled
        self.filling_fraction = Fraction(filling_fraction)
        self.group_name = group_name
        self.irrep_labels = [tuple(l) for l in irrep_labels]
        self.sector_dims = list(sector_dims)
        self.ed_scan_res = ed_scan_res


def save_checkpoint(state: ScanState, path: str) -> None:
    """``save_checkpoint`` (``:1137-1151``): written to the fixed sibling ``path + '.tmp'``
    and renamed over ``path`` (atomic on POSIX), the temporary removed on any failure, so an
    interrupted write leaves the previous checkpoint intact."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    meta = {"solver_revision": SOLVER_REVISION, "model": model_to_json(state.model), "n_filled": state.n_filled,
            "filling_fraction": [state.filling_fraction.numerator, state.filling_fraction.denominator],
            "group_name": state.group_name, "irrep_labels": [list(l) for l in state.irrep_labels],
            "sector_dims": [int(x) for x in state.sector_dims], "done": sorted(int(i) for i in state.ed_scan_res)}
    arrays = {"meta": np.frombuffer(json.dumps(meta).encode("utf-8"), dtype=np.uint8)}
    for i, (vals, vecs) in state.ed_scan_res.items():
        arrays[f"vals_{int(i)}"] = np.asarray(vals, dtype=np.float64)
        arrays[f"vecs_{int(i)}"] = np.asarray(vecs, dtype=np.complex128)
    tmp = path + ".tmp"
    try:
        with open(tmp, "wb") as f:
            np.savez(f, **arrays)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def load_checkpoint(path: str) -> ScanState:
    """``load_checkpoint`` (``:1154-1160``): refuses files of another solver revision."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode("utf-8"))
        if meta.get("solver_revision", 0) != SOLVER_REVISION:
            raise RuntimeError(f"Legacy ED checkpoint {path}: recompute with overwrite=True (flux/projection fix).")
        res = {int(i): (np.array(z[f"vals_{i}"]), np.array(z[f"vecs_{i}"])) for i in meta["done"]}
    ff = Fraction(meta["filling_fraction"][0], meta["filling_fraction"][1])
    return ScanState(model_from_json(meta["model"]), int(meta["n_filled"]), ff, meta["group_name"],
                     [tuple(l) for l in meta["irrep_labels"]], meta["sector_dims"], res)
