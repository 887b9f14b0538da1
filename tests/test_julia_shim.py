"""CPU check of the Julia binding (julia/SectorED.jl) against the C ABI (include/sed.h):
every ccall names an exported entry point of the header, with its argument count and
argument / return types compatible with the C prototype, and the Julia mirror of
sed_solve_stats has the C struct's fields in order and type.  (Julia is not installed
here; the shim's calls are the ones tests/ makes through ctypes.)"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "julia", "SectorED.jl")
HEADER = os.path.join(ROOT, "include", "sed.h")

HANDLES = ("sed_context", "sed_catalog", "sed_sector", "sed_server")


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def header_prototypes():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"^\s*((?:const\s+)?[\w]+\s*\**)\s*(sed_\w+)\s*\(([^;]*?)\)\s*;", txt, flags=re.M):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3)
        params = [] if args.strip() in ("", "void") else [" ".join(a.split()) for a in args.split(",")]
        types = []
        for p in params:
            p = re.sub(r"\b\w+$", "", p).strip()          # drop the parameter name
            types.append(p.replace(" *", "*").replace("* ", "*"))
        protos[name] = (ret.replace(" *", "*"), types)
    return protos


def shim_ccalls():
    txt = open(SHIM).read()
    calls = []
    pos = 0
    while True:
        i = txt.find("ccall((:", pos)
        if i < 0:
            break
        j = i + len("ccall((:")
        name = re.match(r"\w+", txt[j:]).group(0)
        k = txt.index("),", j) + 2                          # after (:name, LIB),
        rest = txt[k:].lstrip()
        # return type up to the top-level comma before the argument tuple
        depth, r = 0, 0
        while not (rest[r] == "," and depth == 0):
            depth += rest[r] in "({[" and 1 or (rest[r] in ")}]" and -1 or 0)
            r += 1
        ret = rest[:r].strip()
        tup = rest[r + 1:].lstrip()
        assert tup[0] == "(", (name, tup[:40])
        depth = 0
        for e, ch in enumerate(tup):
            depth += 1 if ch == "(" else (-1 if ch == ")" else 0)
            if depth == 0:
                break
        args = _split_top(tup[1:e])
        calls.append((name, ret, args))
        pos = j
    return calls


def compatible(c, jl):
    c = c.replace("const ", "").strip()
    if c == "int":
        return jl == "Cint"
    if c == "int64_t":
        return jl == "Int64"
    if c == "uint64_t":
        return jl == "UInt64"
    if c == "double":
        return jl in ("Cdouble", "Float64")
    if c == "char*":
        return jl == "Cstring"
    base = c.rstrip("*")
    stars = len(c) - len(base)
    if base in HANDLES:
        return jl == "Ptr{Cvoid}" if stars == 1 else jl in ("Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}")
    if stars != 1:
        return False
    table = {"int": ("Ref{Cint}", "Ptr{Cint}"), "int32_t": ("Ptr{Int32}", "Ref{Int32}"),
             "int64_t": ("Ref{Int64}", "Ptr{Int64}"), "uint64_t": ("Ptr{UInt64}",),
             "double": ("Ptr{Float64}", "Ptr{ComplexF64}", "Ref{Float64}"), "float": ("Ptr{Float32}",),
             "void": ("Ptr{Cvoid}", "Ptr{Float64}", "Ptr{ComplexF64}", "Ptr{Float32}", "Ptr{ComplexF32}"),
             "sed_solve_stats": ("Ref{SolveStats}", "Ptr{SolveStats}")}
    return jl in table.get(base, ())


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = shim_ccalls()
    assert len(calls) >= 18
    for name, ret, args in calls:
        assert name in protos, name
        cret, ctypes_ = protos[name]
        assert len(args) == len(ctypes_), (name, args, ctypes_)
        assert compatible(cret, ret), (name, cret, ret)
        for a, c in zip(args, ctypes_):
            assert compatible(c, a), (name, c, a)
    used = {n for n, _, _ in calls}
    # the operator path a drop-in needs
    for fn in ("sed_create", "sed_catalog_build", "sed_sector_build", "sed_sector_set_terms", "sed_matvec",
               "sed_solve", "sed_catalog_copy", "sed_catalog_stabilizers", "sed_last_error"):
        assert fn in used, fn


def test_solve_stats_struct_mirrors_the_header():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct sed_solve_stats \{(.*?)\} sed_solve_stats;", txt, flags=re.S).group(1)
    cfields = re.findall(r"(\w+)\s+(\w+);", body)
    jl = re.search(r"struct SolveStats\n(.*?)\nend", open(SHIM).read(), flags=re.S).group(1)
    jfields = re.findall(r"(\w+)::(\w+)", jl)
    cmap = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64"}
    assert [(n, cmap[t]) for t, n in cfields] == jfields


def test_shim_extends_the_reference_methods_it_replaces():
    txt = open(SHIM).read()
    # methods on the reference's own functions and types, no overwrite of its signatures
    assert "function ED.hamiltonian_linear_operator(basis::ED.Symmetry_Sector_Basis" in txt
    assert "cmap::ED.CanonicalMap" in txt
    assert "function ED.diagonalize_block_matrixfree(H::GPUMatrixFreeHamiltonian, n::Int; nev::Int=5" in txt
    assert "ED.Symmetry_Orbit_Catalog(G, " in txt          # GPU catalog in the reference's own type
    assert "s.cat.n_site" not in txt
