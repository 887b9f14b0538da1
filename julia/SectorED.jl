# SectorED.jl -- routes the matrix-free sector path of RealSpace_ExactDiagonalization to
# libsed.so (gfx950 HIP kernels behind the C ABI of include/sed.h).
#
# Load it after the package:   using RealSpace_ExactDiagonalization; include("julia/SectorED.jl")
#
# It adds methods, it overwrites none: `hamiltonian_linear_operator` gets methods for the
# concrete term-list types a Real_Space_Second_Quantized_Model holds (more specific than
# the package's `Vector{<:Tuple{Int,Int,<:Number}}` method, src/symmetry_resolved_ed.jl:891-897),
# returning a GPU operator, and `diagonalize_block_matrixfree` gets a method for that
# operator type (the package's method takes an untyped operator, :998-1030).  So
# `ed_scan_at_irrep_matrixfree!` (:1092-1128), `ed_scan!` and every caller above them run
# unchanged on the package's own Symmetry_Sector_Basis / CanonicalMap / OneDim_Irrep:
# the GPU orbit catalog is built once per Symmetry_Orbit_Catalog (from the CanonicalMap's
# group and statistics, cached by object identity), the GPU sector once per irrep.
# `build_symmetry_orbit_catalog_gpu` builds the package's Symmetry_Orbit_Catalog itself on
# the GPU (its tables are bit-exact with the CPU build), for build_ed_data's catalog step.
# Set SectorED.ENABLED[] = false to fall back to the CPU methods.
module SectorED

using LinearAlgebra, SparseArrays
using ..RealSpace_ExactDiagonalization
const ED = RealSpace_ExactDiagonalization

const LIB = get(ENV, "SED_LIB", joinpath(@__DIR__, "..", "realspace_exactdiagonalization_amd", "libsed.so"))
const ENABLED = Ref(true)
const PRECISION = Ref{Cint}(0)          # 0 FP64, 2 FP32 + FP64 refinement (sed.h SED_PREC_*)
const DEVICE = Ref{Cint}(0)

check(rc) = rc == 0 || error(unsafe_string(ccall((:sed_last_error, LIB), Cstring, ())))

# ---------------------------------------------------------------- handles
mutable struct Ctx
    h::Ptr{Cvoid}
end
function Ctx(device::Integer)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sed_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r))
    c = Ctx(r[])
    finalizer(x -> ccall((:sed_destroy, LIB), Cint, (Ptr{Cvoid},), x.h), c)
    return c
end
const CTX = Ref{Union{Nothing,Ctx}}(nothing)
context() = (CTX[] === nothing && (CTX[] = Ctx(DEVICE[])); CTX[])

mutable struct GPUCatalog
    h::Ptr{Cvoid}
    ctx::Ctx
    n_site::Int
    n_filled::Int
    n_group::Int
    n_total::Int
    n_orbits::Int
end

mutable struct GPUSector
    h::Ptr{Cvoid}
    cat::GPUCatalog
    n_site::Int
    dim::Int
end

_perm(G::ED.Finite_Symmetry_Group) = Int32[op.perm[i] for op in G.operations for i in 1:G.n_site]
_phases(G::ED.Finite_Symmetry_Group) = ComplexF64[op.perm_phases[i] for op in G.operations for i in 1:G.n_site]
_stats(s::ED.Particle_Statistics) = s isa ED.Fermionic ? Cint(1) : Cint(0)

function gpu_catalog(G::ED.Finite_Symmetry_Group, n_filled::Integer, stats::ED.Particle_Statistics)
    ctx = context()
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sed_catalog_build, LIB), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Int32}, Ptr{ComplexF64}, Ref{Ptr{Cvoid}}),
        ctx.h, G.n_site, n_filled, _stats(stats), length(G.operations), G.identity_idx, _perm(G), _phases(G), r))
    nt, no = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:sed_catalog_info, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), r[], nt, no))
    c = GPUCatalog(r[], ctx, G.n_site, n_filled, length(G.operations), nt[], no[])
    finalizer(x -> ccall((:sed_catalog_destroy, LIB), Cint, (Ptr{Cvoid},), x.h), c)
    return c
end

function gpu_sector(cat::GPUCatalog, irrep::ED.OneDim_Irrep; atol::Float64=1e-12)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    d = Ref{Int64}(0)
    check(ccall((:sed_sector_build, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Cdouble, Ref{Ptr{Cvoid}}, Ref{Int64}),
        cat.h, irrep.values, atol, r, d))
    s = GPUSector(r[], cat, cat.n_site, d[])
    finalizer(x -> ccall((:sed_sector_destroy, LIB), Cint, (Ptr{Cvoid},), x.h), s)
    return s
end

# one GPU catalog per package catalog, one GPU sector per (catalog, irrep label)
const CATALOGS = IdDict{ED.Symmetry_Orbit_Catalog,GPUCatalog}()
const SECTORS = IdDict{ED.Symmetry_Orbit_Catalog,Dict{Any,GPUSector}}()

function gpu_sector_for(basis::ED.Symmetry_Sector_Basis, cmap::ED.CanonicalMap)
    cat = cmap.catalog
    n_filled = isempty(cat.representative_mask_list) ? 0 : count_ones(cat.representative_mask_list[1])
    g = get!(() -> gpu_catalog(cmap.symmetry_group, n_filled, cmap.particle_statistics), CATALOGS, cat)
    secs = get!(() -> Dict{Any,GPUSector}(), SECTORS, cat)
    s = get!(() -> gpu_sector(g, basis.irrep), secs, basis.irrep.label)
    s.dim == length(basis.representative_mask_list) || error("GPU sector basis differs from the package's")
    return s
end

# ---------------------------------------------------------------- operator
"GPU sector operator: callable as H(x) like the package's MatrixFreeHamiltonian (:880-884)."
struct GPUMatrixFreeHamiltonian
    s::GPUSector
end

function set_terms!(s::GPUSector, bilinear_terms, density_terms)
    i1 = Int32[a for (a, _, _) in bilinear_terms]
    i2 = Int32[b for (_, b, _) in bilinear_terms]
    t = ComplexF64[c for (_, _, c) in bilinear_terms]
    d1 = Int32[a for (a, _, _) in density_terms]
    d2 = Int32[b for (_, b, _) in density_terms]
    v = ComplexF64[c for (_, _, c) in density_terms]
    check(ccall((:sed_sector_set_terms, LIB), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{Int32}, Ptr{ComplexF64}, Cint, Ptr{Int32}, Ptr{Int32}, Ptr{ComplexF64}),
        s.h, length(i1), i1, i2, t, length(d1), d1, d2, v))
    return s
end

const TermList{T} = Vector{Tuple{Int,Int,T}}

for T in (ComplexF64, Float64)
    @eval function ED.hamiltonian_linear_operator(basis::ED.Symmetry_Sector_Basis, bilinear_terms::TermList{$T},
                                                  density_terms::TermList{$T}, cmap::ED.CanonicalMap)
        ENABLED[] || return invoke(ED.hamiltonian_linear_operator,
            Tuple{ED.Symmetry_Sector_Basis,Vector{<:Tuple{Int,Int,<:Number}},Vector{<:Tuple{Int,Int,<:Number}},
                  ED.CanonicalMap}, basis, bilinear_terms, density_terms, cmap)
        s = set_terms!(gpu_sector_for(basis, cmap), bilinear_terms, density_terms)
        return GPUMatrixFreeHamiltonian(s), s.dim
    end
end

function (H::GPUMatrixFreeHamiltonian)(x::Vector{ComplexF64})
    y = similar(x)
    check(ccall((:sed_matvec, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Cint), H.s.h, x, y, 0))
    return y
end

# sed_solve_stats (include/sed.h): 4 x int32, int64, 2 x double, 2 x int32 = 48 bytes
struct SolveStats
    converged::Int32
    iters_fp32::Int32
    iters_fp64::Int32
    restarts::Int32
    matvecs::Int64
    residual_max::Float64
    seconds::Float64
    full_orth_steps::Int32
    repeated::Int32
end
SolveStats() = SolveStats(0, 0, 0, 0, 0, 0.0, 0.0, 0, 0)

const LAST_STATS = Ref(SolveStats())

"diagonalize_block_matrixfree (:998-1030) on the GPU: dense below dimension 500 as the package does, thick-restart Lanczos above, residual tolerance tol."
function ED.diagonalize_block_matrixfree(H::GPUMatrixFreeHamiltonian, n::Int; nev::Int=5, tol::Float64=1e-10,
                                         seed::UInt64=UInt64(1234), maxiter::Int=300,
                                         precision::Integer=PRECISION[])
    n == 0 && return Float64[], Matrix{ComplexF64}(undef, 0, 0)
    k = min(nev, n)
    vals = zeros(Float64, k)
    vecs = zeros(ComplexF64, n, k)
    st = Ref(SolveStats())
    check(ccall((:sed_solve, LIB), Cint,
        (Ptr{Cvoid}, Cint, Cdouble, UInt64, Cint, Cint, Ptr{Float64}, Ptr{ComplexF64}, Ref{SolveStats}),
        H.s.h, k, tol, seed, maxiter, precision, vals, vecs, st))
    LAST_STATS[] = st[]
    return vals, vecs
end

# ---------------------------------------------------------------- catalog on the GPU
"""
    build_symmetry_orbit_catalog_gpu(; second_quantized_model, n_filled, symmetry_group, particle_statistics)

The package's `Symmetry_Orbit_Catalog` (:305-327), built on the GPU and copied back in the
package's layout (rep_rank / rep_sym / rep_amp tables, representatives, stabilizers); same
keywords and result as `build_symmetry_orbit_catalog` (:371-487).
"""
function build_symmetry_orbit_catalog_gpu(; second_quantized_model, n_filled::Int,
                                          symmetry_group::ED.Finite_Symmetry_Group,
                                          particle_statistics::ED.Particle_Statistics)
    G = symmetry_group
    @assert G.n_site == second_quantized_model.lattice.n_site
    g = gpu_catalog(G, n_filled, particle_statistics)
    rr = zeros(Int32, g.n_total)
    rs = zeros(Int32, g.n_total)
    ra = zeros(ComplexF64, g.n_total)
    reps = zeros(UInt64, g.n_orbits)
    so = zeros(Int64, g.n_orbits)
    check(ccall((:sed_catalog_copy, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, Ptr{ComplexF64}, Ptr{UInt64}, Ptr{Int64}), g.h, rr, rs, ra, reps, so))
    gi = zeros(Int32, g.n_group * g.n_orbits)
    gp = zeros(ComplexF64, g.n_group * g.n_orbits)
    check(ccall((:sed_catalog_stabilizers, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{ComplexF64}), g.h, gi, gp))
    gidx = [Int[gi[(o - 1) * g.n_group + j] for j in 1:so[o]] for o in 1:g.n_orbits]
    gph = [ComplexF64[gp[(o - 1) * g.n_group + j] for j in 1:so[o]] for o in 1:g.n_orbits]
    n_site = G.n_site
    lin_left, lin_right, lin_n_right = n_site <= ED.LIN_TABLE_MAX_N_SITE ? ED._lin_table(n_site, n_filled) :
                                       (Int[], Int[], -1)
    cat = ED.Symmetry_Orbit_Catalog(G, Vector{ED.Mask}(reps), Vector{Int}(so), gidx, gph, rr, rs, ra,
                                    lin_left, lin_right, lin_n_right)
    CATALOGS[cat] = g
    return cat
end

# ---------------------------------------------------------------- observables and matrix mode
# eigenvector `which` (1-based) of the last solve on the sector of H
function density_correlation(H::GPUMatrixFreeHamiltonian, which::Int=1)
    n = H.s.n_site
    nn = zeros(Float64, n * n)
    d = zeros(Float64, n)
    check(ccall((:sed_sector_density_correlation, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}),
        H.s.h, which - 1, nn, d))
    return permutedims(reshape(nn, n, n)), d                         # C[i, j] = <n_i n_j>
end

function off_diagonal_long_range_order(H::GPUMatrixFreeHamiltonian, which::Int=1)
    n = H.s.n_site
    rho = zeros(ComplexF64, n * n)
    check(ccall((:sed_sector_one_body, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{ComplexF64}), H.s.h, which - 1, rho))
    return permutedims(reshape(rho, n, n))                           # rho[i, j] = <a_i^dag a_j>
end

function build_ed_Hamiltonian_symmetry_block_gpu(basis::ED.Symmetry_Sector_Basis, bilinear_terms, density_terms,
                                                 cmap::ED.CanonicalMap)
    s = set_terms!(gpu_sector_for(basis, cmap), bilinear_terms, density_terms)
    nnz = Ref{Int64}(0)
    check(ccall((:sed_sector_matrix_csc, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}),
        s.h, nnz, C_NULL, C_NULL, C_NULL))
    colptr = zeros(Int64, s.dim + 1)
    rowval = zeros(Int64, nnz[])
    nzval = zeros(ComplexF64, nnz[])
    check(ccall((:sed_sector_matrix_csc, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}),
        s.h, nnz, colptr, rowval, nzval))
    return SparseMatrixCSC(s.dim, s.dim, colptr, rowval, nzval)     # 1-based already
end

function expand_sector_state_to_fock_amplitudes(c::Vector{ComplexF64}, basis::ED.Symmetry_Sector_Basis,
                                               cmap::ED.CanonicalMap; atol::Float64=1e-12)
    s = gpu_sector_for(basis, cmap)
    n = Ref{Int64}(0)
    check(ccall((:sed_sector_fock_expansion, LIB), Cint,
        (Ptr{Cvoid}, Ptr{ComplexF64}, Cdouble, Int64, Ref{Int64}, Ptr{UInt64}, Ptr{ComplexF64}),
        s.h, c, atol, 0, n, C_NULL, C_NULL))
    masks = zeros(UInt64, n[])
    amps = zeros(ComplexF64, n[])
    check(ccall((:sed_sector_fock_expansion, LIB), Cint,
        (Ptr{Cvoid}, Ptr{ComplexF64}, Cdouble, Int64, Ref{Int64}, Ptr{UInt64}, Ptr{ComplexF64}),
        s.h, c, atol, n[], n, masks, amps))
    return Dict{UInt64,ComplexF64}(zip(masks, amps))
end

function position_operator_matrix(to::ED.Symmetry_Sector_Basis, from::ED.Symmetry_Sector_Basis,
                                  cmap::ED.CanonicalMap, site_phases::Vector{ComplexF64})
    st, sf = gpu_sector_for(to, cmap), gpu_sector_for(from, cmap)
    rows = zeros(Int64, sf.dim)
    vals = zeros(ComplexF64, sf.dim)
    check(ccall((:sed_sector_position_operator, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Int64}, Ptr{ComplexF64}), st.h, sf.h, site_phases, rows, vals))
    keep = findall(j -> rows[j] > 0 && abs(vals[j]) > 1e-13, 1:sf.dim)
    return sparse(rows[keep], keep, vals[keep], st.dim, sf.dim)
end

"update_orbit_stabilizer_phases! (:338-369) on the GPU catalog paired with `cat`."
function update_gpu_phases!(cat::ED.Symmetry_Orbit_Catalog, new_group::ED.Finite_Symmetry_Group)
    g = CATALOGS[cat]
    check(ccall((:sed_catalog_update_phases, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), g.h, _phases(new_group)))
    delete!(SECTORS, cat)
    return cat
end

end # module
