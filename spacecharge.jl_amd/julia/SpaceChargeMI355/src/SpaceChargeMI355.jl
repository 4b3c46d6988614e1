# SpaceChargeMI355.jl -- the MI355X backend of SpaceCharge.jl.
#
# Loaded next to SpaceCharge.jl, it adds a device-array type `MIArray`, a
# KernelAbstractions backend tag `MI355X`, and more-specific methods of the
# reference's exported functions (src/SpaceCharge.jl:17) for meshes whose
# arrays are MIArrays, each a ccall into the C ABI
# (include/spacecharge_mi355.h) of libspacecharge_mi355.so.  The reference's
# own Mesh3D constructors build such meshes unchanged:
#
#   * Mesh3D(grid, x, y, z; T) with MIArray particles: the default backend
#     get_backend(x) (src/mesh.jl:101) is MI355X(), extrema runs on the device
#     (:120-122), and KernelAbstractions.zeros(MI355X(), T, ...) allocates
#     rho and efield as MIArrays (:159-160);
#   * Mesh3D(grid, x, y, z; T, backend=MI355X()) with host particles, and
#     Mesh3D(grid, min_bounds, max_bounds; T, backend=MI355X()) (:196-238).
#
# The library's mesh object (which owns the solver workspace) is created on
# first use and kept in the mesh's `_workspace` slot, where the reference
# keeps its FFT workspace (src/mesh.jl:49-72); it is freed with the mesh.
# Meshes on Julia's CPU or CUDA backends keep the reference methods.
#
#   using SpaceCharge, SpaceChargeMI355
#   x, y, z, q = MIArray(xh), MIArray(yh), MIArray(zh), MIArray(qh)
#   mesh = Mesh3D((128, 128, 128), x, y, z; T=Float32)
#   deposit!(mesh, x, y, z, q); solve!(mesh); Ex, Ey, Ez = interpolate_field(mesh, x, y, z)
#   KernelAbstractions.synchronize(MI355X())
#
# Differentiation: space_charge_fields (the whole step) and the functional
# forms deposit_charge, solve_field and interpolate_at have ChainRulesCore
# rrules in the package extension ext/SpaceChargeMI355ChainRulesCoreExt.jl,
# loaded when ChainRulesCore is.
module SpaceChargeMI355

using SpaceCharge
import SpaceCharge: deposit!, clear_mesh!, solve!, interpolate_field, Mesh3D
import KernelAbstractions
const KA = KernelAbstractions

export MI355X, MIArray, solve_potential, space_charge_fields, deposit_charge, solve_field,
       interpolate_at, deposit_adjoint, solve_adjoint, interpolate_adjoint, autotune_deposit!,
       deposit_batched!, solve_batched!, interpolate_batched, set_grouped_launch

# the shared library: $SPACECHARGE_MI355_LIB, else the in-tree build next to the package
const libsc = get(ENV, "SPACECHARGE_MI355_LIB",
                  normpath(joinpath(@__DIR__, "..", "..", "..", "libspacecharge_mi355.so")))
const SC_F32 = Cint(4)
const SC_F64 = Cint(8)

dtcode(::Type{Float32}) = SC_F32
dtcode(::Type{Float64}) = SC_F64

# ---------------------------------------------------------------- context
const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function ctx()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:sc_ctx_create, libsc), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, h), C_NULL)
        CTX[] = h[]
    end
    return CTX[]
end

function check(rc::Cint, c::Ptr{Cvoid})
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sc_last_error, libsc), Cstring, (Ptr{Cvoid},), c))
    error(msg)  # ErrorException with the reference's message, as src/mesh.jl:107 etc.
end

"""Blocks until the MI355X has finished the work queued so far."""
synchronize() = check(ccall((:sc_synchronize, libsc), Cint, (Ptr{Cvoid},), ctx()), ctx())

# ---------------------------------------------------------------- device arrays
"""Device array on the MI355X (column-major, like Julia's Array)."""
mutable struct MIArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    owner::Union{Nothing,Any}   # a parent that owns the memory, or nothing (owned here)
end

Base.size(a::MIArray) = a.dims
Base.IndexStyle(::Type{<:MIArray}) = IndexLinear()
Base.getindex(::MIArray, ::Int) = error("scalar indexing of MIArray: use Array(a)")
Base.setindex!(::MIArray, v, ::Int) = error("scalar indexing of MIArray: use copyto!")
Base.pointer(a::MIArray) = a.ptr
Base.unsafe_convert(::Type{Ptr{T}}, a::MIArray{T}) where {T} = a.ptr

function MIArray{T}(::UndefInitializer, dims::Vararg{Int,N}) where {T,N}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sc_alloc, libsc), Cint, (Ptr{Cvoid}, Csize_t, Ptr{Ptr{Cvoid}}), ctx(),
                max(sizeof(T) * prod(dims; init=1), 1), p), ctx())
    a = MIArray{T,N}(Ptr{T}(p[]), dims, nothing)
    finalizer(a) do b
        b.owner === nothing &&
            ccall((:sc_free, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx(), b.ptr)
    end
    return a
end
MIArray{T}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N} = MIArray{T}(undef, dims...)

function MIArray(h::Array{T,N}) where {T<:Union{Float32,Float64},N}
    a = MIArray{T}(undef, size(h)...)
    copyto!(a, h)
    return a
end

function Base.copyto!(dst::MIArray{T}, src::Array{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! between different lengths"))
    check(ccall((:sc_memcpy_h2d, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                ctx(), dst.ptr, pointer(src), sizeof(src)), ctx())
    return dst
end

function Base.copyto!(dst::Array{T}, src::MIArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! between different lengths"))
    check(ccall((:sc_memcpy_d2h, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                ctx(), pointer(dst), src.ptr, sizeof(dst)), ctx())
    return dst
end

function Base.copyto!(dst::MIArray{T}, src::MIArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! between different lengths"))
    check(ccall((:sc_memcpy_d2d, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                ctx(), dst.ptr, src.ptr, sizeof(T) * length(src)), ctx())
    return dst
end

Base.Array(a::MIArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, a.dims), a)
Base.similar(a::MIArray{T}) where {T} = MIArray{T}(undef, a.dims...)
Base.similar(::MIArray, ::Type{S}, dims::Dims) where {S} = MIArray{S}(undef, dims...)
Base.copy(a::MIArray) = copyto!(similar(a), a)

"""fill!(a, 0): the only fill the device arrays need (clear_mesh!, src/deposition.jl:10-12,
and KernelAbstractions.zeros)."""
function Base.fill!(a::MIArray{T}, v) where {T}
    iszero(v) || error("fill! of an MIArray supports zero only (copyto! from a host array)")
    check(ccall((:sc_memset, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Csize_t),
                ctx(), a.ptr, 0, sizeof(T) * length(a)), ctx())
    return a
end

"""extrema of a device array (src/mesh.jl:120-122 calls it on the particles)."""
function Base.extrema(a::MIArray{T}) where {T<:Union{Float32,Float64}}
    mn, mx = Ref(0.0), Ref(0.0)
    check(ccall((:sc_extrema, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}),
                ctx(), dtcode(T), a.ptr, length(a), mn, mx), ctx())
    return T(mn[]), T(mx[])
end

# broadcasting: `dst .= src` (a device copy, as in the reference's GPU tests,
# test/test_gpu.jl:58) and `dst .= 0`; anything else needs a host round trip
struct MIStyle{N} <: Broadcast.AbstractArrayStyle{N} end
MIStyle{M}(::Val{N}) where {N,M} = MIStyle{N}()
Base.BroadcastStyle(::Type{<:MIArray{T,N}}) where {T,N} = MIStyle{N}()
function Base.copyto!(dst::MIArray, bc::Broadcast.Broadcasted{<:MIStyle})
    bc = Broadcast.flatten(bc)
    if bc.f === identity && length(bc.args) == 1
        src = bc.args[1]
        src isa Number && return fill!(dst, src)
        src isa MIArray && size(src) == size(dst) && return copyto!(dst, src)
    end
    error("broadcasting on MIArray supports copies (dst .= src) and dst .= 0 only")
end

# ---------------------------------------------------------------- backend
"""KernelAbstractions backend tag of the MI355X: get_backend of an MIArray,
and the `backend` keyword of the reference's Mesh3D constructors."""
struct MI355X <: KA.GPU end

KA.get_backend(::MIArray) = MI355X()
KA.allocate(::MI355X, ::Type{T}, dims::Tuple; unified::Bool=false) where {T} =
    (unified && error("unified memory is not supported on the MI355X backend");
     MIArray{T}(undef, dims...))
KA.zeros(b::MI355X, ::Type{T}, dims::Tuple; kwargs...) where {T} =
    fill!(KA.allocate(b, T, dims; kwargs...), zero(T))
KA.synchronize(::MI355X) = synchronize()

# ---------------------------------------------------------------- mesh
# The library's mesh object wraps mesh.rho and mesh.efield and owns the
# solver workspace.  It lives in the reference's lazy workspace slot
# (src/mesh.jl:31, 49-72), so it is created on first use and freed with the
# Mesh3D; it holds the arrays so that they outlive it.
mutable struct MeshHandle
    h::Ptr{Cvoid}
    rho::MIArray
    efield::MIArray
end

const MIMesh{T} = Mesh3D{T,<:MIArray,<:MIArray}

function handle(mesh::MIMesh{T}) where {T}
    ws = mesh._workspace
    if ws !== nothing && haskey(ws, :mi355)
        mh = ws.mi355
        mh.rho === mesh.rho && mh.efield === mesh.efield && return mh.h
    end
    size(mesh.rho) == mesh.grid_size && size(mesh.efield) == (mesh.grid_size..., 3) ||
        error("rho and efield do not match the mesh's grid_size")
    h = Ref{Ptr{Cvoid}}(C_NULL)
    nx, ny, nz = mesh.grid_size
    check(ccall((:sc_mesh_create_with_buffers, libsc), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                ctx(), dtcode(T), nx, ny, nz, mesh.rho.ptr, mesh.efield.ptr, h), ctx())
    mh = MeshHandle(h[], mesh.rho, mesh.efield)
    finalizer(m -> ccall((:sc_mesh_destroy, libsc), Cint, (Ptr{Cvoid},), m.h), mh)
    mesh._workspace = (mi355 = mh,)
    return mh.h
end

d3(t) = Float64[t[1], t[2], t[3]]

# Particles of a mesh on the MI355X must be on the device; host arrays are
# copied there (the reference would fail launching a GPU kernel on them).
device(a::MIArray) = a
device(a::AbstractArray{T}) where {T<:Union{Float32,Float64}} = MIArray(Array(a))
device(a::AbstractArray) = MIArray(Array{Float64}(a))

# ---------------------------------------------------------------- the reference's calls
clear_mesh!(mesh::MIMesh) = check(ccall((:sc_clear_mesh, libsc), Cint, (Ptr{Cvoid},),
                                        handle(mesh)), ctx())

function deposit!(mesh::MIMesh, x::MIArray{P}, y::MIArray{P}, z::MIArray{P}, q::MIArray{P};
                  clear::Bool=true) where {P}
    length(x) == length(y) == length(z) == length(q) ||
        error("Particle coordinate and charge arrays must have the same length.")
    check(ccall((:sc_deposit, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                 Ptr{Float64}, Ptr{Float64}, Cint),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, q.ptr, length(x),
                d3(mesh.min_bounds), d3(mesh.delta), clear), ctx())
    return nothing
end

function deposit!(mesh::MIMesh, x::AbstractVector, y::AbstractVector, z::AbstractVector,
                  q::AbstractVector; clear::Bool=true)
    length(x) == length(y) == length(z) == length(q) ||
        error("Particle coordinate and charge arrays must have the same length.")
    xd = device(x)
    P = eltype(xd)
    deposit!(mesh, xd, device(P.(Array(y))), device(P.(Array(z))), device(P.(Array(q)));
             clear=clear)
end

function solve!(mesh::MIMesh; at_cathode::Bool=false)
    check(ccall((:sc_solve, libsc), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Cint),
                handle(mesh), d3(mesh.min_bounds), d3(mesh.max_bounds), d3(mesh.delta),
                Float64(mesh.gamma), at_cathode), ctx())
    return nothing
end

function interpolate_field(mesh::MIMesh, x::MIArray{P}, y::MIArray{P}, z::MIArray{P}) where {P}
    length(x) == length(y) == length(z) ||
        error("Particle coordinate arrays must have the same length.")
    Ex, Ey, Ez = similar(x), similar(x), similar(x)   # particle eltype, src/interpolation.jl:110-112
    check(ccall((:sc_interpolate, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Float64},
                 Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, length(x), d3(mesh.min_bounds),
                d3(mesh.delta), Ex.ptr, Ey.ptr, Ez.ptr), ctx())
    return Ex, Ey, Ez
end

# host particles: the fields come back as host arrays, like similar(particles_x)
function interpolate_field(mesh::MIMesh, x::AbstractVector, y::AbstractVector,
                           z::AbstractVector)
    xd = device(x)
    P = eltype(xd)
    E = interpolate_field(mesh, xd, device(P.(Array(y))), device(P.(Array(z))))
    return Array(E[1]), Array(E[2]), Array(E[3])
end

# Scalar potential in free space (no reference counterpart; its Green function
# is SpaceCharge.potential_green_function, src/green_functions.jl:13-22):
# the lab-frame potential, an MIArray of the mesh's grid size and eltype.
function solve_potential(mesh::MIMesh{T}) where {T}
    phi = MIArray{T}(undef, mesh.grid_size...)
    check(ccall((:sc_solve_potential, libsc), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Cvoid}),
                handle(mesh), d3(mesh.delta), Float64(mesh.gamma), phi.ptr), ctx())
    return phi
end

# ---------------------------------------------------------------- differentiation
# No reference counterpart.  Vector-Jacobian products of the three calls
# with the mesh geometry held fixed; the ChainRulesCore extension builds its
# rrules from them.

"""g_q, g_x, g_y, g_z of deposit! at the charge cotangent g_rho (grid
size, mesh type).  With `into = (gx, gy, gz)` the position terms are added
to those arrays instead of new ones."""
function deposit_adjoint(mesh::MIMesh{T}, x::MIArray{P}, y::MIArray{P}, z::MIArray{P},
                         q::MIArray{P}, g_rho::MIArray{T}; into=nothing) where {T,P}
    gq = similar(x)
    gx, gy, gz = into === nothing ? (similar(x), similar(x), similar(x)) : into
    check(ccall((:sc_deposit_adjoint, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                 Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Cint),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, q.ptr, length(x),
                d3(mesh.min_bounds), d3(mesh.delta), g_rho.ptr, gq.ptr, gx.ptr, gy.ptr, gz.ptr,
                into === nothing ? Cint(0) : Cint(1)), ctx())
    return gq, gx, gy, gz
end

"""g_rho = S^T g_efield for the field solve S of solve!(mesh; at_cathode)."""
function solve_adjoint(mesh::MIMesh{T}, g_efield::MIArray{T}; at_cathode::Bool=false) where {T}
    g_rho = MIArray{T}(undef, mesh.grid_size...)
    check(ccall((:sc_solve_adjoint, libsc), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Cint, Ptr{Cvoid},
                 Ptr{Cvoid}),
                handle(mesh), d3(mesh.min_bounds), d3(mesh.max_bounds), d3(mesh.delta),
                Float64(mesh.gamma), at_cathode, g_efield.ptr, g_rho.ptr), ctx())
    return g_rho
end

"""g_efield, g_x, g_y, g_z of interpolate_field at the mesh's current
efield, for the field cotangents gEx, gEy, gEz (particle type)."""
function interpolate_adjoint(mesh::MIMesh{T}, x::MIArray{P}, y::MIArray{P}, z::MIArray{P},
                             gEx::MIArray{P}, gEy::MIArray{P}, gEz::MIArray{P}) where {T,P}
    g_ef = MIArray{T}(undef, mesh.grid_size..., 3)
    gx, gy, gz = similar(x), similar(x), similar(x)
    check(ccall((:sc_interpolate_adjoint, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Float64},
                 Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, length(x), d3(mesh.min_bounds),
                d3(mesh.delta), gEx.ptr, gEy.ptr, gEz.ptr, g_ef.ptr, gx.ptr, gy.ptr, gz.ptr,
                Cint(0)), ctx())
    return g_ef, gx, gy, gz
end

"""
    space_charge_fields(mesh, x, y, z, q; at_cathode=false) -> (Ex, Ey, Ez)

One space-charge step as a function: deposit!(mesh, x, y, z, q),
solve!(mesh; at_cathode), interpolate_field(mesh, x, y, z).  Differentiable
in x, y, z and q through ChainRulesCore (the mesh geometry is a constant).
"""
function space_charge_fields(mesh::MIMesh, x::MIArray, y::MIArray, z::MIArray, q::MIArray;
                             at_cathode::Bool=false)
    deposit!(mesh, x, y, z, q)
    solve!(mesh; at_cathode=at_cathode)
    return interpolate_field(mesh, x, y, z)
end

"""deposit! as a function: a copy of the deposited rho (grid size, mesh type)."""
function deposit_charge(mesh::MIMesh, x::MIArray, y::MIArray, z::MIArray, q::MIArray)
    deposit!(mesh, x, y, z, q)
    return copy(mesh.rho)
end

"""solve! of a given rho as a function: a copy of the field (grid size x 3)."""
function solve_field(mesh::MIMesh{T}, rho::MIArray{T}; at_cathode::Bool=false) where {T}
    copyto!(mesh.rho, rho)
    solve!(mesh; at_cathode=at_cathode)
    return copy(mesh.efield)
end

"""interpolate_field of a given field (grid size x 3, mesh type) as a function."""
function interpolate_at(mesh::MIMesh{T}, efield::MIArray{T}, x::MIArray, y::MIArray,
                        z::MIArray) where {T}
    copyto!(mesh.efield, efield)
    return interpolate_field(mesh, x, y, z)
end

# ---------------------------------------------------------------- no reference counterpart
# Picks the binned deposit's tiling for this mesh by timing the deposit of
# these particles at every setting; leaves mesh.rho holding that deposit.
# Returns the 5x4 mean times in ms (chunk_log2 10:14 by slice_log2 10:13).
function autotune_deposit!(mesh::MIMesh, x::MIArray{P}, y::MIArray{P}, z::MIArray{P},
                           q::MIArray{P}; reps::Integer=5) where {P}
    times = zeros(Float32, 4, 5)   # column-major: [slice, chunk]
    check(ccall((:sc_mesh_autotune_deposit, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                 Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float32}),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, q.ptr, length(x),
                d3(mesh.min_bounds), d3(mesh.delta), Cint(reps), times), ctx())
    return permutedims(times)
end

# Many bunches per call: the grouped launches of the C ABI (one launch per
# phase for all bunches; same-shape small meshes solve as one class).
ptrs(as) = Ptr{Cvoid}[a.ptr for a in as]
flat3(f, meshes) = Float64[Float64(v) for m in meshes for v in f(m)]

function deposit_batched!(meshes::Vector{<:MIMesh}, xs::Vector{<:MIArray{P}},
                          ys::Vector{<:MIArray{P}}, zs::Vector{<:MIArray{P}},
                          qs::Vector{<:MIArray{P}}; clear::Bool=true) where {P}
    check(ccall((:sc_deposit_batched, libsc), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                 Ptr{Ptr{Cvoid}}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Cint),
                [handle(m) for m in meshes], length(meshes), dtcode(P), ptrs(xs), ptrs(ys), ptrs(zs),
                ptrs(qs), Int64[length(x) for x in xs], flat3(m -> m.min_bounds, meshes),
                flat3(m -> m.delta, meshes), clear), ctx())
end

function solve_batched!(meshes::Vector{<:MIMesh}; at_cathode=false)
    flags = at_cathode isa Bool ? fill(Cint(at_cathode), length(meshes)) : Cint.(at_cathode)
    check(ccall((:sc_solve_batched, libsc), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Cint}),
                [handle(m) for m in meshes], length(meshes), flat3(m -> m.min_bounds, meshes),
                flat3(m -> m.max_bounds, meshes), flat3(m -> m.delta, meshes),
                Float64[Float64(m.gamma) for m in meshes], flags), ctx())
end

function interpolate_batched(meshes::Vector{<:MIMesh}, xs::Vector{<:MIArray{P}},
                             ys::Vector{<:MIArray{P}}, zs::Vector{<:MIArray{P}}) where {P}
    Exs, Eys, Ezs = ([similar(x) for x in xs] for _ in 1:3)
    check(ccall((:sc_interpolate_batched, libsc), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                 Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                 Ptr{Ptr{Cvoid}}),
                [handle(m) for m in meshes], length(meshes), dtcode(P), ptrs(xs), ptrs(ys), ptrs(zs),
                Int64[length(x) for x in xs], flat3(m -> m.min_bounds, meshes),
                flat3(m -> m.delta, meshes), ptrs(Exs), ptrs(Eys), ptrs(Ezs)), ctx())
    return Exs, Eys, Ezs
end

# phases: 1 = deposit and interpolation phases, 2 = same-shape solves, 3 = both (default)
set_grouped_launch(phases::Integer=3) =
    check(ccall((:sc_ctx_set_grouped_launch, libsc), Cint, (Ptr{Cvoid}, Cint), ctx(), Cint(phases)),
          ctx())

end # module
