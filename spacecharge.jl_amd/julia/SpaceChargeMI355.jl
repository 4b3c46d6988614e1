# SpaceChargeMI355.jl -- Julia binding of libspacecharge_mi355.so for SpaceCharge.jl.
#
# Loaded next to SpaceCharge.jl, it adds a device-array type `MIArray` and
# more-specific methods of the reference's exported functions
# (src/SpaceCharge.jl:17) for meshes whose arrays are MIArrays, each a single
# ccall into the C ABI (include/spacecharge_mi355.h).  Meshes on Julia's CPU
# or CUDA backends keep the reference methods.
#
#   using SpaceCharge, SpaceChargeMI355
#   x, y, z, q = (MIArray(a) for a in (xh, yh, zh, qh))   # host -> MI355X
#   mesh = mi355_mesh((128,128,128), x, y, z; T=Float32)  # bounds as Mesh3D(grid, x, y, z)
#   deposit!(mesh, x, y, z, q); solve!(mesh); Ex, Ey, Ez = interpolate_field(mesh, x, y, z)
#   SpaceChargeMI355.synchronize()
module SpaceChargeMI355

using SpaceCharge
import SpaceCharge: deposit!, clear_mesh!, solve!, interpolate_field, Mesh3D

export MIArray, mi355_mesh, solve_potential, autotune_deposit!, deposit_batched!, solve_batched!,
       interpolate_batched, set_grouped_launch

const libsc = joinpath(@__DIR__, "..", "libspacecharge_mi355.so")
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

synchronize() = check(ccall((:sc_synchronize, libsc), Cint, (Ptr{Cvoid},), ctx()), ctx())

# ---------------------------------------------------------------- device arrays
"""Device array on the MI355X (column-major, like Julia's Array)."""
mutable struct MIArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    owner::Union{Nothing,Any}   # keeps a parent (mesh) alive for views it owns
end

Base.size(a::MIArray) = a.dims
Base.IndexStyle(::Type{<:MIArray}) = IndexLinear()
Base.getindex(::MIArray, ::Int) = error("scalar indexing of MIArray: use Array(a)")
Base.pointer(a::MIArray) = a.ptr

function MIArray{T}(::UndefInitializer, dims::Vararg{Int,N}) where {T,N}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sc_alloc, libsc), Cint, (Ptr{Cvoid}, Csize_t, Ptr{Ptr{Cvoid}}), ctx(),
                sizeof(T) * prod(dims), p), ctx())
    a = MIArray{T,N}(Ptr{T}(p[]), dims, nothing)
    finalizer(a) do b
        b.owner === nothing && ccall((:sc_free, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx(), b.ptr)
    end
    return a
end

function MIArray(h::Array{T,N}) where {T<:Union{Float32,Float64},N}
    a = MIArray{T}(undef, size(h)...)
    check(ccall((:sc_memcpy_h2d, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                ctx(), a.ptr, pointer(h), sizeof(h)), ctx())
    return a
end

function Base.Array(a::MIArray{T,N}) where {T,N}
    h = Array{T,N}(undef, a.dims)
    check(ccall((:sc_memcpy_d2h, libsc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                ctx(), pointer(h), a.ptr, sizeof(h)), ctx())
    return h
end

Base.similar(a::MIArray{T}) where {T} = MIArray{T}(undef, a.dims...)

# ---------------------------------------------------------------- mesh
# The C mesh owns rho/efield and the solver workspace; the Julia Mesh3D holds
# MIArray views of them and keeps the handle alive through `owner`.
mutable struct MeshHandle
    h::Ptr{Cvoid}
end

const HANDLES = IdDict{Any,MeshHandle}()

function new_mesh(::Type{T}, grid::NTuple{3,Int}) where {T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sc_mesh_create, libsc), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Ptr{Cvoid}}),
                ctx(), dtcode(T), grid[1], grid[2], grid[3], h), ctx())
    mh = MeshHandle(h[])
    finalizer(m -> ccall((:sc_mesh_destroy, libsc), Cint, (Ptr{Cvoid},), m.h), mh)
    rho = MIArray{T,3}(Ptr{T}(ccall((:sc_mesh_rho, libsc), Ptr{Cvoid}, (Ptr{Cvoid},), mh.h)),
                       grid, mh)
    ef = MIArray{T,4}(Ptr{T}(ccall((:sc_mesh_efield, libsc), Ptr{Cvoid}, (Ptr{Cvoid},), mh.h)),
                      (grid..., 3), mh)
    return mh, rho, ef
end

"""Mesh3D on the MI355X with bounds from the particle extrema (src/mesh.jl:95-174)."""
function mi355_mesh(grid::NTuple{3,Int}, x::MIArray{P}, y::MIArray{P}, z::MIArray{P};
                    T::Type{<:AbstractFloat}=Float64, gamma::Real=1.0,
                    total_charge::Real=0.0) where {P}
    length(x) == length(y) == length(z) ||
        error("Particle coordinate arrays must have the same length.")
    lo, hi, dl = zeros(3), zeros(3), zeros(3)
    check(ccall((:sc_bounds_from_particles, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Cint,
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx(), dtcode(P), x.ptr, y.ptr, z.ptr, length(x), collect(Int64, grid), dtcode(T),
                lo, hi, dl), ctx())
    mh, rho, ef = new_mesh(T, grid)
    mesh = Mesh3D{T,typeof(rho),typeof(ef)}(grid, Tuple(T.(lo)), Tuple(T.(hi)), Tuple(T.(dl)),
                                            T(gamma), T(total_charge), rho, ef, nothing)
    HANDLES[mesh] = mh
    return mesh
end

"""Mesh3D on the MI355X with manual bounds (src/mesh.jl:196-238)."""
function mi355_mesh(grid::NTuple{3,Int}, min_bounds::NTuple{3,Real}, max_bounds::NTuple{3,Real};
                    T::Type{<:AbstractFloat}=Float64, gamma::Real=1.0, total_charge::Real=0.0)
    any(grid .<= 1) && error("All elements of grid_size must be at least 2.")
    any(max_bounds .<= min_bounds) &&
        error("max_bounds must be strictly greater than min_bounds for all dimensions.")
    lo, hi = T.(min_bounds), T.(max_bounds)
    dl = (hi .- lo) ./ (grid .- 1)
    mh, rho, ef = new_mesh(T, grid)
    mesh = Mesh3D{T,typeof(rho),typeof(ef)}(grid, lo, hi, dl, T(gamma), T(total_charge), rho, ef,
                                            nothing)
    HANDLES[mesh] = mh
    return mesh
end

handle(mesh) = HANDLES[mesh].h
d3(t) = Float64[t[1], t[2], t[3]]

# ---------------------------------------------------------------- the four calls
const MIMesh{T} = Mesh3D{T,<:MIArray,<:MIArray}

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
end

function solve!(mesh::MIMesh; at_cathode::Bool=false)
    check(ccall((:sc_solve, libsc), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Cint),
                handle(mesh), d3(mesh.min_bounds), d3(mesh.max_bounds), d3(mesh.delta),
                Float64(mesh.gamma), at_cathode), ctx())
end

function interpolate_field(mesh::MIMesh, x::MIArray{P}, y::MIArray{P}, z::MIArray{P}) where {P}
    Ex, Ey, Ez = similar(x), similar(x), similar(x)   # particle eltype, src/interpolation.jl:110-112
    check(ccall((:sc_interpolate, libsc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Float64},
                 Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                handle(mesh), dtcode(P), x.ptr, y.ptr, z.ptr, length(x), d3(mesh.min_bounds),
                d3(mesh.delta), Ex.ptr, Ey.ptr, Ez.ptr), ctx())
    return Ex, Ey, Ez
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
