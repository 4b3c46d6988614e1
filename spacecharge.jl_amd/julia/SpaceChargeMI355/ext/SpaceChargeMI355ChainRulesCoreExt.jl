# ChainRulesCore rules of the MI355X space-charge step (loaded with ChainRulesCore).
#
# The reference has no derivatives.  These rules make the step usable in
# differentiable tracking (Zygote and other ChainRules-based AD): each pullback
# is the library's hand-written vector-Jacobian products, with the mesh
# geometry (bounds, spacing, gamma) held fixed:
#   interpolation^T = a deposit of the field cotangents with the same CIC
#                     weights, plus the weights' position derivatives;
#   solve^T         = correlation with the Green function (with the cathode
#                     image when at_cathode is set);
#   deposit^T       = a gather of the charge cotangent, plus q times the
#                     weights' position derivatives.
# The mesh is mutable state, so each rule saves the field its pullback needs
# (the interpolation's position derivatives read it) and restores it there.
module SpaceChargeMI355ChainRulesCoreExt

using ChainRulesCore
using SpaceChargeMI355
using SpaceChargeMI355: MIArray, MIMesh, deposit_adjoint, solve_adjoint, interpolate_adjoint

# A cotangent as a device array like `like`: zero tangents become zeros.
function dev_cotangent(d, like::MIArray{T}) where {T}
    d = unthunk(d)
    if d isa AbstractZero
        return fill!(similar(like), zero(T))
    end
    d isa MIArray{T} && size(d) == size(like) && return d
    return copyto!(similar(like), Array{T}(d))
end

# The three field cotangents of an (Ex, Ey, Ez) output.
function field_cotangents(dE, x::MIArray)
    dE = unthunk(dE)
    if dE isa AbstractZero
        z = dev_cotangent(dE, x)
        return z, copy(z), copy(z)
    end
    return dev_cotangent(dE[1], x), dev_cotangent(dE[2], x), dev_cotangent(dE[3], x)
end

function ChainRulesCore.rrule(::typeof(space_charge_fields), mesh::MIMesh, x::MIArray,
                              y::MIArray, z::MIArray, q::MIArray; at_cathode::Bool=false)
    E = space_charge_fields(mesh, x, y, z, q; at_cathode=at_cathode)
    efield = copy(mesh.efield)
    function space_charge_fields_pullback(dE)
        gEx, gEy, gEz = field_cotangents(dE, x)
        copyto!(mesh.efield, efield)
        g_ef, gx, gy, gz = interpolate_adjoint(mesh, x, y, z, gEx, gEy, gEz)
        g_rho = solve_adjoint(mesh, g_ef; at_cathode=at_cathode)
        gq, = deposit_adjoint(mesh, x, y, z, q, g_rho; into=(gx, gy, gz))
        return NoTangent(), NoTangent(), gx, gy, gz, gq
    end
    return E, space_charge_fields_pullback
end

function ChainRulesCore.rrule(::typeof(deposit_charge), mesh::MIMesh, x::MIArray, y::MIArray,
                              z::MIArray, q::MIArray)
    rho = deposit_charge(mesh, x, y, z, q)
    function deposit_charge_pullback(drho)
        g_rho = dev_cotangent(drho, rho)
        gq, gx, gy, gz = deposit_adjoint(mesh, x, y, z, q, g_rho)
        return NoTangent(), NoTangent(), gx, gy, gz, gq
    end
    return rho, deposit_charge_pullback
end

function ChainRulesCore.rrule(::typeof(solve_field), mesh::MIMesh, rho::MIArray;
                              at_cathode::Bool=false)
    efield = solve_field(mesh, rho; at_cathode=at_cathode)
    function solve_field_pullback(defield)
        g_ef = dev_cotangent(defield, efield)
        return NoTangent(), NoTangent(), solve_adjoint(mesh, g_ef; at_cathode=at_cathode)
    end
    return efield, solve_field_pullback
end

function ChainRulesCore.rrule(::typeof(interpolate_at), mesh::MIMesh, efield::MIArray,
                              x::MIArray, y::MIArray, z::MIArray)
    E = interpolate_at(mesh, efield, x, y, z)
    function interpolate_at_pullback(dE)
        gEx, gEy, gEz = field_cotangents(dE, x)
        copyto!(mesh.efield, efield)
        g_ef, gx, gy, gz = interpolate_adjoint(mesh, x, y, z, gEx, gEy, gEz)
        return NoTangent(), NoTangent(), g_ef, gx, gy, gz
    end
    return E, interpolate_at_pullback
end

end # module
