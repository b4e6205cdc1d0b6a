# HelmholtzAMD.jl — Julia-side drop-in for Helmholtz.jl's SL-MG solver path,
# calling libhelmholtz_amd.so (include/helmholtz_amd.h) through ccall.
#
# A jInv plug-in exactly as the reference's solver is one
# (src/Helmholtz.jl:5-8, src/ShiftedLaplacianMultigridSolver.jl:4,17,32,104):
# ShiftedLaplacianMultigridSolverAMD <: jInv.LinearSolvers.AbstractSolver, and
# solveLinearSystem / copySolver / clear! are methods of jInv's own generic
# functions, so a jInv caller's solveLinearSystem(A, B, Ainv) dispatches here.
# Mirrors src/ShiftedLaplacianMultigridSolver.jl:4-109 (struct, constructor,
# copySolver, solveLinearSystem, clear!) and the operator apply of
# src/GetHelmholtz.jl:85-95.  No Julia toolchain exists where this file was
# written; it is checked against the C header (argument order and types) and
# for the plug-in declarations, and exercised through the equivalent Python
# ctypes binding.
module HelmholtzAMD

using LinearAlgebra
using SparseArrays
using jInv.Mesh
using jInv.LinearSolvers
import jInv.LinearSolvers.AbstractSolver
import jInv.LinearSolvers.solveLinearSystem
import jInv.LinearSolvers.copySolver
import jInv.Utils.clear!

export ShiftedLaplacianMultigridSolverAMD, getShiftedLaplacianMultigridSolverAMD
export solveLinearSystem, copySolver, clear!

const LIB = joinpath(@__DIR__, "..", "libhelmholtz_amd.so")

const HH_C128 = Cint(0); const HH_C64 = Cint(1)
const HH_HOST = Cint(0); const HH_DEVICE = Cint(1)
const HH_FGMRES = Cint(0)
const HH_BICGSTAB = Cint(1)

struct HHMGParams
    levels::Int32; pre::Int32; post::Int32; relax::Float64; cycle::Int32; max_outer::Int32
    tol::Float64; inner::Int32; krylov::Int32; prec::Int32; learned::Int32; verbose::Int32
    graph::Int32; autotune::Int32; fused::Int32; wlevels::Int32
    per_level::Int32; pre_l::NTuple{16,Int32}; post_l::NTuple{16,Int32}
    relax_type::Int32
    basis::Int32
end
const HH_RELAX_JAC = Cint(0); const HH_RELAX_JAC_GMRES = Cint(1)

# relaxPre / relaxPost: a count or a function of the level (finest = 1)
sweeps_at(v, l) = v isa Function ? Int32(v(l)) : Int32(v)
per_level_sweeps(v, levels) = ntuple(l -> l <= levels ? sweeps_at(v, l) : Int32(0), 16)

mutable struct HHSolveStats
    iters::Ptr{Int32}; status::Ptr{Int32}; true_res::Ptr{Float64}
    max_iters::Int32; n_converged::Int32; setup_time::Float64; solve_time::Float64; n_prec::Int64
    learned_used::Int32; fell_back::Int32
end

check(rc) = rc == 0 || error("helmholtz_amd: " * unsafe_string(ccall((:hh_last_error, LIB), Cstring, ())))

mutable struct HHOperator
    ptr::Ptr{Cvoid}
    ho_beta::Union{Nothing,Tuple{Float64,Float64}}  # (beta_lap, beta_mass) of a spread operator
    function HHOperator(nnodes::Vector{Int}, h::Vector{Float64}, m::Vector{Float64}, gamma::Vector{Float64},
                        omega::Union{Float64,ComplexF64}, neumann::Bool, sommerfeld::Bool, order::Int=2)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        nn = Int64[nnodes; ones(Int64, 3 - length(nnodes))]
        hh = Float64[h; ones(3 - length(h))]
        GC.@preserve m gamma check(ccall((:hh_op_create, LIB), Cint,
            (Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint, Cint, Cint,
             Ptr{Ptr{Cvoid}}),
            length(nnodes), nn, hh, m, gamma, real(omega), imag(omega), neumann, sommerfeld, order, out))
        op = new(out[], nothing)
        finalizer(o -> ccall((:hh_op_destroy, LIB), Cint, (Ptr{Cvoid},), o.ptr), op)
        return op
    end
    HHOperator(ptr::Ptr{Cvoid}, ho_beta) = new(ptr, ho_beta)
end

# GetHelmholtzOperatorHO(Msh, m, omega, gamma, NeumannAtFirstDim, Sommerfeld, beta)
# (src/GetHelmholtz.jl:54-72): the spread operator; beta scalar (2-D) or [beta_lap; beta_mass] (3-D)
function HHOperatorHO(nnodes::Vector{Int}, h::Vector{Float64}, m::Vector{Float64}, gamma::Vector{Float64},
                      omega::Union{Float64,ComplexF64}, neumann::Bool, sommerfeld::Bool, beta=1.0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    nn = Int64[nnodes; ones(Int64, 3 - length(nnodes))]
    hh = Float64[h; ones(3 - length(h))]
    bl, bm = length(beta) == 2 ? (beta[1], beta[2]) : (beta[1], beta[1])
    GC.@preserve m gamma check(ccall((:hh_op_create_ho, LIB), Cint,
        (Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint, Cint, Float64,
         Float64, Ptr{Ptr{Cvoid}}),
        length(nnodes), nn, hh, m, gamma, real(omega), imag(omega), neumann, sommerfeld, bl, bm, out))
    op = HHOperator(out[], (Float64(bl), Float64(bm)))
    finalizer(o -> ccall((:hh_op_destroy, LIB), Cint, (Ptr{Cvoid},), o.ptr), op)
    return op
end

# Y = (H + i*shift*w^2*diag(m)) X  — the product of getHelmholtzFun (src/GetHelmholtz.jl:85-95)
function apply!(Y::Array{ComplexF64}, op::HHOperator, X::Array{ComplexF64}, shift::Float64=0.0)
    nrhs = size(X, 2)
    GC.@preserve X Y check(ccall((:hh_op_apply, LIB), Cint,
        (Ptr{Cvoid}, Float64, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Cint, Cint),
        op.ptr, shift, X, Y, nrhs, HH_C128, HH_HOST))
    return Y
end

# jInv AbstractSolver plug-in with the fields of ShiftedLaplacianMultigridSolver
# (src/ShiftedLaplacianMultigridSolver.jl:4-15) plus the library handles
mutable struct ShiftedLaplacianMultigridSolverAMD <: AbstractSolver
    helmParam::Any          # Helmholtz.HelmholtzParam
    MG::Any                 # Multigrid.MGparam (levels, relaxParam, relaxPre/Post, cycleType, maxOuterIter, relativeTol)
    shift::Vector{Float64}
    Krylov::String
    inner::Int64
    doClear::Int64
    verbose::Bool
    setupTime::Real
    nPrec::Int
    solveTime::Real
    op::Union{Nothing,HHOperator}
    handle::Ptr{Cvoid}
    prec::Cint              # HH_C128 (plain) or HH_C64 (mixed precision preconditioner)
    basis::Cint             # FGMRES basis storage: HH_C128, or HH_C64 (needs prec = HH_C64)
end

getShiftedLaplacianMultigridSolverAMD(helmParam, MG, shift::Array{Float64}, Krylov::String="BiCGSTAB",
                                      inner::Int64=5, verbose::Bool=false; prec=HH_C128, basis=HH_C128) =
    ShiftedLaplacianMultigridSolverAMD(helmParam, MG, copy(vec(shift)), Krylov, inner, 0, verbose,
                                       0.0, 0, 0.0, nothing, C_NULL, prec, basis)
# a scalar shift is broadcast to every level (src/ShiftedLaplacianMultigridSolver.jl:28-30)
getShiftedLaplacianMultigridSolverAMD(helmParam, MG, shift::Float64, Krylov::String="BiCGSTAB",
                                      inner::Int64=5, verbose::Bool=false; prec=HH_C128, basis=HH_C128) =
    getShiftedLaplacianMultigridSolverAMD(helmParam, MG, fill(shift, MG.levels), Krylov, inner, verbose;
                                          prec=prec, basis=basis)

param_operator(p, order::Int) =
    HHOperator(collect(p.Mesh.n .+ 1), collect(p.Mesh.h), vec(p.m), vec(p.gamma), p.omega,
               p.NeumannOnTop, p.Sommerfeld, order)

# The operator a first solve builds from.  The reference's hierarchy comes from
# the matrix it is passed (MGsetup(ShiftedHT, ...), src/ShiftedLaplacianMultigridSolver.jl:65),
# so ShiftedHT is checked, not ignored: its shifted product on a fixed vector
# must equal that of helmParam's operator with orderNeumannBC 2 or 1 (the one
# that matches is used); ShiftedHT is the reference's sparse(SH') (SH = H +
# GetHelmholtzShiftOP(m, w, shift[1])), or an HHOperator applied with shift[1].
function matching_operator(s::ShiftedLaplacianMultigridSolverAMD, ShiftedHT)
    p = s.helmParam
    ShiftedHT === nothing && return param_operator(p, 2)
    n = prod(p.Mesh.n .+ 1)
    x = ComplexF64[complex(sin(0.7k), cos(1.3k)) for k in 1:n]
    want = ShiftedHT isa HHOperator ? apply!(similar(x), ShiftedHT, x, s.shift[1]) : ShiftedHT' * x
    size(want) == size(x) || error("ShiftedHT has the wrong size for helmParam.Mesh")
    # a spread operator (GetHelmholtzOperatorHO) is the outer operator of the
    # solve; it must be helmParam's, with its own beta
    cands = ShiftedHT isa HHOperator && ShiftedHT.ho_beta !== nothing ?
        [HHOperatorHO(collect(p.Mesh.n .+ 1), collect(p.Mesh.h), vec(p.m), vec(p.gamma), p.omega,
                      p.NeumannOnTop, p.Sommerfeld, collect(ShiftedHT.ho_beta))] :
        [param_operator(p, 2), param_operator(p, 1)]
    for op in cands
        got = apply!(similar(x), op, x, s.shift[1])
        norm(got - want) <= 1e-10 * norm(want) && return op
    end
    error("ShiftedHT does not match helmParam's shifted operator (orderNeumannBC 1 or 2, shift[1] = $(s.shift[1]))")
end

function ensure_handle!(s::ShiftedLaplacianMultigridSolverAMD, ShiftedHT=nothing)
    if s.op === nothing
        s.op = matching_operator(s, ShiftedHT)
    end
    if s.handle == C_NULL
        mg = s.MG
        pl = mg.relaxPre isa Function || mg.relaxPost isa Function
        prm = Ref(HHMGParams(mg.levels, sweeps_at(mg.relaxPre, 1), sweeps_at(mg.relaxPost, 1), mg.relaxParam,
                             Int32(mg.cycleType), mg.maxOuterIter, mg.relativeTol, s.inner,
                             s.Krylov == "BiCGSTAB" ? HH_BICGSTAB : HH_FGMRES, s.prec, 0, s.verbose, 0, 0, 1, 0,
                             pl ? 1 : 0, per_level_sweeps(mg.relaxPre, mg.levels),
                             per_level_sweeps(mg.relaxPost, mg.levels),
                             mg.relaxType == "Jac-GMRES" ? HH_RELAX_JAC_GMRES : HH_RELAX_JAC, s.basis))
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:hh_solver_create, LIB), Cint, (Ptr{Cvoid}, Ref{HHMGParams}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                    s.op.ptr, prm, s.shift, out))
        s.handle = out[]
        finalizer(x -> x.handle != C_NULL && ccall((:hh_solver_destroy, LIB), Cint, (Ptr{Cvoid},), x.handle), s)
    end
end

# solveLinearSystem(ShiftedHT, B, param, doTranspose) (src/ShiftedLaplacianMultigridSolver.jl:33-102),
# a method of jInv.LinearSolvers.solveLinearSystem.  doTranspose = 1 solves
# the adjoint system H^H X = B (:68-80).  As in the reference, ShiftedHT is
# read when the hierarchy is (re)built and ignored while it exists (:74).
function solveLinearSystem(ShiftedHT, B, param::ShiftedLaplacianMultigridSolverAMD, doTranspose::Int64=0)
    doTranspose in (0, 1) || error("doTranspose must be 0 or 1")
    if param.doClear == 1
        clear!(param)
    end
    if norm(B) == 0.0
        return zeros(eltype(B), size(B)), param
    end
    ensure_handle!(param, ShiftedHT)
    Bc = convert(Array{ComplexF64}, B)
    nrhs = size(Bc, 2)
    X = similar(Bc)
    iters = zeros(Int32, nrhs); status = zeros(Int32, nrhs); res = zeros(Float64, nrhs)
    GC.@preserve Bc X iters status res begin
        st = Ref(HHSolveStats(pointer(iters), pointer(status), pointer(res), 0, 0, 0.0, 0.0, 0, 0, 0))
        check(ccall((:hh_solve, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Cint, Cint, Ref{HHSolveStats}),
                    param.handle, Bc, X, nrhs, doTranspose, HH_HOST, st))
        param.setupTime += st[].setup_time
        param.solveTime += st[].solve_time
        param.nPrec += st[].n_prec
    end
    return (size(B, 2) == 1 ? vec(X) : X), param
end

# clear! (src/ShiftedLaplacianMultigridSolver.jl:104-109), a method of
# jInv.Utils.clear!: drops the hierarchy; the next solve rebuilds it and
# re-reads ShiftedHT
function clear!(s::ShiftedLaplacianMultigridSolverAMD)
    s.handle != C_NULL && check(ccall((:hh_solver_clear, LIB), Cint, (Ptr{Cvoid},), s.handle))
    s.op = nothing
    if s.handle != C_NULL
        ccall((:hh_solver_destroy, LIB), Cint, (Ptr{Cvoid},), s.handle)
        s.handle = C_NULL
    end
    s.doClear = 0
end

# copySolver (src/ShiftedLaplacianMultigridSolver.jl:17-22), a method of
# jInv.LinearSolvers.copySolver: same settings, no hierarchy
copySolver(s::ShiftedLaplacianMultigridSolverAMD) =
    ShiftedLaplacianMultigridSolverAMD(s.helmParam, s.MG, copy(s.shift), s.Krylov, s.inner, 0, s.verbose,
                                       0.0, 0, 0.0, nothing, C_NULL, s.prec, s.basis)

end # module
