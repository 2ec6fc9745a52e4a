################################################################################
# TwoPIScalar_GPUrrule.jl -- gradients of the late-time occupation n(p, T)
# through the hand-written adjoint of the KB step (kb_adj_* in
# include/qftkb.h), as a ChainRulesCore rrule.  The reference has no
# derivative code; this is the rule an AD system (Zygote, Diffractor, ...)
# uses for a loss built on occupationGPU.
#
# Include it after TwoPIScalar_GPUfull.jl.  ChainRulesCore is not a
# dependency of the reference: add it to the environment first.
#
#   n = occupationGPU(problem, lambda, n0, T)   # n(p, T) per site
#   L = sum(w .* n)
#   dL/dlambda, dL/dn0 from the pullback of occupationGPU, one reverse sweep
#
# Mass2 is held fixed (an RMass counterterm is not differentiated).  The
# adjoint runs on full storage ("GPUfull") for the LambdaLO, LambdaNLO and
# LambdaNLOclassical expansions, on every lattice the back-end takes.
################################################################################
using ChainRulesCore

# the context of a run at coupling Lambda, initialised with the per-site
# occupations n0 at the discretisation mass (initSimData!, Initialization.jl:8-29)
function startGPU(problem::QFTdynamicsProblem, Lambda::Float64, n0::Vector{Float64})
    simdata = gpusimdata(problem, :kb_create; Lambda = Lambda)
    length(n0) == problem.disc.Nx^problem.disc.sdim || throw(DimensionMismatch("n0 needs one value per site"))
    kbcheck(simdata.ctx, ccall((:kb_init_particles, libqftkb), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble),
                               simdata.ctx, n0, problem.disc.Mass))
    return simdata
end

function stepsGPU!(simdata::TwoPIScalarSimDataGPUfull, T::Int)
    for t in 3:T
        kbcheck(simdata.ctx, ccall((:kb_step, libqftkb), Cint, (Ptr{Cvoid}, Cint), simdata.ctx, t))
    end
end

# n(p, T) per site (createMeasurement, Measurement.jl:38-59)
function occupationsGPU(simdata::TwoPIScalarSimDataGPUfull, T::Int)
    V = simdata.Nx^simdata.sdim
    n = zeros(V)
    out = Ref{KBMeas}()
    kbcheck(simdata.ctx, ccall((:kb_measure, libqftkb), Cint,
                               (Ptr{Cvoid}, Cint, Ref{KBMeas}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                               simdata.ctx, T, out, C_NULL, C_NULL, C_NULL, n, C_NULL, C_NULL))
    return n
end

"""
    occupationGPU(problem, Lambda, n0, T) -> Vector{Float64}

The occupation n(p, T) at every site after evolving the particle-number state
n0 (one value per site) with coupling Lambda from t = 3 to T.
"""
function occupationGPU(problem::QFTdynamicsProblem, Lambda::Float64, n0::Vector{Float64}, T::Int)
    simdata = startGPU(problem, Lambda, n0)
    stepsGPU!(simdata, T)
    return occupationsGPU(simdata, T)
end

# checkpoint_every: 0 tapes every column a step drops; C > 0 keeps the window
# every C steps and recomputes each segment; -1 packs the tape to one value
# per cubic orbit (cubic-symmetric states only)
function ChainRulesCore.rrule(::typeof(occupationGPU), problem::QFTdynamicsProblem, Lambda::Float64,
                              n0::Vector{Float64}, T::Int; checkpoint_every::Int = 50)
    simdata = startGPU(problem, Lambda, n0)
    kbcheck(simdata.ctx, ccall((:kb_adj_record, libqftkb), Cint, (Ptr{Cvoid}, Cint, Cint),
                               simdata.ctx, checkpoint_every, T))
    stepsGPU!(simdata, T)
    nT = occupationsGPU(simdata, T)
    swept = Ref(false)
    function occupationGPU_pullback(nbar)
        # a sweep returns the context to its initial rows (recording on): a
        # second pullback call runs the forward steps again first
        swept[] && stepsGPU!(simdata, T)
        swept[] = true
        wgt = collect(Float64, unthunk(nbar))   # dL/dn(p, T): the loss weights
        lbar = Ref{Cdouble}(0)
        n0bar = zeros(length(n0))
        kbcheck(simdata.ctx, ccall((:kb_adj_gradient, libqftkb), Cint,
                                   (Ptr{Cvoid}, Ptr{Float64}, Ref{Cdouble}, Ptr{Float64}, Ptr{Float64}),
                                   simdata.ctx, wgt, lbar, n0bar, C_NULL))
        return NoTangent(), NoTangent(), lbar[], n0bar, NoTangent()
    end
    return nT, occupationGPU_pullback
end
