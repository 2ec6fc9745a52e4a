################################################################################
# TwoPIScalar_GPUfull.jl -- the "GPUfull" numerics back-end for the
# reference's src/TwoPIScalar model: SimData lives on an MI355X and every
# hot-path method calls libqftkb.so (include/qftkb.h) through ccall.
#
# Include it after TwoPIScalar_SimDataStruct.jl / _Initialization.jl /
# _Evolution.jl / _Measurement.jl; three more lines wire it into
# TwoPIScalar_Modelfile.jl (the include, the "Num" => "GPUfull" switch at
# :104-108, and the getTwoPISimData branch).  Methods dispatch on the new
# types exactly like the CPUfull ones they stand beside:
#   getTwoPISimData          (TwoPIScalar_SimDataStruct.jl:669-708; also
#                             "Num" => "GPUorbits" for cubic-orbit storage)
#   getTwoPIScalarTmpData    (:710-721)
#   initialize!              (TwoPIScalar_Initialization.jl:289-325; the
#                             reference's method takes CPU TmpData only)
#   initSimData!             (:8-29, Gauss :95-112)
#   getHartreeMass2          (:207-218)
#   evolve!                  (TwoPIScalar_Evolution.jl:924-981)
#   createMeasurement        (TwoPIScalar_Measurement.jl:31-99)
################################################################################
const libqftkb = get(ENV, "QFTKB_LIB", joinpath(@__DIR__, "..", "qftkb", "libqftkb.so"))

export TwoPIScalarGPUfull, TwoPIScalarGPUorbits
struct TwoPIScalarGPUfull <: TwoPIScalarNumerics
    nchunks::Int      # kept for parameter compatibility; unused on the GPU
end
# one stored value per orbit of the lattice's cubic group (kb_create_sym):
# exact for the cubic-symmetric states every initialisation produces, the
# counterpart of CPUred ("Num" => "GPUorbits")
struct TwoPIScalarGPUorbits <: TwoPIScalarNumerics
    nchunks::Int
end

# mirrors kb_params in include/qftkb.h (field order and C types)
struct KBParams
    Nx::Cint
    sdim::Cint
    NstepsinMemory::Cint
    ONgroup::Cint
    dt::Cdouble
    Lambda::Cdouble
    Mass::Cdouble
    Mass2::Cdouble
    pexp::Cint
    device::Cint
end

struct KBMeas
    time::Cdouble
    E::Cdouble
    E_pot::Cdouble
    E_kin::Cdouble
    E_mem::Cdouble
    M2::Cdouble
    n_unstable::Cint
end

export TwoPIScalarSimDataGPUfull
mutable struct TwoPIScalarSimDataGPUfull <: TwoPIScalarSimDataGPU
    ctx::Ptr{Cvoid}
    kL2values::Vector{Float64}
    NstepsinMemory::Int
    Nx::Int
    sdim::Int
end

struct TwoPIScalarTmpDataGPUfull <: TwoPIScalarTmpData end

function kbcheck(ctx::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:kb_last_error, libqftkb), Cstring, (Ptr{Cvoid},), ctx))
    rc == 2 && throw(BoundsError("libqftkb: " * msg))
    throw(ErrorException("libqftkb error $rc: " * msg))
end

pexpcode(::TwoPIScalarLambdaLO) = Cint(0)
pexpcode(::TwoPIScalarLambdaNLOquantum) = Cint(1)
pexpcode(::TwoPIScalarNinverseLO) = Cint(2)
pexpcode(::TwoPIScalarNinverseNLO) = Cint(3)
pexpcode(::TwoPIScalarLambdaNLOclassical) = Cint(4)
pexpcode(p) = throw(ArgumentError("GPUfull implements LambdaLO, LambdaNLO (quantum and classical), NinverseLO and NinverseNLO, not $(typeof(p))"))

# replaces the CPUfull branch of getTwoPISimData
getTwoPISimData(problem::QFTdynamicsProblem, ::TwoPIScalarGPUfull) = gpusimdata(problem, :kb_create)
getTwoPISimData(problem::QFTdynamicsProblem, ::TwoPIScalarGPUorbits) = gpusimdata(problem, :kb_create_sym)

# Lambda: the coupling of the run (the problem's by default; the rrule in
# TwoPIScalar_GPUrrule.jl evaluates at others)
function gpusimdata(problem::QFTdynamicsProblem, create::Symbol; Lambda::Float64 = problem.model.Lambda)
    d, m = problem.disc, problem.model
    p = Ref(KBParams(d.Nx, d.sdim, problem.simsetup.NstepsinMemory, m.ONgroup, d.dt, Lambda, d.Mass,
                     m.Mass2, pexpcode(problem.pexp), 0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = create === :kb_create_sym ?
        ccall((:kb_create_sym, libqftkb), Cint, (Ref{KBParams}, Ref{Ptr{Cvoid}}), p, h) :
        ccall((:kb_create, libqftkb), Cint, (Ref{KBParams}, Ref{Ptr{Cvoid}}), p, h)
    rc == 0 || throw(ErrorException("kb_create failed with code $rc"))
    kL2 = zeros(d.Nx^d.sdim)
    for element in d.fftwhelper, idx in element.ind
        kL2[idx] = element.lev2
    end
    kbcheck(h[], ccall((:kb_set_kL2, libqftkb), Cint, (Ptr{Cvoid}, Ptr{Float64}), h[], kL2))
    sd = TwoPIScalarSimDataGPUfull(h[], kL2, problem.simsetup.NstepsinMemory, d.Nx, d.sdim)
    finalizer(x -> ccall((:kb_destroy, libqftkb), Cint, (Ptr{Cvoid},), x.ctx), sd)
    return sd
end

getTwoPIScalarTmpData(simdata::TwoPIScalarSimDataGPUfull, problem::QFTdynamicsProblem) = TwoPIScalarTmpDataGPUfull()

function Base.getproperty(s::TwoPIScalarSimDataGPUfull, f::Symbol)
    if f === :indices
        a, b = Ref{Cint}(0), Ref{Cint}(0)
        kbcheck(getfield(s, :ctx), ccall((:kb_window, libqftkb), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}),
                                         getfield(s, :ctx), a, b))
        return [Int(a[]), Int(b[])]
    elseif f === :hartreemass2
        m = Ref{Cdouble}(0)
        kbcheck(getfield(s, :ctx), ccall((:kb_hartree, libqftkb), Cint, (Ptr{Cvoid}, Ref{Cdouble}),
                                         getfield(s, :ctx), m))
        return [m[]]
    end
    return getfield(s, f)
end

# initSimData! for the particle-number initialisations (Initialization.jl:8-29)
function initSimData!(initmass::Float64, simdata::TwoPIScalarSimDataGPUfull, tmpdata::TwoPIScalarTmpDataGPUfull,
                      model::TwoPIScalarPhi4, pexp::TwoPIScalarPertExpansion, disc::TwoPIScalarDiscretization,
                      init::TwoPIScalarInitialization)
    n = getparticlenr(init, disc)
    nsite = zeros(disc.Nx^disc.sdim)
    for i in 1:length(disc.fftwhelper), idx in disc.fftwhelper[i].ind
        nsite[idx] = n[i]
    end
    kbcheck(simdata.ctx, ccall((:kb_init_particles, libqftkb), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble),
                               simdata.ctx, nsite, initmass))
end

# initSimData! Gauss (Initialization.jl:95-112)
function initSimData!(initmass::Float64, simdata::TwoPIScalarSimDataGPUfull, tmpdata::TwoPIScalarTmpDataGPUfull,
                      model::TwoPIScalarPhi4, pexp::TwoPIScalarPertExpansion, disc::TwoPIScalarDiscretization,
                      init::TwoPIScalarGauss)
    V = disc.Nx^disc.sdim
    F11 = fill(init.xi^2, V)
    F21 = init.xi * init.eta * disc.dt .+ F11
    F22 = (init.eta^2 + init.sig^2 / (4 * init.xi^2)) * disc.dt^2 .- F11 .+ 2 .* F21
    kbcheck(simdata.ctx, ccall((:kb_init_rows, libqftkb), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                               simdata.ctx, F11, F21, F22))
end

# initialize! (TwoPIScalar_Initialization.jl:289-325).  The reference defines
# it for TwoPIScalarTmpDataCPU only, so the GPU numerics need their own
# method.  The initial mass follows the reference's choice per model and
# initialisation (:293-322): the discretisation mass for the tachyonic model
# and for the Gauss, Quench, TopHatT1, Thermal and Box initialisations, the LO
# gap mass (getGapMass, :250-266) for every other one.
function initmassGPU(model::TwoPIScalarPhi4, pexp::TwoPIScalarPertExpansion, disc::TwoPIScalarDiscretization,
                     init::TwoPIScalarInitialization)
    if model isa Phi4Tachyonic || init isa TwoPIScalarGauss || init isa TwoPIScalarQuench ||
       init isa TwoPIScalarTopHatT1 || init isa TwoPIScalarThermal || init isa TwoPIScalarBox
        return disc.Mass
    end
    return getGapMass(model, pexp, disc, init)
end

function initialize!(thesolution::QFTdynamicsSolutionTwoPIScalar, tmpdata::TwoPIScalarTmpDataGPUfull)
    @unpack problem, simdata, measurearray = thesolution
    @unpack model, pexp, disc, init, reno, num, simsetup = problem
    if isnewsimulation(simsetup) == true
        initSimData!(initmassGPU(model, pexp, disc, init), simdata, tmpdata, model, pexp, disc, init)
    end
    simsetup.lastEvolStep = 2
    measure!(thesolution, 2)
end

function getHartreeMass2(model::TwoPIScalarPhi4, pexp::TwoPIScalarPertExpansion, disc::TwoPIScalarDiscretization,
                         simdata::TwoPIScalarSimDataGPUfull, idxone::Int64, idxtwo::Int64)
    return simdata.hartreemass2[1]
end

# checkpoint and restart (kb_save_window / kb_load_window): a snapshot of the
# run as bytes, restored into SimData built from the same parameters; evolve!
# then continues at simdata.indices[2] + 1
function saveWindow(simdata::TwoPIScalarSimDataGPUfull)
    n = Ref{Culonglong}(0)
    kbcheck(simdata.ctx, ccall((:kb_window_bytes, libqftkb), Cint, (Ptr{Cvoid}, Ref{Culonglong}), simdata.ctx, n))
    buf = Vector{UInt8}(undef, n[])
    kbcheck(simdata.ctx, ccall((:kb_save_window, libqftkb), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Culonglong),
                               simdata.ctx, buf, n[]))
    return buf
end

function loadWindow!(simdata::TwoPIScalarSimDataGPUfull, buf::Vector{UInt8})
    kbcheck(simdata.ctx, ccall((:kb_load_window, libqftkb), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Culonglong),
                               simdata.ctx, buf, length(buf)))
end

function evolve!(thesolution::QFTdynamicsSolutionTwoPIScalar, tmpdata::TwoPIScalarTmpDataGPUfull, t)
    simdata = thesolution.simdata
    kbcheck(simdata.ctx, ccall((:kb_step, libqftkb), Cint, (Ptr{Cvoid}, Cint), simdata.ctx, t))
    thesolution.problem.simsetup.lastEvolStep = t
end

function createMeasurement(t::Int64, model::TwoPIScalarPhi4, simdata::TwoPIScalarSimDataGPUfull,
                           disc::TwoPIScalarDiscretization, pexp::TwoPIScalarPertExpansion)
    V = disc.Nx^disc.sdim
    ctx = simdata.ctx
    out = Ref{KBMeas}()
    F, ddF, r, n, om, c = (zeros(V) for _ in 1:6)
    kbcheck(ctx, ccall((:kb_measure, libqftkb), Cint,
                       (Ptr{Cvoid}, Cint, Ref{KBMeas}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                        Ptr{Float64}, Ptr{Float64}), ctx, t, out, F, ddF, r, n, om, c))
    L = t - simdata.indices[1]
    SigF, Sigr, Fonet, ronet = (zeros(L) for _ in 1:4)
    kbcheck(ctx, ccall((:kb_measure_history, libqftkb), Cint,
                       (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                       ctx, t, SigF, Sigr, Fonet, ronet))
    # per-level vectors at the representative site ind[1] (Measurement.jl:25)
    nl = length(disc.fftwhelper)
    meas = MeasurementTwoPIScalar{nl, disc.NstepsinMemory}()
    meas.time, meas.E, meas.E_pot, meas.E_kin, meas.E_mem, meas.M2 =
        out[].time, out[].E, out[].E_pot, out[].E_kin, out[].E_mem, out[].M2
    for i in 1:nl
        idx = disc.fftwhelper[i].ind[1]
        meas.F[i], meas.ddF[i], meas.r[i] = F[idx], ddF[idx], r[idx]
        meas.n[i], meas.omega[i], meas.c[i] = n[idx], om[idx], c[idx]
    end
    meas.SigF, meas.Sigr, meas.Fonet, meas.ronet = SigF, Sigr, Fonet, ronet
    return meas
end
