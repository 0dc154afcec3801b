# VPMX.jl -- Julia binding of libvpmx.so (include/vpmx.h) for FLOWUnsteady.
#
# FLOWUnsteady takes the particle-interaction scheme and the time integrator
# as function-valued arguments of run_simulation
# (src/FLOWUnsteady_simulation.jl:124-137), which become fields of
# vpm.ParticleField (:239-253).  Pass the functions below there:
#
#     uns.run_simulation(simulation, nsteps;
#                        vpm_UJ          = VPMX.UJ_fmm_gpu,       # or VPMX.UJ_gpu
#                        vpm_integration = VPMX.rungekutta3_gpu,  # or VPMX.euler_gpu
#                        ...)
#
# Every struct mirrors its C typedef field for field (names, types, order), so
# Julia lays it out as the C compiler does; tests/test_julia_shim.py checks
# that against include/vpmx.h, and every ccall's symbol and argument count.
module VPMX

import FLOWUnsteady: vpm

const lib = joinpath(@__DIR__, "..", "flowunsteady_amd", "libvpmx.so")

# ---- the C structs --------------------------------------------------------------------
struct Layout                 # vpmx_layout
    nrows::Int32
    X::Int32
    Gamma::Int32
    sigma::Int32
    vol::Int32
    U::Int32
    J::Int32
    SFS::Int32
    C::Int32
    static_::Int32
end

struct UJParams               # vpmx_uj_params
    kernel::Int32
    precision::Int32
    transposed::Int32
    reset::Int32
    sfs::Int32
    fmm::Int32
    p::Int32
    ncrit::Int32
    theta::Float64
    sigma_cut::Float64        # 0: geometric acceptance (nonzero_sigma = false);
                              # < 0: the kernel's default number of core radii
end

struct StepParams             # vpmx_step_params
    f::Float64
    g::Float64
    Uinf::NTuple{3,Float64}
    sfs::Int32
    Cs::Float64
    clip_backscatter::Int32
    relax_kind::Int32
    rlxf::Float64
    alpha::Float64            # dynamic SFS
    sfs_rlxf::Float64
    minC::Float64
    maxC::Float64
    viscous::Int32            # 0 inviscid, 1 core spreading, 2 PSE
    nu::Float64
    sgm0::Float64
    beta::Float64
    itmax::Int32
    tol::Float64
end

# ---- context --------------------------------------------------------------------------
const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:vpmx_create, lib), Cint, (Ref{Ptr{Cvoid}}, Int32), r, 0)
        st == 0 || error("vpmx_create: " * unsafe_string(ccall((:vpmx_last_error, lib), Cstring,
                                                                (Ptr{Cvoid},), C_NULL)))
        CTX[] = r[]
        # replay direct-scheme steps from a captured graph (bitwise equal to eager)
        check(ccall((:vpmx_set_graph, lib), Cint, (Ptr{Cvoid}, Int32), CTX[], 1))
    end
    return CTX[]
end

check(st) = st == 0 || error("vpmx: " * unsafe_string(ccall((:vpmx_last_error, lib), Cstring,
                                                              (Ptr{Cvoid},), ctx())))

# ---- what the ParticleField's settings mean for the library --------------------------
const precision = Ref{Int32}(1)     # 0 = FP64, 1 = mixed FP32 pair math / FP64 sums

# FLOWVPM row indices are 1-based; the ABI takes 0-based row offsets.
# C is three rows (coefficient, averaged numerator, averaged denominator).
layout(pf) = Layout(size(pf.particles, 1), vpm.X_INDEX[1] - 1, vpm.GAMMA_INDEX[1] - 1,
                    vpm.SIGMA_INDEX[1] - 1, vpm.VOL_INDEX[1] - 1, vpm.U_INDEX[1] - 1,
                    vpm.J_INDEX[1] - 1, vpm.SFS_INDEX[1] - 1, vpm.C_INDEX[1] - 1,
                    vpm.STATIC_INDEX[1] - 1)

kernel_id(k) = k === vpm.gaussianerf ? 2 : k === vpm.winckelmans ? 3 :
               k === vpm.gaussian ? 1 : 0

prop(x, name, default) = hasproperty(x, name) ? getproperty(x, name) : default

# vpm.FMM(; p, ncrit, theta, nonzero_sigma, ε_tol) (src/FLOWUnsteady_simulation.jl:43)
sigma_cut(fmm) = prop(fmm, :nonzero_sigma, false) ? -1.0 : 0.0

# Expansion order: p, or with ε_tol set the smallest q <= 12 whose leading
# truncation term theta^(q+1) is at most ε_tol (as flowunsteady_amd.vpm.FMM.order).
function fmm_order(fmm)
    tol = prop(fmm, :ε_tol, nothing)
    tol === nothing && return Int32(fmm.p)
    for q in 0:12
        fmm.theta^(q + 1) <= tol && return Int32(q)
    end
    return Int32(12)
end

uj_params(pf, fmm; reset=true, sfs=false, transposed=pf.transposed) =
    UJParams(kernel_id(pf.kernel), precision[], transposed, reset, sfs, fmm,
             fmm_order(pf.fmm), pf.fmm.ncrit, pf.fmm.theta, sigma_cut(pf.fmm))

# SFS models (src/FLOWUnsteady_simulation.jl:39,130): none, a constant
# coefficient, or the dynamic procedure (vpm.DynamicSFS(Estr, pseudo3level_positive;
# alpha, maxC, clippings), examples/rotorhover/rotorhover.jl:53-55).
sfs_kind(s) = s === vpm.SFS_none ? 0 : hasproperty(s, :alpha) ? 2 : 1
sfs_Cs(s) = Float64(prop(s, :Cs, 0.0))
function clips(s)     # backscatter clipping unless a clippings list leaves it out
    c = prop(s, :clippings, nothing)
    return Int32(c === nothing || any(x -> x === vpm.clipping_backscatter, c))
end

relax_kind(r) = r === vpm.norelaxation ? 0 :
                prop(r, :relax, nothing) === vpm.relax_correctedpedrizzetti ? 2 : 1
rlxf(r) = r === vpm.norelaxation ? 0.0 : Float64(r.rlxf)

# vpm_viscous (src/FLOWUnsteady_simulation.jl:42,229-234): nu and a core
# spreading's sgm0 are set by run_simulation before the field is built.
viscous_kind(v) = vpm.isinviscid(v) ? 0 : vpm.iscorespreading(v) ? 1 : 2

function step_params(pf; Uinf=Tuple(Float64.(pf.Uinf(pf.t))))
    F = pf.formulation
    f, g = hasproperty(F, :f) ? (Float64(F.f), Float64(F.g)) : (0.0, 0.0)
    S, R, V = pf.SFS, pf.relaxation, pf.viscous
    return StepParams(f, g, Uinf, sfs_kind(S), sfs_Cs(S), clips(S), relax_kind(R), rlxf(R),
                      Float64(prop(S, :alpha, 0.999)), Float64(prop(S, :rlxf, 0.005)),
                      Float64(prop(S, :minC, 0.0)), Float64(prop(S, :maxC, 1.0)),
                      viscous_kind(V), Float64(prop(V, :nu, 0.0)), Float64(prop(V, :sgm0, 0.0)),
                      Float64(prop(V, :beta, 1.5)), Int32(prop(V, :itmax, 20)),
                      Float64(prop(V, :tol, 1e-3)))
end

# ---- drop-ins for the ParticleField hooks ---------------------------------------------
function _uj(pfield, fmm; reset, sfs, transposed)
    P = uj_params(pfield, fmm; reset, sfs, transposed)
    GC.@preserve pfield begin
        check(ccall((:vpmx_uj, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Layout}, Ref{UJParams}),
                    ctx(), pfield.particles, vpm.get_np(pfield), layout(pfield), P))
    end
    return nothing
end

# vpm.UJ_direct: pfield.UJ(pfield; reset=true, sfs=false, ...)
UJ_gpu(pfield; reset=true, reset_sfs=false, sfs=false, transposed=pfield.transposed,
       optargs...) = _uj(pfield, 0; reset, sfs, transposed)

# vpm.UJ_fmm (FLOWUnsteady's default vpm_UJ) with the field's p, ncrit, theta
UJ_fmm_gpu(pfield; reset=true, reset_sfs=false, sfs=false, transposed=pfield.transposed,
           optargs...) = _uj(pfield, 1; reset, sfs, transposed)

fmm_of(pfield) = pfield.UJ === UJ_fmm_gpu ? 1 : 0

# vpm.rungekutta3: pfield.integration(pfield, dt; relax=false)
function rungekutta3_gpu(pfield, dt; relax=false, custom_UJ=nothing)
    S = step_params(pfield)
    P = uj_params(pfield, fmm_of(pfield); sfs=S.sfs != 0)
    GC.@preserve pfield begin
        check(ccall((:vpmx_rk3_step, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Layout}, Ref{UJParams},
                     Ref{StepParams}, Float64, Int32),
                    ctx(), pfield.particles, vpm.get_np(pfield), layout(pfield), P, S, dt, relax))
    end
    return nothing
end

# vpm.euler
function euler_gpu(pfield, dt; relax=false, custom_UJ=nothing)
    S = step_params(pfield)
    P = uj_params(pfield, fmm_of(pfield); sfs=S.sfs != 0)
    GC.@preserve pfield begin
        check(ccall((:vpmx_euler_step, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Layout}, Ref{UJParams},
                     Ref{StepParams}, Float64, Int32),
                    ctx(), pfield.particles, vpm.get_np(pfield), layout(pfield), P, S, dt, relax))
    end
    return nothing
end

# vpm.monitor_enstrophy on the field as it stands (its J must be current)
function enstrophy_gpu(pfield)
    e, en = Ref{Float64}(0.0), Ref{Float64}(0.0)
    GC.@preserve pfield begin
        check(ccall((:vpmx_upload, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Layout}),
                    ctx(), pfield.particles, vpm.get_np(pfield), layout(pfield)))
    end
    check(ccall((:vpmx_diagnostics, lib), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}),
                ctx(), e, en))
    return e[]
end

# ---- probes: FLOWUnsteady's Vvpm_on_Xs (src/FLOWUnsteady_simulation.jl:494-570) -------
# Same arguments and the same steps as the reference -- static particles from
# static_particles_fun and mirror images are appended to the field and removed
# again afterwards, and fsgm's scaling is applied and undone exactly as there
# -- but the probes are not appended: vpmx_uj_targets evaluates U at them from
# the field (the probes carry no circulation, so they add nothing to each
# other).  The evaluation is direct (exact), also when the field's UJ is the
# FMM.  The reference's own function keeps working unchanged with UJ_gpu /
# UJ_fmm_gpu as pfield.UJ; this one is for callers that want the m x N
# evaluation, e.g. bound to FLOWUnsteady's name with
#     @eval uns Vvpm_on_Xs(pf::vpm.ParticleField, Xs::Array{T,1}; kw...) where {T} =
#         $(VPMX).Vvpm_on_Xs(pf, Xs; kw...)
function Vvpm_on_Xs(pfield::vpm.ParticleField, Xs::Array{T,1};
                    static_particles_fun=(args...) -> nothing, dt=0, fsgm=1,
                    mirror=false, mirror_X=nothing, mirror_normal=nothing) where {T}
    m = length(Xs)
    m == 0 && return [zeros(3) for i in 1:m]
    org_np = vpm.get_np(pfield)
    # as the reference: one linear entry of the matrix per static particle present
    nscaled = 0
    if abs(fsgm) != 1
        for i in 1:pfield.np
            if pfield.particles[vpm.STATIC_INDEX, i] > 0
                pfield.particles[vpm.SIGMA_INDEX] *= fsgm
                nscaled += 1
            end
        end
    end
    static_particles_fun(pfield, pfield.t, dt)
    if mirror
        for i in 1:vpm.get_np(pfield)
            P = vpm.get_particle(pfield, i)
            X = collect(vpm.get_X(P))
            G = collect(vpm.get_Gamma(P))
            Xm = X - 2 * sum((X - mirror_X) .* mirror_normal) * mirror_normal
            Gm = 2 * sum(G .* mirror_normal) * G / sqrt(sum(G .^ 2)) - G
            C1, C2, C3 = vpm.get_C(P)
            vpm.add_particle(pfield, Xm, Gm, vpm.get_sigma(P)[]; vol=vpm.get_vol(P)[],
                             circulation=vpm.get_circulation(P)[], C=[C1, C2, C3], static=true)
        end
    end
    Xm = Matrix{Float64}(undef, 3, m)
    for (i, X) in enumerate(Xs)
        Xm[:, i] .= X
    end
    U = zeros(3, m)
    J = zeros(9, m)
    GC.@preserve pfield Xm U J begin
        check(ccall((:vpmx_upload, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Layout}),
                    ctx(), pfield.particles, vpm.get_np(pfield), layout(pfield)))
        check(ccall((:vpmx_uj_targets, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{UJParams}),
                    ctx(), Xm, m, U, J, uj_params(pfield, 0)))
    end
    for p_i in vpm.get_np(pfield):-1:(org_np + 1)
        vpm.remove_particle(pfield, p_i)
    end
    for k in 1:nscaled
        pfield.particles[vpm.SIGMA_INDEX] /= fsgm
    end
    return [U[:, i] for i in 1:m]
end

# ---- a sweep of independent cases advanced together (one launch for all) ---------------
function rungekutta3_gpu_batch(pfields::Vector, dt; relax=false)
    a = pfields[1]
    S = step_params(a; Uinf=(0.0, 0.0, 0.0))
    Uinf = reduce(vcat, [collect(Float64.(pf.Uinf(pf.t))) for pf in pfields])   # 3 per field
    ptrs = [pointer(pf.particles) for pf in pfields]
    nps = Int64[vpm.get_np(pf) for pf in pfields]
    GC.@preserve pfields ptrs nps Uinf begin
        check(ccall((:vpmx_batch_rk3_step, lib), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Ptr{Float64}}, Ptr{Int64}, Ref{Layout},
                     Ref{UJParams}, Ref{StepParams}, Ptr{Float64}, Float64, Int32),
                    ctx(), length(pfields), ptrs, nps, layout(a), uj_params(a, 0), S, Uinf, dt,
                    relax))
    end
    return nothing
end

end # module
