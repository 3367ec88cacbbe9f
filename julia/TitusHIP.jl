"""
    TitusHIP

Julia binding of `libtitus_hip.so` (the C-ABI declared in `include/titus_hip.h`)
for TITUS.  It keeps the reference's solver object and entry points: a TITUS
driver builds `obj = SolverCSD(settings)` as before (code/SolverCSD.jl:63-146)
and calls `TitusHIP.SolveFirstCollisionSourceDLR(obj)` (or the BUG / parallel
forms) in place of the CPU function of the same name.  The return tuples, the
mutation of `obj` (dose accumulation, `settings.dE`, the reduced ordinate
sets) and the cell numbering are those of the reference functions.

`titus_problem(obj, integrator)` is the operand builder: it evaluates the
reference's own `PsiBeam`, `SigmaAtEnergy`, `SetupIC` and system matrices on
`obj` and fills every field of `titus_problem_v1`.  All arrays cross the ABI as
Float64 in Julia's column-major layout, without copies in the library's
direction (the library copies them at `titus_create`).

Julia is not part of this repository's test environment: the file is
checked against the header by `tests/test_julia_binding.py` (struct field
order and types, every `ccall`'s symbol, argument count and argument types).
"""
module TitusHIP

using LinearAlgebra

const libtitus = get(ENV, "TITUS_HIP_LIB",
                     joinpath(@__DIR__, "..", "titus_amd", "libtitus_hip.so"))

const TITUS_ABI_VERSION = Int32(3)
const TITUS_INTEGRATOR_FIXED = Int32(0)
const TITUS_INTEGRATOR_BUG = Int32(1)
const TITUS_INTEGRATOR_PARALLEL = Int32(2)
const TITUS_INTEGRATOR_RK4 = Int32(3)
const TITUS_PRECISION_FP64 = Int32(0)
const TITUS_PRECISION_MIXED = Int32(1)
const TITUS_PRECISION_MIXED_FULL_FP32 = Int32(2)

# mirror of `titus_problem_v1` (include/titus_hip.h); an isbits struct has
# the C layout, so `Ref(p)` passes it to ccall as `const titus_problem_v1*`
struct TitusProblemV1
    version::Int32
    nx::Int32
    ny::Int32
    dx::Float64
    dy::Float64
    density::Ptr{Float64}
    nm::Int32
    degree::Ptr{Int32}
    Ax::Ptr{Float64}
    Az::Ptr{Float64}
    absAx::Ptr{Float64}
    absAz::Ptr{Float64}
    nq::Int32
    omega::Ptr{Float64}
    M::Ptr{Float64}
    ne::Int32
    dE::Float64
    stopping::Ptr{Float64}
    nsigma::Int32
    sigma::Ptr{Float64}
    nb::Int32
    bcells::Ptr{Int32}
    bc_base::Ptr{Float64}
    bc_scale::Ptr{Float64}
    nrestore::Int32
    restore_cells::Ptr{Int32}
    ncand::Int32
    cand_cells::Ptr{Int32}
    integrator::Int32
    precision::Int32
    r::Int32
    rank_cap::Int32
    theta::Float64
    theta_index::Float64
    rank_tol::Float64
    X0::Ptr{Float64}
    S0::Ptr{Float64}
    W0::Ptr{Float64}
    nz::Int32
    dz::Float64
    Ay::Ptr{Float64}
    absAy::Ptr{Float64}
    psi0::Ptr{Float64}
    flux_order::Int32
end

# mirror of `titus_beam_v1`: the RK4 variant's analytic pencil beam
struct TitusBeamV1
    x0::Float64
    y0::Float64
    Omega1::Float64
    Omega3::Float64
    xmid::Ptr{Float64}
    ymid::Ptr{Float64}
    tau::Ptr{Float64}
end

struct TitusError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::TitusError) = print(io, "TitusError(", e.code, "): ", e.msg)

function check(code::Integer)
    if code != 0
        msg = unsafe_string(ccall((:titus_last_error, libtitus), Cstring, ()))
        throw(TitusError(code, msg))
    end
    return nothing
end

# GlobalIndex(l, k) = l^2 + k + l (PNSystem.jl:124-128), 1-based here
global_index1(l, k) = l^2 + k + l + 1

"""
    moment_subset(N) -> (index, degree)

The k >= 0 real spherical harmonics of degree <= N in the reference's order
(l ascending, then k ascending) as 1-based positions in the full (N+1)^2 set,
and their degrees.  In 2D the k >= 0 and k < 0 moments do not couple through
Ax and Az, and a beam symmetric in y keeps the k < 0 half at zero.
"""
function moment_subset(N::Integer)
    idx = Int[]
    deg = Int32[]
    for l in 0:N, k in 0:l
        push!(idx, global_index1(l, k))
        push!(deg, Int32(l))
    end
    return idx, deg
end

# Roe matrix V |Lambda| V^-1 of a symmetric A in Float64 with the reference's
# 1e-10 drop (SolverCSD.jl:83-97)
function roe(A::AbstractMatrix)
    F = eigen(Symmetric(Matrix{Float64}(A)))
    R = F.vectors * Diagonal(abs.(F.values)) * F.vectors'
    R = 0.5 .* (R .+ R')
    R[abs.(R) .<= 1e-10] .= 0.0
    return R
end

"""
    titus_problem(obj, integrator; precision, rank_tol, flux_order)
        -> (problem::TitusProblemV1, keep)

Operands of the library's energy loop for the reference solver object `obj`
(the `SolverCSD` constructor plus the set-up part of the `Solve*` functions,
SolverCSD.jl:63-146, 904-961).  `keep` holds every array the struct points
into; it must stay reachable until `titus_create` returns (`GC.@preserve`).
Side effects on `obj` are those of the reference's set-up: `qReduced`,
`MReduced`, `OReduced` and `settings.dE`.  `density` (C values, row-major cells as `densityVec`)
replaces the object's density, as the plan objective's scaled phantom does.
"""
function titus_problem(obj, integrator::Integer; precision::Integer=TITUS_PRECISION_FP64,
                       rank_tol::Real=1e-7, flux_order::Integer=1,
                       density::Union{Nothing,AbstractVector}=nothing)
    s = obj.settings
    nx, ny = s.NCellsX, s.NCellsY
    C = nx * ny
    Nfull = obj.pn.nTotalEntries
    N = s.nPN

    # ordinate reduction: keep the directions the beam reaches (:913-925)
    floorPsi = 1e-17
    nqf = obj.Q.nquadpoints
    T = eltype(obj.Q.pointsxyz)
    psiBeam = [Float64(PsiBeam(obj, T.(obj.Q.pointsxyz[k, :]), T(s.eMax), s.x0, s.y0, 1))
               for k in 1:nqf]
    idxBeam = findall(psiBeam .> floorPsi * maximum(psiBeam))
    obj.qReduced = obj.Q.pointsxyz[idxBeam, :]
    obj.MReduced = obj.M[:, idxBeam]
    obj.OReduced = obj.O[idxBeam, :]
    nq = length(idxBeam)

    # moments: the k >= 0 subset in 2D
    sub, degree = moment_subset(N)
    nm = length(sub)
    Ax, _, Az = SetupSystemMatrices(obj.pn)
    Axs = Matrix{Float64}(Ax[sub, sub])
    Azs = Matrix{Float64}(Az[sub, sub])
    absAx, absAz = roe(Axs), roe(Azs)
    omega = Matrix{Float64}(obj.qReduced)                      # nq x 3
    M = Matrix{Float64}(obj.M[sub, idxBeam])                   # nm x nq

    # energy grid and the scattering table sigma_l(eGrid[n]) (CSD.jl:97-127)
    eTrafo = Float64.(obj.csd.eTrafo)
    energy = obj.csd.eGrid
    ne = length(eTrafo)
    dE = eTrafo[2] - eTrafo[1]
    s.dE = dE
    stopping = Float64.(obj.csd.S)
    sigma = Matrix{Float64}(undef, N + 1, ne)                 # column n = row n of the C table
    for n in 1:ne
        sigma[:, n] .= Float64.(SigmaAtEnergy(obj.csd, energy[n]))
    end

    # ghost cells and the beam source, split as bc_base[b, q] * bc_scale[n]:
    # PsiBeam factors into angle x space x f(E, n); f is read off PsiBeam at
    # the beam's own direction and position (where it is largest)
    bidx = unique(obj.boundaryIdx)                             # 1-based, (i-1) ny + j
    nb = length(bidx)
    bcells = Int32.(bidx .- 1)
    Om0 = T.([s.Omega1, 0.0, s.Omega3])
    f(E, n) = Float64(PsiBeam(obj, Om0, T(E), s.x0, s.y0, n))
    f0 = f(s.eMax, 1)
    f0 > 0 || throw(ArgumentError("PsiBeam vanishes at the beam's own direction and position"))
    bc_base = Matrix{Float64}(undef, nb, nq)
    for q in 1:nq, b in 1:nb
        i = (bidx[b] - 1) ÷ ny + 1
        j = (bidx[b] - 1) % ny + 1
        bc_base[b, q] = Float64(PsiBeam(obj, T.(obj.qReduced[q, :]), T(s.eMax),
                                        s.xMid[i], s.yMid[j], 1))
    end
    bc_scale = [f(energy[n], n) / f0 for n in 1:ne]

    # psi[obj.boundaryIdx] linear-indexes the (nx, ny, nq) array with
    # (i-1) ny + j: ordinate 1 at (i', j') = ((c-1) % nx + 1, (c-1) ÷ nx + 1)
    # keeps its value through the sweep (:1002, 1005)
    restore = Int32[]
    for c in obj.boundaryIdx
        ip, jp = (c - 1) % nx, (c - 1) ÷ nx
        push!(restore, Int32(ip * ny + jp))
    end
    restore = sort(unique(restore))
    ghost = falses(C)
    ghost[bidx] .= true
    cand = Int32.(findall(.!ghost) .- 1)                       # interior cells, 0-based

    # densityVec (row-major cells, :143), or a caller's density on the same cells
    density = Float64.(density === nothing ? obj.densityVec : density)
    length(density) == C || throw(DimensionMismatch("density must have one value per cell"))
    adaptive = integrator != TITUS_INTEGRATOR_FIXED
    r0 = adaptive ? s.r ÷ 2 : s.r                              # :2013, :1728
    psi0 = s.problem == "waterBeam" ?
        Float64.(Mat2Vec(SetupIC(obj, obj.qReduced))) : nothing # :916, C x nq

    keep = (density, degree, Axs, Azs, absAx, absAz, omega, M, stopping, sigma, bcells,
            bc_base, bc_scale, restore, cand, psi0)
    ptr(a) = a === nothing ? Ptr{Float64}(C_NULL) : pointer(a)
    p = TitusProblemV1(TITUS_ABI_VERSION, Int32(nx), Int32(ny), Float64(s.dx), Float64(s.dy),
                       pointer(density), Int32(nm), pointer(degree), pointer(Axs), pointer(Azs),
                       pointer(absAx), pointer(absAz), Int32(nq), pointer(omega), pointer(M),
                       Int32(ne), dE, pointer(stopping), Int32(N + 1), pointer(sigma),
                       Int32(nb), pointer(bcells), pointer(bc_base), pointer(bc_scale),
                       Int32(length(restore)), pointer(restore), Int32(length(cand)),
                       pointer(cand), Int32(integrator), Int32(precision), Int32(r0),
                       Int32(s.r), Float64(s.ϑ), Float64(s.ϑIndex), Float64(rank_tol),
                       Ptr{Float64}(C_NULL), Ptr{Float64}(C_NULL), Ptr{Float64}(C_NULL),
                       Int32(1), 1.0, Ptr{Float64}(C_NULL), Ptr{Float64}(C_NULL), ptr(psi0),
                       Int32(flux_order))
    return p, keep, sub
end

"""A device-resident solve (one `titus_ctx`), destroyed by the finalizer."""
mutable struct TitusContext
    ptr::Ptr{Cvoid}
    C::Int
    nm::Int
    nq::Int
    ne::Int
    sub::Vector{Int}
end

function destroy!(ctx::TitusContext)
    if ctx.ptr != C_NULL
        ccall((:titus_destroy, libtitus), Cvoid, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    return nothing
end

function create(obj, integrator::Integer; device::Integer=0, kwargs...)
    p, keep, sub = titus_problem(obj, integrator; kwargs...)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        check(ccall((:titus_create, libtitus), Cint,
                    (Ptr{Ptr{Cvoid}}, Ref{TitusProblemV1}, Int32), out, Ref(p), Int32(device)))
    end
    ctx = TitusContext(out[], Int(p.nx) * Int(p.ny), Int(p.nm), Int(p.nq), Int(p.ne), sub)
    finalizer(destroy!, ctx)
    return ctx
end

function cfl(ctx::TitusContext)
    v = Ref{Float64}(0.0)
    check(ccall((:titus_get_cfl, libtitus), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, v))
    return v[]
end

solve!(ctx::TitusContext, last_step::Integer=0) =
    check(ccall((:titus_solve, libtitus), Cint, (Ptr{Cvoid}, Int32), ctx.ptr, Int32(last_step)))

function progress(ctx::TitusContext)
    done, rank = Ref{Int32}(0), Ref{Int32}(0)
    check(ccall((:titus_progress, libtitus), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}),
                ctx.ptr, done, rank))
    return Int(done[]), Int(rank[])
end

"""(dose, X, S, W, psi) with psi in the library's C x nq (`Mat2Vec`) layout."""
function get_state(ctx::TitusContext)
    _, r = progress(ctx)
    dose = zeros(ctx.C)
    X = zeros(ctx.C, r)
    S = zeros(r, r)
    W = zeros(ctx.nm, r)
    psi = zeros(ctx.C, ctx.nq)
    check(ccall((:titus_get_state, libtitus), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.ptr, dose, X, S, W, psi))
    return dose, X, S, W, psi
end

function history(ctx::TitusContext, n::Integer)
    ranks = zeros(Int32, n)
    etas = zeros(n)
    check(ccall((:titus_get_history, libtitus), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Float64}, Int32),
                ctx.ptr, ranks, etas, Int32(n)))
    return ranks, etas
end

# outputs of the reference's Solve* functions (:1116-1118): X U, Sigma / sqrt 2,
# O W V on all (N+1)^2 moments, W V, the dose and psi as (nx, ny, nq)
function reference_outputs(obj, ctx::TitusContext)
    dose, X, S, W, psi = get_state(ctx)
    U, Sig, V = svd(S)
    Wfull = zeros(obj.pn.nTotalEntries, size(W, 2))
    Wfull[ctx.sub, :] .= W
    obj.dose .+= dose                                          # Solve* accumulates into obj.dose
    nx, ny = obj.settings.NCellsX, obj.settings.NCellsY
    psi3 = zeros(nx, ny, ctx.nq)                               # Vec2Mat per ordinate (utils.jl:57-64)
    for q in 1:ctx.nq, i in 1:nx, j in 1:ny
        psi3[i, j, q] = psi[(i - 1) * ny + j, q]
    end
    return X * U, 0.5 * sqrt(obj.gamma[1]) * Sig, obj.O * Wfull * V, Wfull * V, dose, psi3
end

"""SolveFirstCollisionSourceDLR(obj) on the GPU (SolverCSD.jl:891-1120)."""
function SolveFirstCollisionSourceDLR(obj; precision=TITUS_PRECISION_FP64)
    ctx = create(obj, TITUS_INTEGRATOR_FIXED; precision=precision)
    try
        solve!(ctx)
        return reference_outputs(obj, ctx)
    finally
        destroy!(ctx)
    end
end

"""CSolveFirstCollisionSourceDLR(obj): the fixed-rank solve with the limited
flux `solveFlux!` (SolverCSD.jl:1122-1360)."""
function CSolveFirstCollisionSourceDLR(obj; precision=TITUS_PRECISION_FP64)
    ctx = create(obj, TITUS_INTEGRATOR_FIXED; precision=precision, flux_order=2)
    try
        solve!(ctx)
        return reference_outputs(obj, ctx)
    finally
        destroy!(ctx)
    end
end

"""Turn the step rejection of the rank-adaptive integrator on (c_eta > 0) or
off (SolverCSD.jl:2204-2233, 2282-2301), before the first step."""
set_rejection!(ctx::TitusContext, c_eta::Real) =
    check(ccall((:titus_set_rejection, libtitus), Cint, (Ptr{Cvoid}, Float64), ctx.ptr, Float64(c_eta)))

function rejections(ctx::TitusContext)
    n = Ref{Int64}(0)
    check(ccall((:titus_get_rejections, libtitus), Cint, (Ptr{Cvoid}, Ptr{Int64}), ctx.ptr, n))
    return Int(n[])
end

"""CudaSolveFirstCollisionSourceDLR4thOrder(obj) on the MI355X (SolverCSD.jl:1422-1724):
fixed rank, RK4 on the central streaming terms, the analytic uncollided flux
of the pencil beam at (x0, y0) along (Omega1, Omega3).  Returns the outputs of
the fixed-rank form with psi as C x nq, as the reference returns Matrix(psi)."""
function CudaSolveFirstCollisionSourceDLR4thOrder(obj; precision=TITUS_PRECISION_FP64)
    ctx = create(obj, TITUS_INTEGRATOR_RK4; precision=precision)
    try
        s = obj.settings
        xmid, ymid = Float64.(s.xMid), Float64.(s.yMid)
        tau = Float64.(obj.csd.eTrafo)
        GC.@preserve xmid ymid tau begin
            b = TitusBeamV1(Float64(s.x0), Float64(s.y0), Float64(s.Omega1), Float64(s.Omega3),
                            pointer(xmid), pointer(ymid), pointer(tau))
            check(ccall((:titus_set_analytic_beam, libtitus), Cint,
                        (Ptr{Cvoid}, Ref{TitusBeamV1}), ctx.ptr, Ref(b)))
        end
        solve!(ctx)
        out = reference_outputs(obj, ctx)
        _, _, _, _, psi = get_state(ctx)
        return out[1:5]..., psi
    finally
        destroy!(ctx)
    end
end

"""SolveFirstCollisionSourceDLRBUGRejection(obj) on the GPU (SolverCSD.jl:2011-2374):
the outputs of the fixed-rank form plus [E rank] and [E eta].  `rejection=true`
turns on the step rejection with `settings.cη` (a rejected step restarts
from its own start at the enlarged rank, so nothing is counted twice)."""
function SolveFirstCollisionSourceDLRBUGRejection(obj, recordGif::Bool=false;
                                                   precision=TITUS_PRECISION_FP64,
                                                   rejection::Bool=false)
    ctx = create(obj, TITUS_INTEGRATOR_BUG; precision=precision)
    try
        rejection && set_rejection!(ctx, obj.settings.cη)
        solve!(ctx)
        done, _ = progress(ctx)
        ranks, etas = history(ctx, done)
        E = obj.csd.eGrid[2:done + 1]
        return reference_outputs(obj, ctx)..., [E Float64.(ranks)]', [E etas]
    finally
        destroy!(ctx)
    end
end

"""SolveFirstCollisionSourceDLRParallel(obj) on the GPU (SolverCSD.jl:1726-2008):
the outputs of the fixed-rank form plus rVec (energies and ranks after each step)."""
function SolveFirstCollisionSourceDLRParallel(obj, recordGif::Bool=false;
                                               precision=TITUS_PRECISION_FP64)
    ctx = create(obj, TITUS_INTEGRATOR_PARALLEL; precision=precision)
    try
        solve!(ctx)
        done, _ = progress(ctx)
        ranks, _ = history(ctx, done)
        rVec = [obj.csd.eGrid[2:done + 1]'; Float64.(ranks)']
        return reference_outputs(obj, ctx)..., rVec
    finally
        destroy!(ctx)
    end
end

# ---- batches and gradients -------------------------------------------------
"""Solve many contexts (e.g. the beamlets of a plan) side by side."""
function solve_batch!(ctxs::Vector{TitusContext}; nthreads::Integer=0)
    ptrs = [c.ptr for c in ctxs]
    status = zeros(Int32, length(ctxs))
    check(ccall((:titus_solve_batch, libtitus), Cint, (Ptr{Ptr{Cvoid}}, Int32, Int32, Ptr{Int32}),
                ptrs, Int32(length(ctxs)), Int32(nthreads), status))
    return status
end

"""Solve many 2D contexts with shared launches: every kernel of the energy
step runs once for all contexts of a group (`groups` groups of about
`width / groups` contexts; `width` <= 0: all at once).  Returns the status
codes, the batch's counters (flushes, launches, recorded launches,
nanoseconds in flushes, of them waiting for the streams, and the fewest
launches the flushes could take) and each group's tuned tiles (4 values per
group)."""
function solve_batch_lockstep!(ctxs::Vector{TitusContext}; width::Integer=0, groups::Integer=4)
    ptrs = [c.ptr for c in ctxs]
    status = zeros(Int32, length(ctxs))
    stats = zeros(Int64, 6)
    tiles = zeros(Int32, 4 * max(1, groups))
    check(ccall((:titus_solve_batch_lockstep, libtitus), Cint,
                (Ptr{Ptr{Cvoid}}, Int32, Int32, Int32, Ptr{Int32}, Ptr{Int64}, Ptr{Int32}),
                ptrs, Int32(length(ctxs)), Int32(width), Int32(groups), status, stats, tiles))
    return status, stats, tiles
end

set_checkpoints!(ctx::TitusContext, every::Integer) =
    check(ccall((:titus_set_checkpoints, libtitus), Cint, (Ptr{Cvoid}, Int32), ctx.ptr, Int32(every)))

"""d <dose_bar, dose> / d density for each context (C values each): with
`lockstep`, the recomputations and reverse sweeps of all contexts share their
launches (in `groups` groups); otherwise one stream per context on up to
`nthreads` host threads.  Both give the gradients of a lone adjoint bit for
bit."""
function adjoint_batch(ctxs::Vector{TitusContext}, dose_bars::Vector{Vector{Float64}};
                       nthreads::Integer=0, lockstep::Bool=false, groups::Integer=8)
    grads = [zeros(c.C) for c in ctxs]
    ptrs = [c.ptr for c in ctxs]
    status = zeros(Int32, length(ctxs))
    GC.@preserve dose_bars grads begin
        bars = [pointer(b) for b in dose_bars]
        gps = [pointer(g) for g in grads]
        if lockstep
            check(ccall((:titus_adjoint_batch_lockstep, libtitus), Cint,
                        (Ptr{Ptr{Cvoid}}, Int32, Int32, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}},
                         Ptr{Int32}),
                        ptrs, Int32(length(ctxs)), Int32(0), Int32(groups), bars, gps, status))
        else
            check(ccall((:titus_adjoint_batch, libtitus), Cint,
                        (Ptr{Ptr{Cvoid}}, Int32, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Int32}),
                        ptrs, Int32(length(ctxs)), Int32(nthreads), bars, gps, status))
        end
    end
    return grads
end

# ---- the rrule of a plan objective ------------------------------------------
"""
    plan_objective(objs, w, s, target; every=16, nthreads=0) -> J

J(w, s) = || sum_b w_b D_b(rho_0 .* s) - target ||^2 over the beamlet weights w
and the voxel density scalings s, where D_b is the fixed-rank DLRA dose of
beamlet solver `objs[b]` (SolveFirstCollisionSourceDLR) and rho_0 the density
of `objs[1]` (the beamlets share the phantom).  Differentiable through the
`rrule` below (ChainRulesCore 1.15, in the reference's Manifest).  The
forward solves and the adjoints run with shared launches (8 groups); with
`nthreads > 0` they run one stream per beamlet on that many host threads.
"""
function plan_objective(objs, w::AbstractVector, s::AbstractVector, target::AbstractVector;
                        every::Integer=16, nthreads::Integer=0)
    J, pb = plan_forward(objs, w, s, target; every=every, nthreads=nthreads)
    pb.release()
    return J
end

function plan_forward(objs, w, s, target; every::Integer=16, nthreads::Integer=0)
    rho0 = Float64.(objs[1].densityVec)
    rho = rho0 .* Float64.(s)
    ctxs = TitusContext[]
    try
        for obj in objs
            push!(ctxs, create(obj, TITUS_INTEGRATOR_FIXED; density=rho))
        end
        foreach(c -> set_checkpoints!(c, every), ctxs)
        if nthreads > 0
            solve_batch!(ctxs; nthreads=nthreads)
        else
            solve_batch_lockstep!(ctxs; groups=8)
        end
    catch
        foreach(destroy!, ctxs)
        rethrow()
    end
    D = reduce(hcat, [get_state(c)[1] for c in ctxs])          # C x B
    resid = D * Float64.(w) .- Float64.(target)
    J = dot(resid, resid)
    function pullback(Jbar)
        try
            wbar = Jbar .* 2 .* (D' * resid)
            bars = [Jbar .* 2 .* Float64(w[b]) .* resid for b in eachindex(ctxs)]
            grads = adjoint_batch(ctxs, bars; nthreads=nthreads, lockstep=nthreads <= 0)
            sbar = rho0 .* reduce(+, grads)
            return wbar, sbar
        finally
            foreach(destroy!, ctxs)
        end
    end
    release() = foreach(destroy!, ctxs)
    return J, (pullback=pullback, release=release)
end

import ChainRulesCore
function ChainRulesCore.rrule(::typeof(plan_objective), objs, w, s, target;
                              every::Integer=16, nthreads::Integer=0)
    J, pb = plan_forward(objs, w, s, target; every=every, nthreads=nthreads)
    function plan_objective_pullback(Jbar)
        wbar, sbar = pb.pullback(ChainRulesCore.unthunk(Jbar))
        return ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent(), wbar, sbar,
               ChainRulesCore.NoTangent()
    end
    return J, plan_objective_pullback
end

end # module
