# QBHip: binds QuantumBilliards.jl's Kress/Beyn hot path to libqbhip (HIP, gfx950)
# through its C ABI (include/qb.h).  Include it after QuantumBilliards:
#
#     using QuantumBilliards
#     include("QBHip.jl"); using .QBHip
#     QBHip.enable!()          # route construct_boundary_matrices!, construct_B_matrix and
#                              # residual_and_norm_select here
#
# Not executed in this repository (there is no Julia toolchain here); the C ABI
# it calls is exercised by the Python tests.
module QBHip

using LinearAlgebra
import QuantumBilliards
const QB = QuantumBilliards

const LIB = get(ENV, "QBHIP_LIB", joinpath(@__DIR__, "..", "libqbhip.so"))

const QB_CFIE_KRESS = Int32(1)
const QB_DLP_KRESS = Int32(2)
const QB_FP64 = Int32(0)
const QB_MIXED = Int32(1)
const QB_MIXED_INT8RES = Int32(2)

struct ChebParams
    n_panels_h::Int32
    M_h::Int32
    n_panels_j::Int32
    M_j::Int32
end

struct Stats
    N::Int64
    nq::Int32
    r::Int32
    precision::Int32
    refine_iters_max::Int32
    nodes_fp64_fallback::Int32
    refine_resid_max::Float64
    t_assemble_ms::Float64
    t_factor_ms::Float64
    t_solve_ms::Float64
    rmin::Float64
    rmax::Float64
end
Stats() = Stats(0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)

const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function ctx()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:qb_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        rc == 0 || error("qb_ctx_create failed ($rc)")
        CTX[] = h[]
    end
    return CTX[]
end

# nonzero status -> the exception type the reference raises in the same case
function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:qb_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx()))
    rc == -1 ? throw(AssertionError(msg)) : error("libqbhip ($rc): $msg")
end

opcode(::Union{QB.CFIE_kress,QB.CFIE_kress_corners,QB.CFIE_kress_global_corners}) = QB_CFIE_KRESS
opcode(::Union{QB.DLP_kress,QB.DLP_kress_global_corners}) = QB_DLP_KRESS

# Upload a Vector{BoundaryPointsCFIE} (holes already reversed by evaluate_points).
function geom(pts::Vector{QB.BoundaryPointsCFIE{Float64}})
    n = Int64[length(p.xy) for p in pts]
    cat(f) = reduce(vcat, [f(p) for p in pts])
    x = cat(p -> getindex.(p.xy, 1));         y = cat(p -> getindex.(p.xy, 2))
    dx = cat(p -> getindex.(p.tangent, 1));   dy = cat(p -> getindex.(p.tangent, 2))
    ddx = cat(p -> getindex.(p.tangent_2, 1)); ddy = cat(p -> getindex.(p.tangent_2, 2))
    ts = cat(p -> p.ts);                      ws = cat(p -> p.ws)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:qb_geom_create, LIB), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ref{Ptr{Cvoid}}),
                ctx(), Int32(length(pts)), n, x, y, dx, dy, ddx, ddy, ts, ws, h))
    return h[]
end
geom(p::QB.BoundaryPointsCFIE{Float64}) = geom([p])
geom_destroy(g) = ccall((:qb_geom_destroy, LIB), Cvoid, (Ptr{Cvoid},), g)

# Orbit table of the solver's reduced workspace, from the reference's own
# orbit builders: CFIE_kress visits the nodes in order with the coordinate maps
# of build_symmetry_maps (the loop of build_CFIE_kress_reduced_workspace,
# cheb-CFIE:482-502); DLP_kress takes symmetry_index_orbits (dlp_kress.jl:703).
function orbit_table(solver, pts)
    sym = solver.symmetry
    if opcode(solver) == QB_DLP_KRESS
        p = pts isa Vector ? only(pts) : pts
        _, _, _, f2full, f2scale = QB.symmetry_index_orbits(Float64, p, sym, solver.billiard)
    else
        pv = pts isa Vector ? pts : [pts]
        maps = QB.build_symmetry_maps(pv, sym)
        Nfull = sum(length(p.xy) for p in pv)
        visited = falses(Nfull)
        f2full = Vector{Vector{Int}}()
        f2scale = Vector{Vector{ComplexF64}}()
        for g in 1:Nfull
            visited[g] && continue
            orb = QB.apply_symmetry_orbit_from_maps(maps, sym, g)
            push!(f2full, orb.indices)
            push!(f2scale, ComplexF64.(orb.scales))
            visited[orb.indices] .= true
        end
    end
    G = length(f2full[1])
    img = Int64[f2full[b][l] - 1 for l in 1:G, b in eachindex(f2full)]   # img[l, b], 0-based
    sc = ComplexF64[f2scale[b][l] for l in 1:G, b in eachindex(f2full)]
    return img, sc, length(f2full), G
end

# geometry of a solver: with a symmetry, the orbits are attached and every
# assembly on it is the reduced one (qb_geom_set_orbits)
function geom(solver, pts)
    g = geom(pts)
    if !isnothing(solver.symmetry)
        img, sc, m, G = orbit_table(solver, pts)
        rc = ccall((:qb_geom_set_orbits, LIB), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int32, Ptr{Int64}, Ptr{ComplexF64}),
                   ctx(), g, Int64(m), Int32(G), img, sc)
        rc == 0 || (geom_destroy(g); check(rc))
    end
    return g
end
geom_dim(g) = Int(ccall((:qb_geom_dim, LIB), Int64, (Ptr{Cvoid},), g))

# construct_boundary_matrices! (cheb-CFIE:1431-1450, cheb-DLP:1110-1124)
function construct_boundary_matrices!(Tbufs::Vector{Matrix{ComplexF64}}, solver, pts,
                                      zj::AbstractVector{ComplexF64}; multithreaded::Bool=true,
                                      use_chebyshev::Bool=true, n_panels_h::Int=15000,
                                      M_h::Int=5, n_panels_j::Int=3000, M_j::Int=5,
                                      timeit::Bool=false)
    use_chebyshev || error("Direct complex-k construction is not implemented.")
    Mk = length(zj)
    @assert length(Tbufs) == Mk
    g = geom(solver, pts)
    try
        N = geom_dim(g)
        for T in Tbufs
            @assert size(T) == (N, N)
        end
        out = Array{ComplexF64}(undef, N, N, Mk)      # one contiguous block
        cp = Ref(ChebParams(n_panels_h, M_h, n_panels_j, M_j))
        check(ccall((:qb_assemble, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Int64, Ref{ChebParams},
                     Ptr{ComplexF64}, Int32, Ptr{Float64}),
                    ctx(), g, opcode(solver), collect(zj), Int64(Mk), cp, out, Int32(0), C_NULL))
        for m in 1:Mk
            copyto!(Tbufs[m], view(out, :, :, m))
        end
    finally
        geom_destroy(g)
    end
    return nothing
end

# construct_boundary_matrices_with_derivatives! (cheb-CFIE:1507-1530,
# cheb-DLP:1126-1144): T, dT/dk, d2T/dk2 in one launch.  Between boundary
# components the two double-layer derivative terms carry the reference's sign
# (cheb-CFIE:1031-1032) by default, so callers get the reference's numbers;
# reference_sign=false gives them the opposite sign, the derivative of T.
function construct_boundary_matrices_with_derivatives!(Tbufs::Vector{Matrix{ComplexF64}},
                                                       dTbufs::Vector{Matrix{ComplexF64}},
                                                       ddTbufs::Vector{Matrix{ComplexF64}},
                                                       solver, pts,
                                                       zj::AbstractVector{ComplexF64};
                                                       multithreaded::Bool=true,
                                                       use_chebyshev::Bool=true,
                                                       n_panels_h::Int=15000, M_h::Int=5,
                                                       n_panels_j::Int=3000, M_j::Int=5,
                                                       timeit::Bool=false,
                                                       reference_sign::Bool=true)
    use_chebyshev || error("Direct derivative matrix construction is only for real k currently")
    Mk = length(zj)
    @assert length(Tbufs) == Mk && length(dTbufs) == Mk && length(ddTbufs) == Mk
    g = geom(solver, pts)
    try
        N = geom_dim(g)
        outs = [Array{ComplexF64}(undef, N, N, Mk) for _ in 1:3]
        cp = Ref(ChebParams(n_panels_h, M_h, n_panels_j, M_j))
        check(ccall((:qb_assemble_deriv, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Int64, Ref{ChebParams},
                     Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int32, Ptr{Float64}),
                    ctx(), g, opcode(solver), collect(zj), Int64(Mk), cp, outs[1], outs[2],
                    outs[3], Int32(reference_sign ? 2 : 8), C_NULL))
        for m in 1:Mk
            copyto!(Tbufs[m], view(outs[1], :, :, m))
            copyto!(dTbufs[m], view(outs[2], :, :, m))
            copyto!(ddTbufs[m], view(outs[3], :, :, m))
        end
    finally
        geom_destroy(g)
    end
    return nothing
end

# construct_B_matrix (beyn.jl:257-322): window on the device, SVD/eig on the host
function construct_B_matrix(solver, pts, N::Int, k0::Complex{Float64}, R::Float64; nq::Int=64,
                            r::Int=48, svd_tol=1e-14, rng=QB.MersenneTwister(0),
                            use_chebyshev=true, n_panels_h::Int=15000, M_h::Int=5,
                            n_panels_j::Int=3000, M_j::Int=5, info::Bool=false,
                            multithreaded::Bool=true, precision::Symbol=:fp64,
                            refine_tol::Float64=1e-12)
    g = geom(solver, pts)                          # N = geom_dim(g) with a symmetry
    w = Ref{Ptr{Cvoid}}(C_NULL)
    st = Ref(Stats())
    cp = Ref(ChebParams(n_panels_h, M_h, n_panels_j, M_j))
    prec = precision === :mixed ? QB_MIXED : precision === :mixed_int8res ? QB_MIXED_INT8RES : QB_FP64
    try
        check(ccall((:qb_window_create, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Float64, Float64, Float64, Int32,
                     Ref{ChebParams}, Int32, Ref{Ptr{Cvoid}}, Ref{Stats}),
                    ctx(), g, opcode(solver), real(k0), imag(k0), R, Int32(nq), cp, prec, w, st))
        # moments reduced on the device: A0 = Q R, C = (Q^H A1)[1:r, :]
        function moments_qr(V)
            r = size(V, 2)
            Rm = Matrix{ComplexF64}(undef, r, r); Cm = Matrix{ComplexF64}(undef, r, r)
            check(ccall((:qb_window_moments_qr, LIB), Cint,
                        (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64},
                         Float64, Ref{Stats}),
                        w[], Int32(r), V, Rm, Cm, refine_tol, st))
            return Rm, Cm
        end
        V = randn(rng, ComplexF64, N, r)
        Rm, Cm = moments_qr(V)
        U, Σ, W = svd(Rm)                          # same Σ, W as svd(A0)
        rk = count(>=(svd_tol), Σ)
        rk == 0 && return Matrix{ComplexF64}(undef, N, 0), Matrix{ComplexF64}(undef, N, 0)
        if rk == r && 2r < N                       # the single retry of beyn.jl:291-309
            V = randn(rng, ComplexF64, N, 2r)
            Rm, Cm = moments_qr(V)
            U, Σ, W = svd(Rm)
            rk = count(>=(svd_tol), Σ)
        end
        Ur = U[:, 1:rk]; Wk = W[:, 1:rk]
        B = adjoint(Ur) * ((Cm * Wk) ./ transpose(Σ[1:rk]))
        Uk = Matrix{ComplexF64}(undef, N, rk)      # U_k = Q [Ur; 0] on the device
        check(ccall((:qb_window_basis, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Int32, Ptr{ComplexF64}),
                    w[], Matrix(Ur), Int32(rk), Uk))
        return B, Uk
    finally
        w[] != C_NULL && ccall((:qb_window_destroy, LIB), Cvoid, (Ptr{Cvoid},), w[])
        geom_destroy(g)
    end
end

# Several disks on one geometry as one device window (qb_window_create_multi):
# no counterpart in the reference, whose driver runs its disks one after
# another; per disk the (B, U_k) of construct_B_matrix with the probes V.
function construct_B_matrices(solver, pts, N::Int, k0s::Vector{ComplexF64},
                              Rs::Vector{Float64}; nq::Int=64, r::Int=48, svd_tol=1e-14,
                              rng=QB.MersenneTwister(0), n_panels_h::Int=15000, M_h::Int=5,
                              n_panels_j::Int=3000, M_j::Int=5, precision::Symbol=:fp64,
                              refine_tol::Float64=1e-12)
    g = geom(solver, pts)
    w = Ref{Ptr{Cvoid}}(C_NULL)
    st = Ref(Stats())
    cp = Ref(ChebParams(n_panels_h, M_h, n_panels_j, M_j))
    prec = precision === :mixed ? QB_MIXED : precision === :mixed_int8res ? QB_MIXED_INT8RES : QB_FP64
    nw = length(k0s)
    k0v = collect(reinterpret(Float64, k0s))
    try
        check(ccall((:qb_window_create_multi, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32,
                     Ref{ChebParams}, Int32, Ref{Ptr{Cvoid}}, Ref{Stats}),
                    ctx(), g, opcode(solver), Int32(nw), k0v, Rs, Int32(nq), cp, prec, w, st))
        nb = Int(ccall((:qb_window_blocks, LIB), Int32, (Ptr{Cvoid},), w[]))
        V = randn(rng, ComplexF64, N, r)
        Rm = Array{ComplexF64}(undef, r, r, nb); Cm = Array{ComplexF64}(undef, r, r, nb)
        check(ccall((:qb_window_moments_qr, LIB), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64},
                     Float64, Ref{Stats}),
                    w[], Int32(r), V, Rm, Cm, refine_tol, st))
        out = Vector{Tuple{Matrix{ComplexF64},Matrix{ComplexF64}}}(undef, nb)
        for c in 1:nb
            U, Σ, W = svd(Rm[:, :, c])
            rk = count(>=(svd_tol), Σ)
            Ur = U[:, 1:rk]; Wk = W[:, 1:rk]
            B = adjoint(Ur) * ((Cm[:, :, c] * Wk) ./ transpose(Σ[1:rk]))
            Uk = Matrix{ComplexF64}(undef, N, rk)
            rk > 0 && check(ccall((:qb_window_basis_irrep, LIB), Cint,
                                  (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Int32, Ptr{ComplexF64}),
                                  w[], Int32(c - 1), Matrix(Ur), Int32(rk), Uk))
            out[c] = (B, Uk)
        end
        return out
    finally
        w[] != C_NULL && ccall((:qb_window_destroy, LIB), Cvoid, (Ptr{Cvoid},), w[])
        geom_destroy(g)
    end
end

# residual_and_norm_select (beyn.jl:557-593): T(λ_j) assembled and applied to
# φ_j = U_k Y[:, j] on the device; only the norms come back
function residual_and_norm_select(solver, λ::AbstractVector{ComplexF64},
                                  Uk::AbstractMatrix{ComplexF64}, Y::AbstractMatrix{ComplexF64},
                                  k0::ComplexF64, R::Float64, pts; res_tol::Float64,
                                  matnorm::Symbol=:one, epss::Real=1e-15,
                                  auto_discard_spurious::Bool=true, collect_logs::Bool=false,
                                  use_chebyshev::Bool=true, n_panels_h::Int=15000, M_h::Int=5,
                                  n_panels_j::Int=3000, M_j::Int=5, multithreaded::Bool=true)
    matnorm in (:one, :inf) || error("matnorm :two is not provided by libqbhip")
    N, rk = size(Uk)
    Φ = Uk * Y
    inside = findall(j -> abs(λ[j] - k0) <= R, eachindex(λ))
    tens = fill(NaN, rk); tensN = fill(NaN, rk); keep = falses(rk)
    logs = String[]
    if !isempty(inside)
        nl = length(inside)
        r2 = zeros(nl); r1 = zeros(nl); rinf = zeros(nl); n1 = zeros(nl); ninf = zeros(nl)
        g = geom(solver, pts)
        try
            cp = Ref(ChebParams(n_panels_h, M_h, n_panels_j, M_j))
            check(ccall((:qb_residuals, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Int64, Ref{ChebParams},
                         Ptr{ComplexF64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                         Ptr{Float64}, Ptr{Float64}),
                        ctx(), g, opcode(solver), λ[inside], Int64(nl), cp, Φ[:, inside],
                        Int32(0), r2, r1, rinf, n1, ninf))
        finally
            geom_destroy(g)
        end
        for (q, j) in pairs(inside)
            nA, φn, yn = matnorm === :one ? (n1[q], norm(Φ[:, j], 1), r1[q]) :
                                           (ninf[q], norm(Φ[:, j], Inf), rinf[q])
            tens[j] = r2[q]
            tensN[j] = yn / (nA * (φn + epss) + epss)
            if auto_discard_spurious && r2[q] >= res_tol
                collect_logs && push!(logs, "λ=$(λ[j]) ||Aφ||=$(r2[q]) > $res_tol → DROP")
            else
                keep[j] = true
                collect_logs && push!(logs, "λ=$(λ[j]) ||Aφ||=$(r2[q]) < $res_tol ← KEEP")
            end
        end
    end
    idx = findall(keep)
    return idx, Φ[:, idx], tens[idx], tensN[idx], logs
end

# Route the reference's entry points here for the Kress solvers.
function enable!()
    @eval QB function construct_boundary_matrices!(Tbufs::Vector{Matrix{ComplexF64}},
                                                   solver::Union{CFIE_kress,DLP_kress}, pts,
                                                   zj::AbstractVector{ComplexF64}; kw...)
        return $(QBHip).construct_boundary_matrices!(Tbufs, solver, pts, zj; kw...)
    end
    @eval QB function construct_B_matrix(solver::Union{CFIE_kress,DLP_kress}, pts, N::Int,
                                         k0::Complex{Float64}, R::Float64; kw...)
        return $(QBHip).construct_B_matrix(solver, pts, N, k0, R; kw...)
    end
    @eval QB function construct_boundary_matrices_with_derivatives!(
            Tbufs::Vector{Matrix{ComplexF64}}, dTbufs::Vector{Matrix{ComplexF64}},
            ddTbufs::Vector{Matrix{ComplexF64}}, solver::Union{CFIE_kress,DLP_kress}, pts,
            zj::AbstractVector{ComplexF64}; kw...)
        return $(QBHip).construct_boundary_matrices_with_derivatives!(Tbufs, dTbufs, ddTbufs,
                                                                      solver, pts, zj; kw...)
    end
    @eval QB function residual_and_norm_select(solver::Union{CFIE_kress,DLP_kress},
                                               λ::AbstractVector{ComplexF64},
                                               Uk::AbstractMatrix{ComplexF64},
                                               Y::AbstractMatrix{ComplexF64}, k0::ComplexF64,
                                               R::Float64, pts; kw...)
        return $(QBHip).residual_and_norm_select(solver, λ, Uk, Y, k0, R, pts; kw...)
    end
    return nothing
end

end # module
