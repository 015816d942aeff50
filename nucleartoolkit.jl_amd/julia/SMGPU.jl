# Julia binding of libsmgpu (include/smgpu.h) for NuclearToolkit.jl's main_sm path.
#
# The reference's seam is the body of `main_sm` (src/ShellModel/shellmodel_main.jl:89-281):
# readsmsnt -> occ -> prep_Hamil_T1 / prep_2bjumps / prep_Hamil_pn / prep_pn -> TRL
# -> Ritz assembly.  `SMGPU.main_sm` keeps that signature (target_J = [] or [2J],
# is_block / q) and returns main_sm's observable en[1]; `SMGPU.operate_H!` keeps operate_H!'s
# accumulate-into-twf contract (src/ShellModel/lanczos_methods.jl:153-248).
# Errors come back as Julia exceptions with the library's message.
module SMGPU

using Printf

const libsmgpu = get(ENV, "SMGPU_LIB", joinpath(@__DIR__, "..", "build", "libsmgpu.so"))

const FP64 = Int32(0)
const MIXED = Int32(1)

struct SolveParams              # smgpu_solve_params
    num_ev::Int32
    lm::Int32
    ls::Int32
    num_history::Int32
    tol::Float64
    precision::Int32
    refine::Int32
    max_iter::Int32
    seed::UInt64
    reorth::Int32                # REORTH_DEFAULT / _FULL / _PARTIAL
    steps::Int32                 # STEPS_AUTO / _LAUNCH / _GRAPH (captured Lanczos steps)
end
SolveParams(num_ev, lm, ls, num_history, tol, precision, refine, max_iter, seed) =
    SolveParams(num_ev, lm, ls, num_history, tol, precision, refine, max_iter, seed, Int32(0), Int32(0))

const REORTH_DEFAULT = Int32(0)
const REORTH_FULL = Int32(1)
const REORTH_PARTIAL = Int32(2)

struct SolveStats               # smgpu_solve_stats
    iterations::Int32
    n_hv::Int32
    n_hv64::Int32
    n_restart::Int32
    wall_ms::Float64
    ritz::NTuple{64,Float64}
    n_reorth2::Int32
    n_jproj::Int32
    n_reorth::Int32
    n_fallback::Int32
end

struct ProblemInfo              # smgpu_problem_info
    Z::Int32; N::Int32; A::Int32; vp::Int32; vn::Int32; twoM::Int32
    lp::Int32; ln::Int32; cp::Int32; cn::Int32
    nblocks::Int32; n_msps_p::Int32; n_msps_n::Int32
    mdim::Int64; n_psd::Int64; n_nsd::Int64
    n_pp_ops::Int64; n_nn_ops::Int64; n_pn_ops::Int64; n_hop_entries::Int64
end

function check(code::Integer)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:smgpu_last_error, libsmgpu), Cstring, ()))
    error("smgpu status $code: $msg")
end

mutable struct Problem
    h::Ptr{Cvoid}
end

function Problem(sntf::AbstractString, target_nuc::AbstractString; twoM::Integer=-1,
                 ignore_massop::Bool=false, nthreads::Integer=Threads.nthreads())
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_problem_create, libsmgpu), Int32,
                (Cstring, Cstring, Int32, Int32, Int32, Ref{Ptr{Cvoid}}),
                sntf, target_nuc, twoM, ignore_massop, nthreads, out))
    p = Problem(out[])
    finalizer(x -> ccall((:smgpu_problem_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), p)
    return p
end

function info(p::Problem)
    r = Ref{ProblemInfo}()
    check(ccall((:smgpu_problem_get_info, libsmgpu), Int32, (Ptr{Cvoid}, Ref{ProblemInfo}), p.h, r))
    return r[]
end

mutable struct Ham
    h::Ptr{Cvoid}
    mdim::Int
end

function Ham(p::Problem; device::Integer=0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_ham_create, libsmgpu), Int32, (Ptr{Cvoid}, Int32, Ref{Ptr{Cvoid}}), p.h, device, out))
    h = Ham(out[], info(p).mdim)
    finalizer(x -> ccall((:smgpu_ham_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), h)
    return h
end

# J^2 on the problem's basis (prep_J / operate_J!, shellmodel_main.jl:149-157, 226-230)
function HamJ2(p::Problem; device::Integer=0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_ham_create_j2, libsmgpu), Int32, (Ptr{Cvoid}, Int32, Ref{Ptr{Cvoid}}), p.h, device, out))
    h = Ham(out[], info(p).mdim)
    finalizer(x -> ccall((:smgpu_ham_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), h)
    return h
end

# Another interaction on base's basis that shares base's device index tables and
# uploads only p's coefficients (prepEC's samples on one basis,
# src/ShellModel/eigenvector_continuation.jl:140-175); p must have base's basis.
function HamVariant(base::Ham, p::Problem)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_ham_create_variant, libsmgpu), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                base.h, p.h, out))
    h = Ham(out[], base.mdim)
    finalizer(x -> ccall((:smgpu_ham_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), h)
    return h
end

# twf += H wf  (operate_H!'s contract; the device computes H wf in `precision`)
function operate_H!(wf::Vector{Float64}, twf::Vector{Float64}, H::Ham; precision::Int32=FP64)
    @assert length(wf) == H.mdim && length(twf) == H.mdim
    w = similar(wf)
    check(ccall((:smgpu_apply_H_host, libsmgpu), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                H.h, precision, wf, w))
    twf .+= w
    return nothing
end

function default_params()
    r = Ref{SolveParams}()
    ccall((:smgpu_default_params, libsmgpu), Cvoid, (Ref{SolveParams},), r)
    return r[]
end

# TRL + Ritz assembly; returns (en[1], Ritz vectors as mdim x num_ev or nothing, stats).
# j2 (HamJ2 on the same basis, built with twoM = twoJ) and twoJ: J-projected TRL (doubleLanczos).
function trl(H::Ham, num_ev::Integer; lm=300, ls=20, tol=1e-8, num_history=3, precision::Int32=FP64,
             refine::Bool=true, seed::Integer=1234, vectors::Bool=false, in_wf=nothing,
             j2::Union{Ham,Nothing}=nothing, twoJ::Integer=0)
    d = default_params()
    p = SolveParams(num_ev, lm, ls, num_history, tol, precision, refine, d.max_iter, UInt64(seed))
    nev = min(num_ev, H.mdim)
    evals = zeros(Float64, nev)
    evecs = vectors ? zeros(Float64, H.mdim, nev) : nothing   # column k = C row k of num_ev x mdim
    st = Ref{SolveStats}()
    v0 = in_wf === nothing ? C_NULL : pointer(in_wf)
    GC.@preserve in_wf evecs begin
        if j2 === nothing
            check(ccall((:smgpu_trl_solve_v0, libsmgpu), Int32,
                        (Ptr{Cvoid}, Ref{SolveParams}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
                        H.h, p, v0, evals, evecs === nothing ? C_NULL : pointer(evecs), st))
        else
            check(ccall((:smgpu_trl_solve_j, libsmgpu), Int32,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ref{SolveParams}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                         Ref{SolveStats}),
                        H.h, j2.h, twoJ, p, v0, evals, evecs === nothing ? C_NULL : pointer(evecs), st))
        end
    end
    return evals, evecs, st[]
end

# Thick-restart block Lanczos (TRBL) with block size q; in_wf: mdim x q start block
function trbl(H::Ham, num_ev::Integer, q::Integer; lm=300, ls=20, tol=1e-8, num_history=3,
              precision::Int32=FP64, refine::Bool=true, seed::Integer=1234, vectors::Bool=false,
              in_wf=nothing, j2::Union{Ham,Nothing}=nothing, twoJ::Integer=0)
    d = default_params()
    p = SolveParams(num_ev, lm, ls, num_history, tol, precision, refine, d.max_iter, UInt64(seed))
    nev = min(num_ev, H.mdim)
    evals = zeros(Float64, nev)
    evecs = vectors ? zeros(Float64, H.mdim, nev) : nothing
    st = Ref{SolveStats}()
    in_wf === nothing || @assert size(in_wf) == (H.mdim, q)
    v0 = in_wf === nothing ? C_NULL : pointer(in_wf)
    GC.@preserve in_wf evecs begin
        check(ccall((:smgpu_trbl_solve, libsmgpu), Int32,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ref{SolveParams}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ref{SolveStats}),
                    H.h, j2 === nothing ? C_NULL : j2.h, twoJ, q, p, v0, evals,
                    evecs === nothing ? C_NULL : pointer(evecs), st))
    end
    return evals, evecs, st[]
end

# TRL solves of interaction variants on one basis (HamVariant handles; a handle may
# repeat) in lockstep, the H·v of all members in shared launches each round
# (smgpu_trl_solve_variants); member b starts from seed seeds[b].  Returns
# [(evals, stats)] per member, each equal to trl(Hs[b], ...) with that seed.
function trl_variants(Hs::Vector{Ham}, num_ev::Integer; seeds=[1234 + b for b in 0:length(Hs)-1], lm=300,
                      ls=20, tol=1e-8, num_history=3, precision::Int32=FP64, refine::Bool=true)
    nb = length(Hs)
    @assert length(seeds) == nb
    d = default_params()
    ps = [SolveParams(num_ev, lm, ls, num_history, tol, precision, refine, d.max_iter, UInt64(sd)) for sd in seeds]
    nev = min(num_ev, Hs[1].mdim)
    evals = [zeros(Float64, nev) for _ in 1:nb]
    stats = Vector{SolveStats}(undef, nb)
    status = zeros(Int32, nb)
    hs = [H.h for H in Hs]
    GC.@preserve evals begin
        evp = [pointer(e) for e in evals]
        check(ccall((:smgpu_trl_solve_variants, libsmgpu), Int32,
                    (Ptr{Ptr{Cvoid}}, Int32, Ptr{SolveParams}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}},
                     Ptr{Ptr{Float64}}, Ptr{SolveStats}, Ptr{Int32}),
                    hs, nb, ps, C_NULL, evp, C_NULL, stats, status))
    end
    return [(evals[b], stats[b]) for b in 1:nb]
end

# main_sm(sntf, target_nuc, num_ev, target_J; ...): en[1].  target_J = [2J] runs the
# J-projected solve at 2M = 2J (:109-116); is_block=true the block Lanczos with block size q.
# main_sm's dimension line (mdim_print, src/ShellModel/shellmodel_main.jl:1316-1321)
function mdim_print(target_nuc, inf::ProblemInfo, tJ=-1)
    s0 = @sprintf "%6s %14s %10s %10s" target_nuc "Z,N=($(inf.Z),$(inf.N))" "c($(inf.cp),$(inf.cn))" "v($(inf.vp),$(inf.vn))"
    if tJ != -1; s0 *= @sprintf(" 2J = %3i", tJ); end
    s1 = @sprintf "%s %12i %s %5.2f %s " "  mdim:" inf.mdim "(" log10(max(inf.mdim, 1)) ")"
    print(s0, s1)
end

# mdimmode=true prints the dimension line and returns nothing (:129-130)
# M1 / E2 one-body operators of main_sm's calc_moment on a problem's basis
# (prep_obtr_diag, src/ShellModel/transit.jl:82-245), on one GPU
mutable struct Moments
    h::Ptr{Cvoid}
end
function Moments(p::Problem; device::Integer=0, gfactors=[1.0, 0.0, 5.586, -3.826], effcharge=[1.5, 0.5],
                 bpar::Real=0.0)
    g = Vector{Float64}(gfactors)
    e = Vector{Float64}(effcharge)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_moments_create, libsmgpu), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Float64, Ref{Ptr{Cvoid}}),
                p.h, device, g, e, Float64(bpar), out))
    m = Moments(out[])
    finalizer(x -> ccall((:smgpu_moments_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), m)
    return m
end

# eval_moment (transit.jl:702-765): mu and Q of the columns of vecs (mdim x n) with J
# values totJs; NaN where the reference prints nothing
function eval_moment(m::Moments, vecs::Matrix{Float64}, totJs::Vector{Float64})
    n = size(vecs, 2)
    mu = zeros(Float64, n)
    Q = zeros(Float64, n)
    check(ccall((:smgpu_eval_moment, libsmgpu), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                m.h, n, vecs, totJs, mu, Q))
    return mu, Q
end

# M1 / E2 operators of transit_main between two problems on one interaction and nucleus,
# the left at 2M_l and the right at 2M_r <= 2M_l (prep_obtr_nondiag, transit.jl:286-445)
mutable struct Transitions
    h::Ptr{Cvoid}
end
function Transitions(left::Problem, right::Problem; device::Integer=0, gfactors=[1.0, 0.0, 5.586, -3.826],
                     effcharge=[1.5, 0.5], bpar::Real=0.0)
    g = Vector{Float64}(gfactors)
    e = Vector{Float64}(effcharge)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:smgpu_transitions_create, libsmgpu), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Float64, Ref{Ptr{Cvoid}}),
                left.h, right.h, device, g, e, Float64(bpar), out))
    t = Transitions(out[])
    finalizer(x -> ccall((:smgpu_transitions_destroy, libsmgpu), Int32, (Ptr{Cvoid},), x.h), t)
    return t
end

# reduced matrix elements <l||O||r> of the columns of wl (mdim_l x nl) and wr (mdim_r x nr):
# (M1, E2), each nl x nr, NaN where the transition is not allowed
function transit(t::Transitions, jl2::Integer, jr2::Integer, wl::Matrix{Float64}, wr::Matrix{Float64})
    nl = size(wl, 2); nr = size(wr, 2)
    m1 = zeros(Float64, nr, nl)          # C row-major [l * nr + r] = Julia column-major (r, l)
    e2 = zeros(Float64, nr, nl)
    check(ccall((:smgpu_transit, libsmgpu), Int32,
                (Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                t.h, jl2, jr2, nl, wl, nr, wr, m1, e2))
    return permutedims(m1), permutedims(e2)
end

# read_wav(all=true) (shellmodel_main.jl:979-1003): energies, 2J values, vectors (mdim x n)
function read_wav(path::AbstractString, mdim::Integer)
    open(path, "r") do io
        n = read(io, Int32); mtot = read(io, Int32)
        es = [read(io, Float64) for _ in 1:n]
        jj = [read(io, Int32) for _ in 1:n]
        V = zeros(Float64, mdim, n)
        read!(io, V)
        return es, jj, V, mtot
    end
end

# Wigner 3j symbol of doubled arguments (Racah's formula)
function w3j(j1, j2, j3, m1, m2, m3)
    (m1 + m2 + m3 != 0 || abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) && return 0.0
    (j3 < abs(j1 - j2) || j3 > j1 + j2 || isodd(j1 + j2 + j3)) && return 0.0
    f(x) = factorial(big(x))
    a = (j1 + j2 - j3) ÷ 2; b = (j1 - j2 + j3) ÷ 2; c = (-j1 + j2 + j3) ÷ 2
    pre = f(a) * f(b) * f(c) // f((j1 + j2 + j3) ÷ 2 + 1) * f((j1 + m1) ÷ 2) * f((j1 - m1) ÷ 2) *
          f((j2 + m2) ÷ 2) * f((j2 - m2) ÷ 2) * f((j3 + m3) ÷ 2) * f((j3 - m3) ÷ 2)
    s = big(0) // 1
    for k in 0:a
        d = (k, a - k, (j1 - m1) ÷ 2 - k, (j2 + m2) ÷ 2 - k, (j3 - j2 + m1) ÷ 2 + k, (j3 - j1 - m2) ÷ 2 + k)
        minimum(d) < 0 && continue
        s += (-1)^k // prod(f.(d))
    end
    sgn = isodd((j1 - j2 - m3) ÷ 2) ? -1.0 : 1.0
    return sgn * Float64(s) * sqrt(Float64(pre))
end

# transit_main (transit.jl:561-700): M1 and E2 transitions between the states of two .wav
# files, the left at 2M = 2J = jl2 and the right at jr2 (swapped when jl2 < jr2); prints
# the reference's lines and returns (Mred M1, Mred E2) as nl x nr matrices
function transit_main(sntf, target_nuc, jl2, jr2, in_wfs; num_ev_l=100, num_ev_r=100, calc_EM=true,
                      gfactors=[1.0, 0.0, 5.586, -3.826], eff_charge=[1.5, 0.5], device::Integer=0)
    length(in_wfs) == 2 || error("input wfs must be specified")
    widx = [1, 2]
    if jl2 < jr2
        jl2, jr2 = jr2, jl2; widx = [2, 1]
    end
    (jl2 - jr2) ÷ 2 > 2 && error("Only the M1&E2 are supported now")
    left = Problem(sntf, target_nuc; twoM=jl2)
    right = Problem(sntf, target_nuc; twoM=jr2)
    _, _, Vl, _ = read_wav(in_wfs[widx[1]], info(left).mdim)
    _, _, Vr, _ = read_wav(in_wfs[widx[2]], info(right).mdim)
    Vl = Vl[:, 1:min(num_ev_l, size(Vl, 2))]; Vr = Vr[:, 1:min(num_ev_r, size(Vr, 2))]
    m1, e2 = transit(Transitions(left, right; device=device, gfactors=gfactors, effcharge=eff_charge), jl2, jr2,
                     Vl, Vr)
    same = in_wfs[1] == in_wfs[2]
    for (mode, lam2, red, cmom, tag) in (("M1", 2, m1, sqrt(4pi / 3), "<gm>"), ("E2", 4, e2, sqrt(16pi / 5), "<eQ>"))
        all(isnan, red) && continue
        cg = w3j(jr2, lam2, jl2, -jr2, jl2 - jr2, jl2)
        for r in 1:size(red, 2), l in 1:size(red, 1)
            (!calc_EM && l != r) && continue
            Mred = red[l, r]
            if calc_EM
                tx = "Mred.($mode) ($l,$r) " * @sprintf("%9.3f ", Mred) * "  "
                tx *= "B($mode) ($l<=$r) " * @sprintf("%9.3f ", Mred^2 / (jr2 + 1)) * " B($mode) ($l=>$r) " *
                      @sprintf("%9.3f ", Mred^2 / (jl2 + 1))
                (same && l == r) && (tx *= " $tag= " * @sprintf("%9.3f ", Mred * cg * cmom))
            else
                tx = " $tag= " * @sprintf("%9.3f ", Mred * cg * cmom)
            end
            println(tx)
        end
    end
    return m1, e2
end

# J from <J^2> (J_from_JJ1, shellmodel_main.jl:1105-1113): the J in 0, 1/2, ..., 50 with
# |J(J+1) - jj| < tol, else -1
function j_from_jj1(jj::Real; tol=1e-6)
    for J2 in 0:100
        hJ = 0.5 * J2
        abs(hJ * (hJ + 1.0) - jj) < tol && return hJ
    end
    return -1.0
end

# <v|J^2|v> of the columns of vecs with the J^2 operator on the same basis
function expect_j2(j2::Ham, vecs::Matrix{Float64})
    n = size(vecs, 2)
    w = similar(vecs)
    check(ccall((:smgpu_apply_H_host_multi, libsmgpu), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}), j2.h, FP64, n, vecs, w))
    return [sum(view(vecs, :, k) .* view(w, :, k)) for k in 1:n]
end

function main_sm(sntf, target_nuc, num_ev, target_J; num_history=3, lm=300, ls=20, tol=1e-8,
                 precision::Int32=FP64, device::Integer=0, q::Integer=1, is_block::Bool=false,
                 mdimmode::Bool=false, calc_moment::Bool=false, gfactors=[1.0, 0.0, 5.586, -3.826],
                 effcharge=[1.5, 0.5])
    length(target_J) <= 1 || error("Multiple J is not supported")
    tJ = isempty(target_J) ? -1 : Int(target_J[1])
    prob = Problem(sntf, target_nuc; twoM=tJ)
    inf = info(prob)
    if mdimmode
        mdim_print(target_nuc, inf, tJ)
        println()
        return nothing
    end
    mdim = inf.mdim
    mdim == 0 && return Float64[]
    H = Ham(prob; device=device)
    j2 = tJ < 0 ? nothing : HamJ2(prob; device=device)
    nev = min(num_ev, mdim)
    if is_block
        en, vecs, _ = trbl(H, nev, q; lm=lm, ls=ls, tol=tol, num_history=num_history, precision=precision,
                           j2=j2, twoJ=max(tJ, 0), vectors=calc_moment)
    else
        en, vecs, _ = trl(H, nev; lm=lm, ls=ls, tol=tol, num_history=num_history, precision=precision,
                          j2=j2, twoJ=max(tJ, 0), vectors=calc_moment)
    end
    if calc_moment
        # main_sm's moments (shellmodel_main.jl:242-246, 267-269)
        jop = j2 === nothing ? HamJ2(prob; device=device) : j2
        totJs = [j_from_jj1(x) for x in expect_j2(jop, vecs)]
        mu, Q = eval_moment(Moments(prob; device=device, gfactors=gfactors, effcharge=effcharge), vecs, totJs)
        tx = ""
        for k in eachindex(mu)
            t = ""
            isnan(mu[k]) || (t *= "mu (n=$k) " * @sprintf("%12.5f ", mu[k]) * "\t")
            isnan(Q[k]) || (t *= "Q (n=$k) " * @sprintf("%12.5f ", Q[k]))
            t == "" || (tx *= t * "\n")
        end
        tx == "" || println(tx)
    end
    return en
end

end # module
