# LowRankApproxMI355X.jl -- Julia host shim over liblra_hip.so (include/lra.h).
#
# Adds methods to LowRankApprox.jl's own functions for device matrices (AMDGPU.jl ROCArrays)
# and for device operators, so that a caller of the reference keeps its API:
#
#     using LowRankApprox, AMDGPU, LowRankApproxMI355X
#     A = ROCArray(rand(8192, 8192))
#     V = idfact(A, LRAOptions(rtol=1e-10))          # IDPackedV with device sk, rd, T
#     F = psvdfact(A; rtol=1e-10)                     # PartialSVD with device factors
#     L = ROCLinOp(m, n, (Y, X) -> mul!(Y, M, X), (Y, X) -> mul!(Y, M', X))
#     F = pqrfact(:n, L; sketch_randn_niter=1)        # matrix-free, power iteration via L, L'
#
# The entry points and the reference interfaces they replace:
#   idfact / idfact(::LinOp)      src/id.jl:434-447              lra_idfact / lra_idfact_op
#   pqrfact                       src/pqr.jl:285-320             lra_pqrfact / lra_pqrfact_op
#   psvdfact, psvdvals            src/psvd.jl:238-290            lra_psvdfact(_op), lra_psvdvals(_op)
#   pheigfact, pheigvals          src/pheig.jl:276-314           lra_pheigfact(_op), lra_pheigvals(_op)
#   curfact                       src/cur.jl:527-564             lra_curfact / lra_curfact_op
#   prange                        src/prange.jl:14-77            lra_prange
#   snorm, snormdiff              src/snorm.jl:14-53             lra_snorm, lra_snormdiff(_id), lra_snorm_op
#   _LAPACK.laqps!                src/lapack.jl:117-139          lra_laqps_f64
#   geqp3_adap!                   src/pqr.jl:348-418             lra_geqp3_adap_f64
#
# Conventions of the library: column-major Float64, Int64 sizes, 1-based indices, device
# pointers, status 0 / -(bad argument) / LRA_E_*.  Complex and Float32 device matrices are
# not taken by the library: Float32 is converted to Float64 here with LRAOptions(Float32)'s
# tolerances; complex inputs keep the reference's CPU path.
#
# This module is written against the C ABI and AMDGPU.jl's ROCArray; the build container has
# no Julia, so it is exercised only through the identical Python ctypes mirror
# (lowrankapprox.jl_amd/__init__.py) that the test suite runs.
module LowRankApproxMI355X

using LinearAlgebra
using AMDGPU
import LowRankApprox
import LowRankApprox: LRAOptions, IDPackedV, PartialQR, PartialSVD, PartialHermEigen,
                      CURPackedU, idfact, pqrfact, psvdfact, psvdvals, pheigfact, pheigvals,
                      curfact, prange, snorm, snormdiff

export LraCtx, ROCLinOp, laqps_gpu!, geqp3_adap_gpu!

const liblra = get(ENV, "LRA_LIB", joinpath(@__DIR__, "..", "liblra_hip.so"))

const SKETCH = Dict(:none => 0, :randn => 1, :srft => 2, :sub => 3, :sprn => 4)
const PREC = Dict(:f64 => 0, :f32 => 1, :bf16 => 2, :bf16x3 => 3)
const LRA_E_UNSUPPORTED = 3

# ---------------------------------------------------------------------------------------
# context and errors

mutable struct LraCtx
    h::Ptr{Cvoid}
end
function LraCtx(device::Integer=0; hbm_budget::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:lra_ctx_create, liblra), Cint, (Cint, Csize_t, Ptr{Ptr{Cvoid}}),
                device, hbm_budget, r), "lra_ctx_create")
    ctx = LraCtx(r[])
    finalizer(c -> ccall((:lra_ctx_destroy, liblra), Cint, (Ptr{Cvoid},), c.h), ctx)
end

const _CTX = Ref{Union{Nothing,LraCtx}}(nothing)
ctx() = (c = _CTX[]; c === nothing ? (_CTX[] = LraCtx(AMDGPU.device_id(AMDGPU.device()) - 1)) : c)

lasterror() = unsafe_string(ccall((:lra_last_error, liblra), Cstring, ()))
function check(st::Integer, what)
    st == 0 && return nothing
    st < 0 && throw(ArgumentError("$what: argument $(-st): $(lasterror())"))
    error("$what failed with status $st: $(lasterror())")
end

# enqueue the library's work on the task's AMDGPU stream, so it orders with the caller's
bind_stream!(c::LraCtx) =
    check(ccall((:lra_ctx_set_stream, liblra), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                c.h, AMDGPU.stream().stream), "lra_ctx_set_stream")

dptr(A::ROCArray) = reinterpret(Ptr{eltype(A)}, pointer(A))
ld(A::ROCMatrix) = max(1, stride(A, 2))

# ---------------------------------------------------------------------------------------
# options: lra_opts mirrors LRAOptions field by field (src/LowRankApprox.jl:77-94)

struct LraOpts
    atol::Float64; maxdet_niter::Int64; maxdet_tol::Float64; nb::Int64
    pheig_orthtol::Float64; pqrfact_retval::Int32; rank::Int64; rtol::Float64
    sketch::Int32; sketch_randn_niter::Int64; sketchfact_adap::Int32
    randn_samp_mul::Int64; randn_samp_add::Int64; srft_samp_mul::Int64; srft_samp_add::Int64
    sub_samp_mul::Int64; sub_samp_add::Int64; snorm_niter::Int64; verb::Int32
    precision::Int32; seed::UInt64
end

# the affine sample maps n -> a n + b of the default sketchfact_*_samp functions
affine(f) = (b = f(0); (f(1) - b, b))

function LraOpts(o::LRAOptions; precision::Symbol=:f64, seed::Integer=0)
    rv = Int32(0)
    occursin('q', o.pqrfact_retval) && (rv |= 1)
    occursin('r', o.pqrfact_retval) && (rv |= 2)
    occursin('t', o.pqrfact_retval) && (rv |= 4)
    ra, rb = affine(o.sketchfact_randn_samp)
    sa, sb = affine(o.sketchfact_srft_samp)
    ua, ub = affine(o.sketchfact_sub_samp)
    LraOpts(o.atol, o.maxdet_niter, o.maxdet_tol, o.nb, o.pheig_orthtol, rv, o.rank, o.rtol,
            SKETCH[o.sketch], o.sketch_randn_niter, o.sketchfact_adap, ra, rb, sa, sb, ua, ub,
            o.snorm_niter, o.verb, PREC[precision], UInt64(seed))
end

opts_of(::Type{T}, opts, args) where {T} =
    (o = isnothing(opts) ? LRAOptions(T) : opts; isempty(args) ? copy(o) : copy(o; args...))

# device Float64 operand: Float32 is widened (the factorization runs in Float64 with the
# caller's Float32 tolerances)
f64(A::ROCMatrix{Float64}) = A
f64(A::ROCMatrix{Float32}) = ROCArray{Float64}(A)

# ---------------------------------------------------------------------------------------
# result records (include/lra.h)

mutable struct LraIdOut; sk::Ptr{Int64}; rd::Ptr{Int64}; T::Ptr{Float64}; ldt::Int64
    k::Int64; rounds::Int64; end
mutable struct LraPqrOut; Q::Ptr{Float64}; ldq::Int64; R::Ptr{Float64}; ldr::Int64
    p::Ptr{Int64}; T::Ptr{Float64}; ldt::Int64; k::Int64; rounds::Int64; end
mutable struct LraPsvdOut; U::Ptr{Float64}; ldu::Int64; S::Ptr{Float64}; Vt::Ptr{Float64}
    ldvt::Int64; k::Int64; end
mutable struct LraPheigOut; values::Ptr{Float64}; vectors::Ptr{Float64}; ldv::Int64
    k::Int64; kn::Int64; kp::Int64; end

# copy a library-owned column-major block (valid until the next call on the context)
function own(p::Ptr{Float64}, rows, cols, ldp)
    (rows == 0 || cols == 0) && return ROCArray{Float64}(undef, rows, cols)
    W = unsafe_wrap(ROCArray, p, (ldp, cols); own=false)
    copy(view(W, 1:rows, :))
end
own(p::Ptr{Int64}, n) = n == 0 ? ROCArray{Int64}(undef, 0) :
    copy(unsafe_wrap(ROCArray, p, (n,); own=false))

# ---------------------------------------------------------------------------------------
# matrix-free operators on device vectors (LinearOperator, src/linop.jl:9-40)

struct ROCLinOp
    m::Int; n::Int
    mul!::Function      # (Y, X) -> Y = L X   on ROCMatrix blocks
    mulc!::Function     # (Y, X) -> Y = L' X
    hermitian::Bool
end
ROCLinOp(m, n, mul!, mulc!=mul!; hermitian::Bool=false) = ROCLinOp(m, n, mul!, mulc!, hermitian)
Base.size(L::ROCLinOp) = (L.m, L.n)

# lra_apply_fn: Y (ldy) = op(L) X (ldx); user = the ROCLinOp
function apply_cb(user::Ptr{Cvoid}, trans::Cint, nrhs::Int64, X::Ptr{Float64}, ldx::Int64,
                  Y::Ptr{Float64}, ldy::Int64, ::Ptr{Cvoid})::Cint
    try
        L = unsafe_pointer_to_objref(user)::ROCLinOp
        rin, rout = trans == 0 ? (L.n, L.m) : (L.m, L.n)
        Xw = view(unsafe_wrap(ROCArray, X, (ldx, nrhs); own=false), 1:rin, :)
        Yw = view(unsafe_wrap(ROCArray, Y, (ldy, nrhs); own=false), 1:rout, :)
        trans == 0 ? L.mul!(Yw, Xw) : L.mulc!(Yw, Xw)
        return Cint(0)
    catch
        return Cint(1)
    end
end
const APPLY = Ref{Ptr{Cvoid}}(C_NULL)
function __init__()
    APPLY[] = @cfunction(apply_cb, Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64,
                                          Ptr{Float64}, Int64, Ptr{Cvoid}))
end

# chkopts!(opts, L) (src/LowRankApprox.jl:142-148)
function linop_opts(o::LRAOptions)
    if o.sketch != :randn
        @warn "invalid sketch method; using \"randn\""
        o.sketch = :randn
    end
    o
end

# ---------------------------------------------------------------------------------------
# ?laqps and geqp3_adap! (src/lapack.jl:117-139, src/pqr.jl:348-418)

function laqps_gpu!(offset::Integer, nb::Integer, A::ROCMatrix{Float64}, jpvt::ROCVector{Int64},
                    tau::ROCVector{Float64}, vn1::ROCVector{Float64}, vn2::ROCVector{Float64})
    c = ctx(); bind_stream!(c)
    m, n = size(A); kb = Ref{Int64}(0)
    check(ccall((:lra_laqps_f64, liblra), Cint,
                (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Float64}, Int64,
                 Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Int64),
                c.h, m, n, offset, nb, kb, dptr(A), ld(A), dptr(jpvt), dptr(tau), dptr(vn1),
                dptr(vn2), C_NULL, C_NULL, 0), "lra_laqps_f64")
    kb[]
end

function geqp3_adap_gpu!(B::ROCMatrix{Float64}, opts::LRAOptions)
    c = ctx(); bind_stream!(c)
    m, n = size(B)
    jpvt = ROCVector{Int64}(undef, n); tau = ROCVector{Float64}(undef, max(1, min(m, n)))
    k = Ref{Int64}(0)
    check(ccall((:lra_geqp3_adap_f64, liblra), Cint,
                (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Int64, Float64, Float64, Int64,
                 Ptr{Int64}, Ptr{Float64}, Ptr{Int64}),
                c.h, m, n, dptr(B), ld(B), opts.rank, opts.atol, opts.rtol, opts.nb, dptr(jpvt),
                dptr(tau), k), "lra_geqp3_adap_f64")
    jpvt, tau, k[]
end

# ---------------------------------------------------------------------------------------
# idfact (src/id.jl:434-447)

function id_result(out::LraIdOut, sk, rd, N)
    k = out.k
    IDPackedV(Array(sk[1:k]), Array(rd[1:N-k]), Array(own(out.T, k, N - k, out.ldt)))
end

function idfact(trans::Symbol, A::ROCMatrix{T}, opts::Union{Nothing,LRAOptions}=nothing;
                precision::Symbol=:f64, seed::Integer=0, args...) where {T<:Union{Float32,Float64}}
    o = opts_of(T, opts, args); o.pqrfact_retval = "t"
    Ad = f64(A); m, n = size(Ad); N = trans == :n ? n : m
    c = ctx(); bind_stream!(c)
    sk = ROCVector{Int64}(undef, max(N, 1)); rd = ROCVector{Int64}(undef, max(N, 1))
    out = LraIdOut(dptr(sk), dptr(rd), C_NULL, 0, 0, 0)
    co = LraOpts(o; precision, seed)
    check(ccall((:lra_idfact, liblra), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts}, Ref{LraIdOut}),
                c.h, trans == :n ? 0 : 1, m, n, dptr(Ad), ld(Ad), co, out), "lra_idfact")
    id_result(out, sk, rd, N)
end
idfact(A::ROCMatrix, args...; kw...) = idfact(:n, A, args...; kw...)

function idfact(trans::Symbol, L::ROCLinOp, opts::Union{Nothing,LRAOptions}=nothing;
                seed::Integer=0, args...)
    o = linop_opts(opts_of(Float64, opts, args)); o.pqrfact_retval = "t"
    N = trans == :n ? L.n : L.m
    c = ctx(); bind_stream!(c)
    sk = ROCVector{Int64}(undef, max(N, 1)); rd = ROCVector{Int64}(undef, max(N, 1))
    out = LraIdOut(dptr(sk), dptr(rd), C_NULL, 0, 0, 0)
    co = LraOpts(o; seed)
    GC.@preserve L check(ccall((:lra_idfact_op, liblra), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Any, Cint, Ref{LraOpts},
                 Ref{LraIdOut}),
                c.h, trans == :n ? 0 : 1, L.m, L.n, APPLY[], L, L.hermitian, co, out),
                "lra_idfact_op")
    id_result(out, sk, rd, N)
end
idfact(L::ROCLinOp, args...; kw...) = idfact(:n, L, args...; kw...)

# ---------------------------------------------------------------------------------------
# pqrfact (src/pqr.jl:285-320): PartialQR for retval "qr", else the named factors

function pqr_result(out::LraPqrOut, rv::AbstractString, P, N)
    k = out.k
    Q = occursin('q', rv) ? Array(own(out.Q, P, k, out.ldq)) : nothing
    R = occursin('r', rv) ? Array(own(out.R, k, N, out.ldr)) : nothing
    p = Array(own(out.p, N))
    T = occursin('t', rv) ? Array(own(out.T, k, N - k, max(out.ldt, 1))) : nothing
    (occursin('q', rv) && occursin('r', rv) && !occursin('t', rv)) ? PartialQR(Q, R, p) :
        (Q=Q, R=R, p=p, k=k, T=T)
end

function pqrfact(trans::Symbol, A::Union{ROCMatrix,ROCLinOp}, opts=nothing; seed::Integer=0,
                 precision::Symbol=:f64, args...)
    isop = A isa ROCLinOp
    o = opts_of(isop ? Float64 : eltype(A), opts, args)
    isop && linop_opts(o)
    m, n = size(A); N = trans == :n ? n : m; P = trans == :n ? m : n
    c = ctx(); bind_stream!(c)
    out = LraPqrOut(C_NULL, 0, C_NULL, 0, C_NULL, C_NULL, 0, 0, 0)
    co = LraOpts(o; precision, seed)
    if isop
        GC.@preserve A check(ccall((:lra_pqrfact_op, liblra), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Any, Cint, Ref{LraOpts},
                     Ref{LraPqrOut}),
                    c.h, trans == :n ? 0 : 1, m, n, APPLY[], A, A.hermitian, co, out),
                    "lra_pqrfact_op")
    else
        Ad = f64(A)
        check(ccall((:lra_pqrfact, liblra), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts},
                     Ref{LraPqrOut}),
                    c.h, trans == :n ? 0 : 1, m, n, dptr(Ad), ld(Ad), co, out), "lra_pqrfact")
    end
    pqr_result(out, o.pqrfact_retval, P, N)
end
pqrfact(A::Union{ROCMatrix,ROCLinOp}, args...; kw...) = pqrfact(:n, A, args...; kw...)

function prange(trans::Symbol, A::ROCMatrix, opts=nothing; seed::Integer=0, args...)
    o = opts_of(eltype(A), opts, args)
    Ad = f64(A); m, n = size(Ad); P = trans == :c ? n : m
    c = ctx(); bind_stream!(c)
    out = LraPqrOut(C_NULL, 0, C_NULL, 0, C_NULL, C_NULL, 0, 0, 0)
    check(ccall((:lra_prange, liblra), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts},
                 Ref{LraPqrOut}),
                c.h, Dict(:n => 0, :c => 1, :b => 2)[trans], m, n, dptr(Ad), ld(Ad),
                LraOpts(o; seed), out), "lra_prange")
    Array(own(out.Q, P, out.k, out.ldq))
end
prange(A::ROCMatrix, args...; kw...) = prange(:n, A, args...; kw...)

# ---------------------------------------------------------------------------------------
# psvdfact / psvdvals (src/psvd.jl:238-290)

function psvdfact(A::Union{ROCMatrix,ROCLinOp}, opts=nothing; seed::Integer=0,
                  precision::Symbol=:f64, args...)
    isop = A isa ROCLinOp
    o = opts_of(isop ? Float64 : eltype(A), opts, args)
    isop && linop_opts(o)
    m, n = size(A)
    c = ctx(); bind_stream!(c)
    out = LraPsvdOut(C_NULL, 0, C_NULL, C_NULL, 0, 0)
    co = LraOpts(o; precision, seed)
    if isop
        GC.@preserve A check(ccall((:lra_psvdfact_op, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Any, Cint, Ref{LraOpts},
                     Ref{LraPsvdOut}),
                    c.h, m, n, APPLY[], A, A.hermitian, co, out), "lra_psvdfact_op")
    else
        Ad = f64(A)
        check(ccall((:lra_psvdfact, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts}, Ref{LraPsvdOut}),
                    c.h, m, n, dptr(Ad), ld(Ad), co, out), "lra_psvdfact")
    end
    k = out.k
    PartialSVD(Array(own(out.U, m, k, out.ldu)), vec(Array(own(out.S, k, 1, max(k, 1)))),
               Array(own(out.Vt, k, n, out.ldvt)))
end

function psvdvals(A::ROCMatrix, opts=nothing; seed::Integer=0, args...)
    o = opts_of(eltype(A), opts, args)
    Ad = f64(A); m, n = size(Ad)
    c = ctx(); bind_stream!(c)
    S = ROCVector{Float64}(undef, max(1, min(m, n))); k = Ref{Int64}(0)
    check(ccall((:lra_psvdvals, liblra), Cint,
                (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts}, Ptr{Float64},
                 Ref{Int64}),
                c.h, m, n, dptr(Ad), ld(Ad), LraOpts(o; seed), dptr(S), k), "lra_psvdvals")
    Array(S[1:k[]])
end

# ---------------------------------------------------------------------------------------
# pheigfact / pheigvals (src/pheig.jl:276-314)

function pheigfact(A::Union{ROCMatrix,ROCLinOp}, opts=nothing; seed::Integer=0, args...)
    isop = A isa ROCLinOp
    isop && !A.hermitian && error("matrix must be Hermitian")
    o = opts_of(isop ? Float64 : eltype(A), opts, args)
    isop && linop_opts(o)
    n = size(A, 1)
    c = ctx(); bind_stream!(c)
    out = LraPheigOut(C_NULL, C_NULL, 0, 0, 0, 0)
    co = LraOpts(o; seed)
    if isop
        GC.@preserve A check(ccall((:lra_pheigfact_op, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Any, Ref{LraOpts}, Ref{LraPheigOut}),
                    c.h, n, APPLY[], A, co, out), "lra_pheigfact_op")
    else
        Ad = f64(A)
        ishermitian(Ad) || error("matrix must be Hermitian")
        check(ccall((:lra_pheigfact, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ref{LraOpts}, Ref{LraPheigOut}),
                    c.h, n, dptr(Ad), ld(Ad), co, out), "lra_pheigfact")
    end
    k = out.k
    PartialHermEigen(vec(Array(own(out.values, k, 1, max(k, 1)))),
                     Array(own(out.vectors, n, k, out.ldv)))
end

# ---------------------------------------------------------------------------------------
# curfact (src/cur.jl:527-564)

function curfact(A::Union{ROCMatrix,ROCLinOp}, opts=nothing; seed::Integer=0, args...)
    isop = A isa ROCLinOp
    o = opts_of(isop ? Float64 : eltype(A), opts, args)
    isop && linop_opts(o)
    m, n = size(A); mn = max(1, min(m, n))
    c = ctx(); bind_stream!(c)
    rows = ROCVector{Int64}(undef, mn); cols = ROCVector{Int64}(undef, mn); k = Ref{Int64}(0)
    co = LraOpts(o; seed)
    if isop
        GC.@preserve A check(ccall((:lra_curfact_op, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Any, Cint, Ref{LraOpts},
                     Ptr{Int64}, Ptr{Int64}, Ref{Int64}),
                    c.h, m, n, APPLY[], A, A.hermitian, co, dptr(rows), dptr(cols), k),
                    "lra_curfact_op")
    else
        Ad = f64(A)
        check(ccall((:lra_curfact, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ref{LraOpts}, Ptr{Int64},
                     Ptr{Int64}, Ref{Int64}),
                    c.h, m, n, dptr(Ad), ld(Ad), co, dptr(rows), dptr(cols), k), "lra_curfact")
    end
    CURPackedU(Array(rows[1:k[]]), Array(cols[1:k[]]))
end

# ---------------------------------------------------------------------------------------
# snorm / snormdiff (src/snorm.jl:14-53)

function snorm(A::Union{ROCMatrix,ROCLinOp}, opts=nothing; seed::Integer=0, args...)
    o = opts_of(Float64, opts, args)
    c = ctx(); bind_stream!(c)
    r = Ref{Float64}(0.0)
    m, n = size(A)
    if A isa ROCLinOp
        GC.@preserve A check(ccall((:lra_snorm_op, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Any, Cint, Ref{LraOpts},
                     Ptr{Float64}, Ref{Float64}),
                    c.h, m, n, APPLY[], A, A.hermitian, LraOpts(o; seed), C_NULL, r),
                    "lra_snorm_op")
    else
        Ad = f64(A)
        check(ccall((:lra_snorm, liblra), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Cint, Ref{LraOpts},
                     Ptr{Float64}, Ref{Float64}),
                    c.h, m, n, dptr(Ad), ld(Ad), ishermitian(Ad), LraOpts(o; seed), C_NULL, r),
                    "lra_snorm")
    end
    r[]
end

# snormdiff(A, V) for an ID of A, applied matrix-free (A - A[:, sk] [I T] P')
function snormdiff(A::ROCMatrix, V::IDPackedV, opts=nothing; seed::Integer=0, args...)
    o = opts_of(Float64, opts, args)
    Ad = f64(A); m, n = size(Ad); k = length(V.sk)
    c = ctx(); bind_stream!(c)
    sk = ROCArray(V.sk); rd = ROCArray(V.rd)
    T = k > 0 && n > k ? ROCArray(V.T) : ROCMatrix{Float64}(undef, max(k, 1), 1)
    r = Ref{Float64}(0.0)
    check(ccall((:lra_snormdiff_id, liblra), Cint,
                (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Int64}, Ptr{Int64},
                 Ptr{Float64}, Int64, Ref{LraOpts}, Ptr{Float64}, Ref{Float64}),
                c.h, m, n, dptr(Ad), ld(Ad), k, dptr(sk), dptr(rd), dptr(T), ld(T),
                LraOpts(o; seed), C_NULL, r), "lra_snormdiff_id")
    r[]
end

function snormdiff(A::ROCMatrix, B::ROCMatrix, opts=nothing; seed::Integer=0, args...)
    o = opts_of(Float64, opts, args)
    Ad = f64(A); Bd = f64(B); m, n = size(Ad)
    c = ctx(); bind_stream!(c)
    r = Ref{Float64}(0.0)
    check(ccall((:lra_snormdiff, liblra), Cint,
                (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cint,
                 Ref{LraOpts}, Ptr{Float64}, Ref{Float64}),
                c.h, m, n, dptr(Ad), ld(Ad), dptr(Bd), ld(Bd),
                ishermitian(Ad) && ishermitian(Bd), LraOpts(o; seed), C_NULL, r),
                "lra_snormdiff")
    r[]
end

end # module
