"""
    NextLAHIP

ROCm-array methods for NextLA.jl's recursive triangular solve / multiply,
calling libnlahip (the gfx950 HIP kernels behind include/nla.h) with `ccall`.

Loaded next to NextLA and AMDGPU, it adds methods of

    NextLA.unified_rectrxm!(side, uplo, transpose, alpha, func, A, B)   src/rectrxm.jl:43
    NextLA.GEMM_ADD!(A, B, C), NextLA.GEMM_SUB!(A, B, C)               src/matmul.jl:69,76
    NextLA.LeftLowerTRSM!(A, B) ... NextLA.RightUpperTRMM!(A, B)       src/trsm.jl:128-150, src/trmm.jl:332-393

for `ROCMatrix{T}` (and unit-stride views of one) with `T <: Union{Float32, Float64}`,
with the reference's argument order and return values.  Work is queued on
`AMDGPU.stream()`; call `AMDGPU.synchronize()` (what
`KernelAbstractions.synchronize` does for a ROCBackend) before reading B.

It also exposes the forms the reference does not have: `batched_rectrxm!`
(many systems in one batched launch sequence), `rectrxm_mixed!` (FP64-accurate
TRSM by low-precision correction solves and FP64 refinement) and
`batched_rectrxm_mixed!` (the same for many systems at once), and methods of
NextLA's `trsm` / `trmm` wrappers that honour `transa` and `diag`, and
`getrf2!` (the recursive LU whose panel solve is the unit-lower TRSM),
`lauum!` (U·U^T / L^T·L on the TRMM), `larfb!` (tile QR's block reflector,
its triangular multiplies on the TRMM), `pamm!` / `parfb!` (tile QR's
triangular-pentagonal reflectors, LAPACK's xTPRFB), `residual!` (the FP64-accurate
triangular residual on the int8 MFMA), `unified_rectrxm_noalloc!` (the
drop-in without device allocation) and `Server` / `submit!` / `step!` /
`drain!` (the continuous-batching request server).

Differences from the reference, by design: arguments are validated before any
device work (`ArgumentError`, B untouched), lower-case characters are
accepted, and `alpha == 0` zeroes B without reading it (BLAS semantics).
"""
module NextLAHIP

using AMDGPU
using NextLA

const libnlahip = get(ENV, "NEXTLA_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libnlahip.so"))

const DeviceMat{T} = Union{ROCMatrix{T}, SubArray{T, 2, <:ROCMatrix{T}}}

dtype_code(::Type{Float32}) = Int32(0)
dtype_code(::Type{Float64}) = Int32(1)

# Column-major pointer and leading dimension; only unit row stride is accepted.
function ptr_ld(A::DeviceMat)
    stride(A, 1) == 1 || throw(ArgumentError("matrix views need unit stride in dimension 1"))
    return pointer(A), Int64(max(stride(A, 2), size(A, 1), 1))
end

function check(st::Int32, what)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:nla_last_error, libnlahip), Cstring, ()))
    st < 0 ? throw(ArgumentError("$what: $msg")) : error("$what: $msg")
end

hip_stream() = AMDGPU.stream().stream

function NextLA.unified_rectrxm!(side::Char, uplo::Char, transpose::Char, alpha::Number, func::Char,
                                 A::DeviceMat{T}, B::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    n = size(A, 1)
    size(A, 2) == n || throw(DimensionMismatch("A must be square"))
    m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
    pa, lda = ptr_ld(A)
    pb, ldb = ptr_ld(B)
    st = GC.@preserve A B ccall((:nla_rectrxm, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Cchar, Int32, Float64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(side), Cchar(uplo), Cchar(transpose), Cchar(func), dtype_code(T), Float64(alpha), Int64(n), Int64(m),
        pa, lda, pb, ldb, hip_stream())
    check(st, "unified_rectrxm!")
    return B
end

# GEMM_ADD!(A, B, C): C += A*B
function NextLA.GEMM_ADD!(A::DeviceMat{T}, B::DeviceMat{T}, C::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    pa, lda = ptr_ld(A); pb, ldb = ptr_ld(B); pc, ldc = ptr_ld(C)
    st = GC.@preserve A B C ccall((:nla_gemm_add, libnlahip), Int32,
        (Int32, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        dtype_code(T), size(A, 1), size(B, 1), size(B, 2), pa, lda, pb, ldb, pc, ldc, hip_stream())
    check(st, "GEMM_ADD!")
    return C
end

# GEMM_SUB!(A, B, C): A -= B*C
function NextLA.GEMM_SUB!(A::DeviceMat{T}, B::DeviceMat{T}, C::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    pa, lda = ptr_ld(A); pb, ldb = ptr_ld(B); pc, ldc = ptr_ld(C)
    st = GC.@preserve A B C ccall((:nla_gemm_sub, libnlahip), Int32,
        (Int32, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        dtype_code(T), size(A, 1), size(C, 1), size(A, 2), pa, lda, pb, ldb, pc, ldc, hip_stream())
    check(st, "GEMM_SUB!")
    return A
end

for (f, func, side, uplo) in ((:LeftLowerTRSM!, 'S', 'L', 'L'), (:LeftUpperTRSM!, 'S', 'L', 'U'),
                              (:RightLowerTRSM!, 'S', 'R', 'L'), (:RightUpperTRSM!, 'S', 'R', 'U'),
                              (:LeftLowerTRMM!, 'M', 'L', 'L'), (:LeftUpperTRMM!, 'M', 'L', 'U'),
                              (:RightLowerTRMM!, 'M', 'R', 'L'), (:RightUpperTRMM!, 'M', 'R', 'U'))
    @eval function NextLA.$f(A::DeviceMat{T}, B::DeviceMat{T}) where {T <: Union{Float32, Float64}}
        NextLA.unified_rectrxm!($side, $uplo, 'N', one(T), $func, A, B)
    end
end

# trsm / trmm wrappers (src/trsm.jl:186, src/trmm.jl:430): transa and diag honoured
for (f, sym) in ((:trsm, :nla_trsm), (:trmm, :nla_trmm))
    @eval function NextLA.$f(side::Char, uplo::Char, transa::Char, diag::Char, A::DeviceMat{T}, B::DeviceMat{T},
                             alpha=one(T)) where {T <: Union{Float32, Float64}}
        n = size(A, 1)
        m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
        pa, lda = ptr_ld(A)
        pb, ldb = ptr_ld(B)
        st = GC.@preserve A B ccall(($(QuoteNode(sym)), libnlahip), Int32,
            (Cchar, Cchar, Cchar, Cchar, Int32, Float64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
             Ptr{Cvoid}),
            Cchar(side), Cchar(uplo), Cchar(transa), Cchar(diag), dtype_code(T), Float64(alpha), Int64(n),
            Int64(m), pa, lda, pb, ldb, hip_stream())
        check(st, $(string(f)))
        return B
    end
end

# getrf2!(A, ipiv, info) (src/lu.jl:185-300): recursive LU with partial pivoting on the device.
# Returns (A, ipiv, info) with info read back (0 or the first zero pivot).
function NextLA.getrf2!(A::DeviceMat{T}, ipiv::AbstractVector{Int}, info::Int) where {T <: Union{Float32, Float64}}
    m, n = size(A)
    pa, lda = ptr_ld(A)
    dpiv = AMDGPU.ROCVector{Int64}(undef, max(min(m, n), 1))
    dinfo = AMDGPU.zeros(Int64, 1)
    st = GC.@preserve A dpiv dinfo ccall((:nla_getrf2, libnlahip), Int32,
        (Int32, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}),
        dtype_code(T), Int64(m), Int64(n), pa, lda, pointer(dpiv), pointer(dinfo), hip_stream())
    check(st, "getrf2!")
    AMDGPU.synchronize()
    inf = Int(Array(dinfo)[1])
    # -1: a multi-work-group leaf timed out at its grid barrier (factors not valid)
    inf < 0 && error("getrf2!: a multi-work-group leaf timed out waiting at its grid barrier")
    k = min(m, n)
    k > 0 && copyto!(ipiv, 1, Array(dpiv), 1, k)
    return A, ipiv, inf
end

# lauum!(uplo, n, A, ib) (src/lauum.jl:52): A := U*U^T or L^T*L on the device; ib is ignored.
function NextLA.lauum!(uplo::Char, n::Integer, A::DeviceMat{T}, ib::Integer) where {T <: Union{Float32, Float64}}
    uplo in ('U', 'L') || throw(ArgumentError("uplo must be 'U' or 'L', got '$uplo'"))
    n >= 0 || throw(ArgumentError("n must be non-negative, got $n"))
    pa, lda = ptr_ld(A)
    bytes = ccall((:nla_lauum_workspace_bytes, libnlahip), Int64, (Int32, Int64), dtype_code(T), Int64(n))
    buf, ws = workspace(max(bytes, 1))
    st = GC.@preserve A buf ccall((:nla_lauum, libnlahip), Int32,
        (Cchar, Int32, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(uplo), dtype_code(T), Int64(n), pa, lda, ws, bytes, hip_stream())
    check(st, "lauum!")
    AMDGPU.synchronize()  # the workspace is released when this returns
    return A
end

# larfb!(side, trans, direct, storev, V, T, C) (src/larfb.jl:166): forward columnwise block
# reflector applied on the device; its triangular multiplies are the library's TRMM.
function NextLA.larfb!(side::Char, trans::Char, direct::Char, storev::Char, V::DeviceMat{T}, Tm::DeviceMat{T},
                       C::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    m, n = size(C)
    k = size(Tm, 1)
    pv, ldv = ptr_ld(V); pt, ldt = ptr_ld(Tm); pc, ldc = ptr_ld(C)
    wr = side in ('L', 'l') ? n : m
    work = AMDGPU.ROCMatrix{T}(undef, max(wr, 1), max(k, 1))
    st = GC.@preserve V Tm C work ccall((:nla_larfb, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Cchar, Int32, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid},
         Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(side), Cchar(trans), Cchar(direct), Cchar(storev), dtype_code(T), Int64(m), Int64(n), Int64(k),
        pv, ldv, pt, ldt, pc, ldc, pointer(work), Int64(max(wr, 1)), hip_stream())
    check(st, "larfb!")
    AMDGPU.synchronize()  # work is released when this returns
    return C
end

# pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W) (src/pamm.jl:50): the pentagonal
# products of tile QR; V's l x l triangle through the device TRMM.  Returns W ('W') or A2 ('A').
function NextLA.pamm!(op::Char, side::Char, storev::Char, direct::Char, m::Integer, n::Integer, k::Integer,
                      l::Integer, A1::DeviceMat{T}, A2::DeviceMat{T}, V::DeviceMat{T},
                      W::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    p1, l1 = ptr_ld(A1); p2, l2 = ptr_ld(A2); pv, ldv = ptr_ld(V); pw, ldw = ptr_ld(W)
    st = GC.@preserve A1 A2 V W ccall((:nla_pamm, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Cchar, Int32, Int64, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
         Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(op), Cchar(side), Cchar(storev), Cchar(direct), dtype_code(T), Int64(m), Int64(n), Int64(k), Int64(l),
        p1, l1, p2, l2, pv, ldv, pw, ldw, hip_stream())
    check(st, "pamm!")
    return op in ('W', 'w') ? W : A2
end

# parfb!(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work) (src/parfb.jl:46):
# LAPACK's xTPRFB on the device.
function NextLA.parfb!(side::Char, trans::Char, direct::Char, storev::Char, m1::Integer, n1::Integer,
                       m2::Integer, n2::Integer, k::Integer, l::Integer, A1::DeviceMat{T}, A2::DeviceMat{T},
                       V::DeviceMat{T}, Tm::DeviceMat{T}, work::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    p1, l1 = ptr_ld(A1); p2, l2 = ptr_ld(A2); pv, ldv = ptr_ld(V); pt, ldt = ptr_ld(Tm); pw, ldw = ptr_ld(work)
    st = GC.@preserve A1 A2 V Tm work ccall((:nla_parfb, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Cchar, Int32, Int64, Int64, Int64, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid},
         Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(side), Cchar(trans), Cchar(direct), Cchar(storev), dtype_code(T), Int64(m1), Int64(n1), Int64(m2),
        Int64(n2), Int64(k), Int64(l), p1, l1, p2, l2, pv, ldv, pt, ldt, pw, ldw, hip_stream())
    check(st, "parfb!")
    return A1, A2
end

# ---- forms without a reference counterpart ---------------------------------

struct NlaBatchItem
    n::Int64
    m::Int64
    A::Ptr{Cvoid}
    lda::Int64
    B::Ptr{Cvoid}
    ldb::Int64
    alpha::Float64
end

function batch_items(side::Char, As, Bs, alphas)
    length(As) == length(Bs) || throw(DimensionMismatch("As and Bs differ in length"))
    items = Vector{NlaBatchItem}(undef, length(As))
    for (i, (A, B)) in enumerate(zip(As, Bs))
        pa, lda = ptr_ld(A)
        pb, ldb = ptr_ld(B)
        m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
        items[i] = NlaBatchItem(size(A, 1), m, pa, lda, pb, ldb, Float64(alphas isa Number ? alphas : alphas[i]))
    end
    return items
end

# 256-byte aligned device workspace of at least `bytes` bytes
function workspace(bytes::Integer)
    buf = AMDGPU.ROCVector{UInt8}(undef, bytes + 256)
    p = pointer(buf)
    off = (256 - (UInt(p) % 256)) % 256
    return buf, p + off
end

"""
    batched_rectrxm!(side, uplo, transpose, func, As, Bs; alphas=1, use_graph=false)

`unified_rectrxm!` applied to every (A_i, B_i) pair in one persistent batched
launch sequence (nla_batch_*).  Returns Bs; work is queued on AMDGPU.stream().
"""
function batched_rectrxm!(side::Char, uplo::Char, transpose::Char, func::Char, As::AbstractVector,
                          Bs::AbstractVector; alphas=1, use_graph::Bool=false)
    T = eltype(first(As))
    items = batch_items(side, As, Bs, alphas)
    GC.@preserve As Bs items begin
        bytes = ccall((:nla_batch_workspace_bytes, libnlahip), Int64, (Cchar, Int32, Ptr{NlaBatchItem}, Int64),
                      Cchar(func), dtype_code(T), items, length(items))
        bytes >= 0 || throw(ArgumentError("invalid batch description"))
        buf, ws = workspace(bytes)
        plan = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve buf begin
            check(ccall((:nla_batch_plan_create, libnlahip), Int32,
                        (Cchar, Cchar, Cchar, Cchar, Int32, Ptr{NlaBatchItem}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid},
                         Ptr{Ptr{Cvoid}}),
                        Cchar(side), Cchar(uplo), Cchar(transpose), Cchar(func), dtype_code(T), items, length(items),
                        ws, bytes, hip_stream(), plan), "nla_batch_plan_create")
            st = ccall((:nla_batch_plan_execute, libnlahip), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int32), plan[],
                       hip_stream(), Int32(use_graph))
            AMDGPU.synchronize()
            ccall((:nla_batch_plan_destroy, libnlahip), Int32, (Ptr{Cvoid},), plan[])
            check(st, "nla_batch_plan_execute")
        end
    end
    return Bs
end

"""
    rectrxm_mixed!(side, uplo, transpose, alpha, A, B; precision=:auto, tol=1e-14, max_iter=20)

FP64-accurate TRSM by bf16 / fp16 / fp32 correction solves (`:auto` picks each
correction's precision by predicted cost) and iterative refinement with
FP64-accurate residuals (int8 modular products or FP64; nla_rectrxm_mixed).
Returns (B, steps, per-system backward error).
"""
function rectrxm_mixed!(side::Char, uplo::Char, transpose::Char, alpha::Number, A::DeviceMat{Float64},
                        B::DeviceMat{Float64}; precision::Symbol=:auto, tol::Real=1e-14, max_iter::Integer=20)
    prec = Dict(:bf16 => Int32(0), :fp16 => Int32(1), :fp32 => Int32(2), :auto => Int32(3),
                :fp64 => Int32(4), :fp16x2 => Int32(5))[precision]
    n = size(A, 1)
    m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
    pa, lda = ptr_ld(A)
    pb, ldb = ptr_ld(B)
    bytes = ccall((:nla_mixed_workspace_bytes, libnlahip), Int64, (Cchar, Int64, Int64), Cchar(side), n, m)
    buf, ws = workspace(bytes)
    # nla_mixed_info: int32 iterations, int32 precision, double backward_error, double history[31],
    # double backward_error_frobenius, int32 residual (+4 bytes of padding to the struct's 8-byte alignment)
    info = zeros(UInt8, 4 + 4 + 8 + 8 * 31 + 8 + 8)
    st = GC.@preserve A B buf info ccall((:nla_rectrxm_mixed, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Int32, Float64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Float64, Int32,
         Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{UInt8}),
        Cchar(side), Cchar(uplo), Cchar(transpose), prec, Float64(alpha), Int64(n), Int64(m), pa, lda, pb, ldb,
        Float64(tol), Int32(max_iter), ws, bytes, hip_stream(), info)
    check(st, "rectrxm_mixed!")
    steps = reinterpret(Int32, info[1:4])[1]
    bwd = reinterpret(Float64, info[9:16])[1]
    return B, steps, bwd
end

# ---- request server (nla_serve_*) ----------------------------------------------------------

struct NlaServeRequest
    side::Cchar; uplo::Cchar; trans::Cchar; func::Cchar
    dtype::Int32
    alpha::Float64
    n::Int64; m::Int64
    A::Ptr{Cvoid}; lda::Int64
    B::Ptr{Cvoid}; ldb::Int64
    a_key::Int64
end

struct NlaServeDone
    ticket::Int64
    status::Int32
    iterations::Int32
    backward_error::Float64
end

"""
    Server(; budget_bytes=64 << 30, max_inflight=4, max_batch=512, refined=false, workers=4)

The continuous-batching server core of libnlahip (nla_serve_*): `submit!` one
`unified_rectrxm!` call at a time (A and B already on the device), `step!` to
retire and launch batches and collect `(ticket, status, iterations,
backward_error)` completions, `drain!` to finish everything.  With
`refined=true`, FP64 TRSMs are served by batched refinement to backward error
1e-14.  The caller keeps A and B alive until their ticket completes.
"""
mutable struct Server
    handle::Ptr{Cvoid}
    function Server(; budget_bytes::Integer=64 << 30, max_inflight::Integer=4, max_batch::Integer=512,
                    refined::Bool=false, workers::Integer=4, use_graph::Bool=false)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:nla_serve_create, libnlahip), Int32, (Int64, Int32, Int32, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
                    Int64(budget_bytes), Int32(max_inflight), Int32(max_batch), Int32(refined ? 1 : 0),
                    Int32(refined ? workers : 0), Int32(use_graph ? 1 : 0), h), "nla_serve_create")
        srv = new(h[])
        finalizer(s -> s.handle == C_NULL || ccall((:nla_serve_destroy, libnlahip), Int32, (Ptr{Cvoid},), s.handle),
                  srv)
        return srv
    end
end

function submit!(srv::Server, side::Char, uplo::Char, transpose::Char, alpha::Number, func::Char,
                 A::DeviceMat{T}, B::DeviceMat{T}; a_key::Integer=0) where {T <: Union{Float32, Float64}}
    n = size(A, 1)
    m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
    pa, lda = ptr_ld(A)
    pb, ldb = ptr_ld(B)
    req = Ref(NlaServeRequest(Cchar(side), Cchar(uplo), Cchar(transpose), Cchar(func), dtype_code(T),
                              Float64(alpha), Int64(n), Int64(m), pa, lda, pb, ldb, Int64(a_key)))
    ticket = Ref{Int64}(0)
    check(ccall((:nla_serve_submit, libnlahip), Int32, (Ptr{Cvoid}, Ptr{NlaServeRequest}, Ptr{Cvoid}, Ptr{Int64}),
                srv.handle, req, hip_stream(), ticket), "nla_serve_submit")
    return ticket[]
end

function step!(srv::Server; capacity::Integer=1024)
    done = Vector{NlaServeDone}(undef, capacity)
    nd = Ref{Int64}(0)
    check(ccall((:nla_serve_step, libnlahip), Int32, (Ptr{Cvoid}, Ptr{NlaServeDone}, Int64, Ptr{Int64}),
                srv.handle, done, Int64(capacity), nd), "nla_serve_step")
    return done[1:nd[]]
end

drain!(srv::Server) = check(ccall((:nla_serve_drain, libnlahip), Int32, (Ptr{Cvoid},), srv.handle), "nla_serve_drain")

"""
    residual!(side, uplo, transpose, alpha, A, X, B, R; method=:int8)

R = alpha*B - op(A)*X (side 'L') or alpha*B - X*op(A) (side 'R') to FP64
accuracy (nla_residual).  `:int8`: Ozaki scheme II on the int8 MFMA; X is
overwritten by its value rounded to >= 55 bits of each system's largest entry,
and R is the residual of that X.  `:fp64`: the FP64 MFMA product.  Returns R.
"""
function residual!(side::Char, uplo::Char, transpose::Char, alpha::Number, A::DeviceMat{Float64},
                   X::DeviceMat{Float64}, B::DeviceMat{Float64}, R::DeviceMat{Float64}; method::Symbol=:int8)
    meth = Dict(:fp64 => Int32(0), :int8 => Int32(1))[method]
    n = size(A, 1)
    m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
    pa, lda = ptr_ld(A)
    px, ldx = ptr_ld(X)
    pb, ldb = ptr_ld(B)
    pr, ldr = ptr_ld(R)
    bytes = ccall((:nla_residual_workspace_bytes, libnlahip), Int64, (Cchar, Int32, Int64, Int64),
                  Cchar(side), meth, n, m)
    bytes >= 0 || throw(ArgumentError("residual!: invalid shape"))
    buf, ws = workspace(bytes)
    st = GC.@preserve A X B R buf ccall((:nla_residual, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Int32, Float64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
         Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(side), Cchar(uplo), Cchar(transpose), meth, Float64(alpha), Int64(n), Int64(m), pa, lda, px, ldx,
        pb, ldb, pr, ldr, ws, bytes, hip_stream())
    check(st, "residual!")
    return R
end

"""
    unified_rectrxm_noalloc!(side, uplo, transpose, alpha, func, A, B)

`unified_rectrxm!` through nla_rectrxm_noalloc: no device allocation (the
TRSM's scratch is a static device array of the library; FP64 n <= 32768,
FP32 n <= 65536), for callers that cannot allocate.  Calls that overlap on one
device race on that array.
"""
function unified_rectrxm_noalloc!(side::Char, uplo::Char, transpose::Char, alpha::Number, func::Char,
                                  A::DeviceMat{T}, B::DeviceMat{T}) where {T <: Union{Float32, Float64}}
    n = size(A, 1)
    size(A, 2) == n || throw(DimensionMismatch("A must be square"))
    m = side in ('L', 'l') ? size(B, 2) : size(B, 1)
    pa, lda = ptr_ld(A)
    pb, ldb = ptr_ld(B)
    st = GC.@preserve A B ccall((:nla_rectrxm_noalloc, libnlahip), Int32,
        (Cchar, Cchar, Cchar, Cchar, Int32, Float64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
        Cchar(side), Cchar(uplo), Cchar(transpose), Cchar(func), dtype_code(T), Float64(alpha), Int64(n), Int64(m),
        pa, lda, pb, ldb, hip_stream())
    check(st, "unified_rectrxm_noalloc!")
    return B
end

"""
    batched_rectrxm_mixed!(side, uplo, transpose, As, Bs; alphas=1, tol=1e-14, max_iter=10)

FP64-accurate TRSM of every (A_i, B_i) pair by batched fp32 correction solves
and FP64 refinement (nla_batch_mixed_*).  Returns (Bs, backward errors, steps).
"""
function batched_rectrxm_mixed!(side::Char, uplo::Char, transpose::Char, As::AbstractVector, Bs::AbstractVector;
                                alphas=1, tol::Real=1e-14, max_iter::Integer=10)
    items = batch_items(side, As, Bs, alphas)
    count = length(items)
    bwd = zeros(Float64, count)
    its = zeros(Int32, count)
    GC.@preserve As Bs items begin
        bytes = ccall((:nla_batch_mixed_workspace_bytes, libnlahip), Int64, (Cchar, Ptr{NlaBatchItem}, Int64),
                      Cchar(side), items, count)
        bytes >= 0 || throw(ArgumentError("invalid batch description"))
        buf, ws = workspace(bytes)
        plan = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve buf begin
            check(ccall((:nla_batch_mixed_plan_create, libnlahip), Int32,
                        (Cchar, Cchar, Cchar, Ptr{NlaBatchItem}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                        Cchar(side), Cchar(uplo), Cchar(transpose), items, count, ws, bytes, hip_stream(), plan),
                  "nla_batch_mixed_plan_create")
            st = ccall((:nla_batch_mixed_plan_execute, libnlahip), Int32,
                       (Ptr{Cvoid}, Float64, Int32, Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Int32}),
                       plan[], Float64(tol), Int32(max_iter), hip_stream(), Int32(1), bwd, its)
            ccall((:nla_batch_mixed_plan_destroy, libnlahip), Int32, (Ptr{Cvoid},), plan[])
            check(st, "nla_batch_mixed_plan_execute")
        end
    end
    return Bs, bwd, its
end

end # module
