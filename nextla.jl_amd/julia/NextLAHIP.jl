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

end # module
