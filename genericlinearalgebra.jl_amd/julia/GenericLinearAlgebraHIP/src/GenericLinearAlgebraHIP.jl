"""
    GenericLinearAlgebraHIP

GenericLinearAlgebra.jl's blocked dense factorisations on an AMD MI355X
(gfx950): methods of the package's own functions (`cholUnblocked!`,
`cholBlocked!`, `cholRecursive!`, `qrUnblocked!`, `qrBlocked!`,
`rankUpdate!`, `QR2` indexing, the compact-WY `lmul!`) for a device-resident
`HIPMatrix`, each one `ccall` into `libgla_hip.so` (the C ABI declared in
`include/gla_hip.h`).  Every C signature used here is in the table
`SIGNATURES`; all calls go through it.

```julia
using GenericLinearAlgebra, GenericLinearAlgebraHIP
A = HIPMatrix(B'B + n*I)                       # upload
L = cholRecursive!(copy(A), Val{:L})           # LowerTriangular{Float64,HIPMatrix{Float64}}
F = qrBlocked!(copy(M), 128)                   # QR2 with device factors
Array(F[Tuple{:R}].data)                       # download
```
"""
module GenericLinearAlgebraHIP

using LinearAlgebra
using Libdl
import ChainRulesCore
using ChainRulesCore: NoTangent, unthunk
import GenericLinearAlgebra
import GenericLinearAlgebra: cholUnblocked!, cholBlocked!, cholRecursive!, qrUnblocked!,
                             qrBlocked!, rankUpdate!, QR2

export HIPArray, HIPMatrix, HIPVector, hipzeros, hipsimilar, hipview, HIPHouseholderBlock,
       chol_solve!, qr_solve!, mixed_solve, BatchPlan, chol_solve_batched!, chol_adjoint,
       chol_adjoint_batched!, qr_adjoint, chol_lower, qr_R, set_tuning!, get_tuning

# ---------------------------------------------------------------- libraries --
# libgla_hip.so as built in-tree (genericlinearalgebra.jl_amd/lib), or the
# path in ENV["GLA_HIP_LIB"] at precompile time.
const libgla = get(ENV, "GLA_HIP_LIB",
                   normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libgla_hip.so")))
const libhip = "libamdhip64"

const GLA_OK = Cint(0)
const GLA_ERR_HIP = Cint(-1000)
const GLA_ERR_UNSUPPORTED = Cint(-1001)
const GLA_ERR_TIMEOUT = Cint(-1003)

# C signature of every entry point of include/gla_hip.h this module calls:
# name => (return type, argument types).  gla_handle_t and gla_batch_plan_t
# are Ptr{Cvoid}; device arrays are Ptr{T}; host out-parameters are Ref{T}.
const SIGNATURES = Dict{Symbol,Tuple{Any,Tuple}}(
    :gla_handle_create => (Cint, (Ref{Ptr{Cvoid}},)),
    :gla_handle_destroy => (Cint, (Ptr{Cvoid},)),
    :gla_last_error_string => (Cstring, ()),
    :gla_version => (Cstring, ()),
    :gla_set_tuning => (Cint, (Ptr{Cvoid}, Cstring, Int64)),
    :gla_get_tuning => (Cint, (Ptr{Cvoid}, Cstring, Ref{Int64})),
    :gla_chol_unblocked_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_unblocked_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_blocked_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_blocked_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_recursive_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_recursive_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_rank_update_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_rank_update_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_trsm_rlt_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_trsm_rlt_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_qr_unblocked_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Cvoid})),
    :gla_qr_unblocked_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Cvoid})),
    :gla_qr_blocked_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_qr_blocked_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_qr_tblock_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_qr_tblock_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_wy_apply_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_wy_apply_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_lmul_unit_lower_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Ptr{Cvoid})),
    :gla_lmul_unit_lower_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32, Ptr{Cvoid})),
    :gla_gemm_d => (Cint, (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_gemm_s => (Cint, (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_batch_plan_create => (Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}, Int32, Int32, Ref{Ptr{Cvoid}})),
    :gla_batch_plan_destroy => (Cint, (Ptr{Cvoid},)),
    :gla_chol_solve_batched_d => (Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Float64}}, Ptr{Int32}, Ptr{Ptr{Float64}}, Ptr{Int32}, Ptr{Cvoid})),
    :gla_chol_solve_batched_s => (Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Float32}}, Ptr{Int32}, Ptr{Ptr{Float32}}, Ptr{Int32}, Ptr{Cvoid})),
    :gla_chol_solve_mixed_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Int32, Ref{Int32}, Ref{Float64}, Ref{Int64}, Ref{Cchar}, Ptr{Cvoid})),
    :gla_gemm_lowp_nt => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_ldlt_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Int64, Ptr{Cvoid})),
    :gla_ldlt_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Int64, Ptr{Cvoid})),
    :gla_ldlt_solve_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_ldlt_solve_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_ldlt_lmul_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_ldlt_lmul_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_symtri_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Cvoid})),
    :gla_symtri_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Cvoid})),
    :gla_bidiag_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid})),
    :gla_bidiag_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid})),
    :gla_chol_unblocked_z => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF64}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_unblocked_c => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF32}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_blocked_z => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF64}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_blocked_c => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF32}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_recursive_z => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF64}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_recursive_c => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{ComplexF32}, Int64, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_rank_update_z => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Float64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Int64, Ptr{Cvoid})),
    :gla_rank_update_c => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Float32, Ptr{ComplexF32}, Int64, Ptr{ComplexF32}, Int64, Ptr{Cvoid})),
    :gla_qr_unblocked_z => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Ptr{Cvoid})),
    :gla_qr_unblocked_c => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF32}, Int64, Ptr{ComplexF32}, Ptr{Cvoid})),
    :gla_qr_blocked_z => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF64}, Int64, Int64, Ptr{ComplexF64}, Ptr{Cvoid})),
    :gla_qr_blocked_c => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF32}, Int64, Int64, Ptr{ComplexF32}, Ptr{Cvoid})),
    :gla_qr_tblock_z => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Ptr{Cvoid})),
    :gla_qr_tblock_c => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF32}, Int64, Ptr{ComplexF32}, Ptr{ComplexF32}, Int64, Ptr{Cvoid})),
    :gla_wy_apply_z => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{ComplexF64}, Int64, Int64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Int64, Ptr{Cvoid})),
    :gla_wy_apply_c => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{ComplexF32}, Int64, Int64, Ptr{ComplexF32}, Int64, Ptr{ComplexF32}, Int64, Ptr{Cvoid})),
    :gla_chol_solve_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_chol_solve_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_qr_solve_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_qr_solve_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Int64, Ref{Int64}, Ptr{Cvoid})),
    :gla_chol_adjoint_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_chol_adjoint_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
    :gla_chol_adjoint_batched_d => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Int32}, Ptr{Cvoid})),
    :gla_chol_adjoint_batched_s => (Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Int64, Int64, Ptr{Int32}, Ptr{Cvoid})),
    :gla_qr_adjoint_d => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid})),
    :gla_qr_adjoint_s => (Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid})),
)

"""
    @gla name args...

`ccall` of the libgla_hip.so entry point `name` with the signature that
`SIGNATURES` records for it (resolved when the calling method is compiled).
"""
macro gla(name, args...)
    name isa Symbol || error("@gla needs a literal entry-point name")
    haskey(SIGNATURES, name) || error("no signature recorded for $name")
    ret, types = SIGNATURES[name]
    return esc(:(ccall(($(QuoteNode(name)), $libgla), $ret, ($(types...),), $(args...))))
end

# --------------------------------------------------------- device memory ----
function _hipcheck(err::Integer, what::AbstractString)
    err == 0 && return nothing
    msg = unsafe_string(ccall((:hipGetErrorString, libhip), Cstring, (Cint,), err))
    error("$what: $msg")
end

const hipMemcpyHostToDevice = Cint(1)
const hipMemcpyDeviceToHost = Cint(2)
const hipMemcpyDeviceToDevice = Cint(3)

"""One hipMalloc allocation, freed by a finalizer."""
mutable struct HIPBuffer
    ptr::Ptr{Cvoid}
    bytes::Int
    function HIPBuffer(bytes::Integer)
        p = Ref{Ptr{Cvoid}}(C_NULL)
        if bytes > 0
            _hipcheck(ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, bytes),
                      "hipMalloc")
        end
        b = new(p[], Int(bytes))
        finalizer(b) do x
            if x.ptr != C_NULL
                ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), x.ptr)
                x.ptr = C_NULL
            end
        end
        return b
    end
end

"""
    HIPArray{T,N}

A column-major device array: element (i, j) sits `(i-1) + (j-1)*ld` elements
past `pointer(A)`; matrices may be views of a larger one (`hipview`), which
keep its allocation alive.  Scalar indexing is disabled: move data with
`HIPArray(::Array)`, `Array(::HIPArray)` and `copyto!`.
"""
struct HIPArray{T,N} <: AbstractArray{T,N}
    buf::HIPBuffer
    offset::Int          # elements from the start of buf
    dims::NTuple{N,Int}
    ld::Int              # leading dimension (elements between columns)
end
const HIPMatrix{T} = HIPArray{T,2}
const HIPVector{T} = HIPArray{T,1}

HIPArray{T}(::UndefInitializer, dims::Vararg{Integer,N}) where {T,N} =
    HIPArray{T,N}(HIPBuffer(sizeof(T) * prod(Int.(dims); init = 1)), 0, Int.(dims),
                  max(1, Int(dims[1])))
HIPArray{T}(::UndefInitializer, dims::NTuple{N,Integer}) where {T,N} = HIPArray{T}(undef, dims...)

Base.size(A::HIPArray) = A.dims
Base.pointer(A::HIPArray{T}) where {T} = Ptr{T}(A.buf.ptr) + A.offset * sizeof(T)
Base.unsafe_convert(::Type{Ptr{T}}, A::HIPArray{T}) where {T} = pointer(A)
Base.getindex(::HIPArray, ::Int...) =
    error("scalar indexing of a HIPArray is disabled; download it with Array(A)")
Base.setindex!(::HIPArray, v, ::Int...) =
    error("scalar indexing of a HIPArray is disabled; upload with copyto!(A, host)")
Base.show(io::IO, A::HIPArray{T,N}) where {T,N} =
    print(io, "HIPArray{$T,$N}(", join(size(A), "×"), ", ld=", A.ld, ")")
Base.show(io::IO, ::MIME"text/plain", A::HIPArray) = show(io, A)
ld(A::HIPArray) = A.ld
_rows(A::HIPArray) = size(A, 1)
_cols(A::HIPArray{T,1}) where {T} = 1
_cols(A::HIPArray{T,2}) where {T} = size(A, 2)
_cols(A::HIPArray{T,3}) where {T} = size(A, 2) * size(A, 3)

"""Device allocation of the given size, zero-filled."""
function hipzeros(::Type{T}, dims::Integer...) where {T}
    A = HIPArray{T}(undef, dims...)
    nbytes = sizeof(T) * prod(Int.(dims); init = 1)
    nbytes > 0 && _hipcheck(ccall((:hipMemset, libhip), Cint, (Ptr{Cvoid}, Cint, Csize_t),
                                  pointer(A), 0, nbytes), "hipMemset")
    return A
end
hipsimilar(A::HIPArray{T}) where {T} = HIPArray{T}(undef, size(A)...)
hipsimilar(A::HIPArray{T}, dims::Integer...) where {T} = HIPArray{T}(undef, dims...)
hipsimilar(A::HIPArray, ::Type{S}, dims::Integer...) where {S} = HIPArray{S}(undef, dims...)

function _copy2d!(dst::Ptr, dld::Integer, src::Ptr, sld::Integer, rows::Integer,
                  cols::Integer, ::Type{T}, kind::Cint) where {T}
    (rows == 0 || cols == 0) && return nothing
    _hipcheck(ccall((:hipMemcpy2D, libhip), Cint,
                    (Ptr{Cvoid}, Csize_t, Ptr{Cvoid}, Csize_t, Csize_t, Csize_t, Cint),
                    dst, dld * sizeof(T), src, sld * sizeof(T), rows * sizeof(T), cols, kind),
              "hipMemcpy2D")
end

function Base.copyto!(B::HIPArray{T}, A::Array{T}) where {T}
    size(B) == size(A) || throw(DimensionMismatch("sizes $(size(B)) and $(size(A)) differ"))
    GC.@preserve A _copy2d!(pointer(B), ld(B), pointer(A), max(1, size(A, 1)), _rows(B),
                            _cols(B), T, hipMemcpyHostToDevice)
    return B
end
function Base.copyto!(B::Array{T}, A::HIPArray{T}) where {T}
    size(B) == size(A) || throw(DimensionMismatch("sizes $(size(B)) and $(size(A)) differ"))
    GC.@preserve B _copy2d!(pointer(B), max(1, size(B, 1)), pointer(A), ld(A), _rows(A),
                            _cols(A), T, hipMemcpyDeviceToHost)
    return B
end
function Base.copyto!(B::HIPArray{T}, A::HIPArray{T}) where {T}
    size(B) == size(A) || throw(DimensionMismatch("sizes $(size(B)) and $(size(A)) differ"))
    _copy2d!(pointer(B), ld(B), pointer(A), ld(A), _rows(A), _cols(A), T,
             hipMemcpyDeviceToDevice)
    return B
end
HIPArray(A::Array{T,N}) where {T,N} = copyto!(HIPArray{T}(undef, size(A)...), A)
HIPMatrix(A::AbstractMatrix{T}) where {T} = HIPArray(Matrix{T}(A))
HIPVector(v::AbstractVector{T}) where {T} = HIPArray(Vector{T}(v))
Base.Array(A::HIPArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, size(A)), A)
Base.Matrix(A::HIPMatrix{T}) where {T} = Array(A)
"""Device-to-device copy (so `cholRecursive!(copy(A), Val{:L})` works on device data)."""
Base.copy(A::HIPArray{T}) where {T} = copyto!(hipsimilar(A), A)

"""A view of rows × cols of a device matrix, sharing its memory."""
function hipview(A::HIPMatrix{T}, rows::AbstractUnitRange{<:Integer},
                 cols::AbstractUnitRange{<:Integer}) where {T}
    checkbounds(Bool, 1:size(A, 1), rows) && checkbounds(Bool, 1:size(A, 2), cols) ||
        throw(BoundsError(A, (rows, cols)))
    off = A.offset + (first(rows) - 1) + (first(cols) - 1) * A.ld
    return HIPArray{T,2}(A.buf, off, (length(rows), length(cols)), A.ld)
end
function hipview(v::HIPVector{T}, r::AbstractUnitRange{<:Integer}) where {T}
    checkbounds(Bool, 1:length(v), r) || throw(BoundsError(v, r))
    return HIPArray{T,1}(v.buf, v.offset + first(r) - 1, (length(r),), max(1, length(r)))
end
Base.view(A::HIPMatrix, rows::AbstractUnitRange{<:Integer}, cols::AbstractUnitRange{<:Integer}) =
    hipview(A, rows, cols)
Base.view(v::HIPVector, r::AbstractUnitRange{<:Integer}) = hipview(v, r)

# ------------------------------------------------------------------ handle --
const HANDLES = Dict{Int,Ptr{Cvoid}}()
const HANDLE_LOCK = ReentrantLock()

struct GLAError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::GLAError) = print(io, "libgla_hip error ", e.code, ": ", e.msg)

function _check(rc::Integer, what::AbstractString)
    rc == GLA_OK && return nothing
    msg = unsafe_string(@gla gla_last_error_string())
    -100 < rc < 0 && throw(ArgumentError("$what: invalid argument $(-rc): $msg"))
    throw(GLAError(Cint(rc), "$what: $msg"))
end

"""
    with_handle(f)

Runs `f(h)` with the library handle of the current HIP device, holding the
lock that keeps one call at a time on it (a handle's workspace, barrier
counter and side stream are shared by its calls).
"""
function with_handle(f)
    lock(HANDLE_LOCK) do
        dev = Ref{Cint}(0)
        _hipcheck(ccall((:hipGetDevice, libhip), Cint, (Ref{Cint},), dev), "hipGetDevice")
        h = get(HANDLES, Int(dev[]), C_NULL)
        if h == C_NULL
            r = Ref{Ptr{Cvoid}}(C_NULL)
            _check(@gla(gla_handle_create, r), "gla_handle_create")
            h = r[]
            HANDLES[Int(dev[])] = h
        end
        return f(h)
    end
end

"""Sets a tuning knob of the current device's handle (see gla_set_tuning)."""
set_tuning!(key::AbstractString, value::Integer) =
    with_handle(h -> _check(@gla(gla_set_tuning, h, key, Int64(value)), "gla_set_tuning"))

"""Reads a tuning knob of the current device's handle (see gla_get_tuning)."""
function get_tuning(key::AbstractString)
    v = Ref{Int64}(0)
    with_handle(h -> _check(@gla(gla_get_tuning, h, key, v), "gla_get_tuning"))
    return v[]
end

const Floats = Union{Float32,Float64}
_uplo(::Type{Val{:L}}) = Cchar('L')
_uplo(::Type{Val{:U}}) = Cchar('U')

function _square(A)
    size(A, 1) == size(A, 2) ||
        throw(DimensionMismatch("matrix is not square: dimensions are $(size(A))"))
    return size(A, 1)
end

_domain(info::Integer) =
    info > 0 && throw(DomainError(info, "sqrt of a negative pivot at position $info"))

# ---------------------------------------------------------------- Cholesky --
# cholUnblocked!, cholBlocked!, cholRecursive! (src/cholesky.jl:3-55), the
# Val{:L} methods of the reference and the Val{:U} extension (U = Lᵀ in the
# upper triangle).  A DomainError at the first negative pivot, as the
# reference's sqrt throws it.
for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    unb = Symbol("gla_chol_unblocked_", sfx)
    blk = Symbol("gla_chol_blocked_", sfx)
    rec = Symbol("gla_chol_recursive_", sfx)
    rku = Symbol("gla_rank_update_", sfx)
    trs = Symbol("gla_trsm_rlt_", sfx)
    @eval begin
        function cholUnblocked!(A::HIPMatrix{$T}, uplo::Type{Val{U}}) where {U}
            n = _square(A)
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($unb, h, _uplo(uplo), n, pointer(A), ld(A), info,
                                         C_NULL), "cholUnblocked!"))
            _domain(info[])
            return A
        end
        function cholBlocked!(A::HIPMatrix{$T}, uplo::Type{Val{U}}, blocksize::Integer) where {U}
            n = _square(A)
            blocksize >= 1 || throw(ArgumentError("blocksize must be positive"))
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($blk, h, _uplo(uplo), n, pointer(A), ld(A),
                                         Int64(blocksize), info, C_NULL), "cholBlocked!"))
            _domain(info[])
            return A
        end
        function cholRecursive!(A::HIPMatrix{$T}, uplo::Type{Val{U}},
                                cutoff::Integer = 1) where {U}
            n = _square(A)
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($rec, h, _uplo(uplo), n, pointer(A), ld(A),
                                         Int64(cutoff), info, C_NULL), "cholRecursive!"))
            _domain(info[])
            return U === :L ? LowerTriangular(A) : UpperTriangular(A)
        end
        # rankUpdate!(C::Hermitian, A, α) — src/juliaBLAS.jl:53-63,89-112
        function rankUpdate!(C::Hermitian{$T,<:HIPMatrix{$T}}, A::HIPArray{$T}, α::Real)
            n = _square(C.data)
            size(A, 1) == n || throw(DimensionMismatch("first dimension of A has wrong size"))
            k = ndims(A) == 1 ? 1 : size(A, 2)
            with_handle(h -> _check(@gla($rku, h, Cchar(C.uplo), n, k, $T(α), pointer(A),
                                         ld(A), pointer(C.data), ld(C.data), C_NULL),
                                    "rankUpdate!"))
            return C
        end
        # rdiv!(B, LowerTriangular(L)') as called at src/cholesky.jl:28,48
        function LinearAlgebra.rdiv!(B::HIPMatrix{$T},
                                     Lt::Adjoint{$T,<:LowerTriangular{$T,<:HIPMatrix{$T}}})
            L = parent(Lt).data
            m, n = size(B)
            size(L) == (n, n) || throw(DimensionMismatch("L must be $n×$n"))
            with_handle(h -> _check(@gla($trs, h, m, n, pointer(L), ld(L), pointer(B), ld(B),
                                         C_NULL), "rdiv!"))
            return B
        end
    end
end

# Complex Hermitian A (ComplexF64/ComplexF32): the complex square root of the
# pivot, as the reference takes it, so no DomainError; her!/herk! updates.
for (T, R, sfx) in ((ComplexF64, Float64, "z"), (ComplexF32, Float32, "c"))
    unb = Symbol("gla_chol_unblocked_", sfx)
    blk = Symbol("gla_chol_blocked_", sfx)
    rec = Symbol("gla_chol_recursive_", sfx)
    rku = Symbol("gla_rank_update_", sfx)
    @eval begin
        function cholUnblocked!(A::HIPMatrix{$T}, uplo::Type{Val{U}}) where {U}
            n = _square(A)
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($unb, h, _uplo(uplo), n, pointer(A), ld(A), info,
                                         C_NULL), "cholUnblocked!"))
            return A
        end
        function cholBlocked!(A::HIPMatrix{$T}, uplo::Type{Val{U}}, blocksize::Integer) where {U}
            n = _square(A)
            blocksize >= 1 || throw(ArgumentError("blocksize must be positive"))
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($blk, h, _uplo(uplo), n, pointer(A), ld(A),
                                         Int64(blocksize), info, C_NULL), "cholBlocked!"))
            return A
        end
        function cholRecursive!(A::HIPMatrix{$T}, uplo::Type{Val{U}},
                                cutoff::Integer = 1) where {U}
            n = _square(A)
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($rec, h, _uplo(uplo), n, pointer(A), ld(A),
                                         Int64(cutoff), info, C_NULL), "cholRecursive!"))
            return U === :L ? LowerTriangular(A) : UpperTriangular(A)
        end
        function rankUpdate!(C::Hermitian{$T,<:HIPMatrix{$T}}, A::HIPArray{$T}, α::Real)
            n = _square(C.data)
            size(A, 1) == n || throw(DimensionMismatch("first dimension of A has wrong size"))
            k = ndims(A) == 1 ? 1 : size(A, 2)
            with_handle(h -> _check(@gla($rku, h, Cchar(C.uplo), n, k, $R(α), pointer(A),
                                         ld(A), pointer(C.data), ld(C.data), C_NULL),
                                    "rankUpdate!"))
            return C
        end
    end
end

# --------------------------------------------------------------------- QR --
"""
    HIPHouseholderBlock(V, T)

The compact-WY block I − V T Vᴴ of GenericLinearAlgebra's `HouseholderBlock`
(src/householder.jl:8-11) with device V and T (the reference's type takes
StridedMatrix fields only).  `H'`, `lmul!(H, A, M)`, `lmul!(H', A, M)`,
`H * A` and `H' * A` behave as the reference's (src/householder.jl:82-159).
"""
struct HIPHouseholderBlock{T}
    V::HIPMatrix{T}
    T::UpperTriangular{T,HIPMatrix{T}}
end
Base.eltype(::HIPHouseholderBlock{T}) where {T} = T
Base.adjoint(H::HIPHouseholderBlock{T}) where {T} = Adjoint{T,typeof(H)}(H)
Base.size(H::HIPHouseholderBlock) = (size(H.V, 1), size(H.V, 1))

for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    unb = Symbol("gla_qr_unblocked_", sfx)
    blk = Symbol("gla_qr_blocked_", sfx)
    tbl = Symbol("gla_qr_tblock_", sfx)
    wya = Symbol("gla_wy_apply_", sfx)
    lul = Symbol("gla_lmul_unit_lower_", sfx)
    qrs = Symbol("gla_qr_solve_", sfx)
    chs = Symbol("gla_chol_solve_", sfx)
    @eval begin
        # qrUnblocked!(A, τ) — src/qr.jl:86-111
        function qrUnblocked!(A::HIPMatrix{$T},
                              τ::HIPVector{$T} = hipzeros($T, min(size(A)...)))
            m, n = size(A)
            length(τ) >= min(m, n) || throw(DimensionMismatch("τ is too short"))
            with_handle(h -> _check(@gla($unb, h, m, n, pointer(A), ld(A), pointer(τ), C_NULL),
                                    "qrUnblocked!"))
            return QR2{$T,typeof(A),typeof(τ)}(A, τ)
        end
        # qrBlocked!(A, blocksize = 12, τ, work) — src/qr.jl:113-146; `work`
        # is left as the reference leaves it
        function qrBlocked!(A::HIPMatrix{$T}, blocksize::Integer = 12,
                            τ::HIPVector{$T} = hipzeros($T, min(size(A)...)),
                            work::HIPMatrix{$T} = hipzeros($T, blocksize, size(A, 2)))
            m, n = size(A)
            blocksize >= 1 || throw(ArgumentError("blocksize must be positive"))
            length(τ) >= min(m, n) || throw(DimensionMismatch("τ is too short"))
            with_handle(h -> _check(@gla($blk, h, m, n, pointer(A), ld(A), Int64(blocksize),
                                         pointer(τ), pointer(work), ld(work), C_NULL),
                                    "qrBlocked!"))
            return QR2{$T,typeof(A),typeof(τ)}(A, τ)
        end
        # F[Tuple{:QBlocked}] — src/qr.jl:64-83
        function Base.getindex(F::QR2{$T,<:HIPMatrix{$T}}, ::Type{Tuple{:QBlocked}})
            m, n = size(F.factors)
            k = min(m, n)
            Tm = hipzeros($T, k, k)
            with_handle(h -> _check(@gla($tbl, h, m, n, pointer(F.factors), ld(F.factors),
                                         pointer(F.τ), pointer(Tm), ld(Tm), C_NULL),
                                    "QR2[QBlocked]"))
            return HIPHouseholderBlock{$T}(F.factors, UpperTriangular(Tm))
        end
        # lmul!(H, A, M) and lmul!(H', A, M) — src/householder.jl:82-115,119-157
        function LinearAlgebra.lmul!(H::HIPHouseholderBlock{$T}, A::HIPMatrix{$T},
                                     M::Union{HIPMatrix{$T},Nothing} = nothing)
            return _wy!(Cchar('N'), H, A, M)
        end
        function LinearAlgebra.lmul!(Ha::Adjoint{$T,HIPHouseholderBlock{$T}}, A::HIPMatrix{$T},
                                     M::Union{HIPMatrix{$T},Nothing} = nothing)
            return _wy!(Cchar('T'), parent(Ha), A, M)
        end
        function _wy!(trans::Cchar, H::HIPHouseholderBlock{$T}, A::HIPMatrix{$T},
                      M::Union{HIPMatrix{$T},Nothing})
            V, Tm = H.V, H.T.data
            mA, nA = size(A)
            size(V, 1) == mA || throw(DimensionMismatch(""))
            Mp = M === nothing ? Ptr{$T}(C_NULL) : pointer(M)
            ldm = M === nothing ? Int64(max(1, min(mA, size(V, 2)))) : Int64(ld(M))
            with_handle(h -> _check(@gla($wya, h, trans, mA, nA, pointer(V), ld(V), size(V, 2),
                                         pointer(Tm), ld(Tm), pointer(A), ld(A), Mp, ldm, C_NULL),
                                    "lmul!"))
            return A
        end
        # lmul!(UnitLowerTriangular(L), B, α) — src/juliaBLAS.jl:154-166
        function LinearAlgebra.lmul!(L::UnitLowerTriangular{$T,<:HIPMatrix{$T}},
                                     B::HIPMatrix{$T}, α::$T)
            m, n = size(B)
            with_handle(h -> _check(@gla($lul, h, m, n, pointer(L.data), ld(L.data), pointer(B),
                                         ld(B), α, C_NULL), "lmul!"))
            return B
        end
        function qr_solve!(F::QR2{$T,<:HIPMatrix{$T}}, B::HIPMatrix{$T})
            m, n = size(F.factors)
            m >= n || throw(ArgumentError("least squares through QR needs m ≥ n"))
            size(B, 1) == m || throw(DimensionMismatch("B has $(size(B, 1)) rows, A has $m"))
            info = Ref{Int64}(0)
            with_handle(h -> _check(@gla($qrs, h, m, n, size(B, 2), pointer(F.factors),
                                         ld(F.factors), pointer(F.τ), pointer(B), ld(B), info,
                                         C_NULL), "qr_solve!"))
            info[] > 0 && throw(ArgumentError("R is singular: zero diagonal entry at $(info[])"))
            return hipview(B, 1:n, 1:size(B, 2))
        end
        function chol_solve!(F::Union{LowerTriangular{$T,<:HIPMatrix{$T}},
                                      UpperTriangular{$T,<:HIPMatrix{$T}}}, B::HIPMatrix{$T})
            A = F.data
            n = _square(A)
            size(B, 1) == n || throw(DimensionMismatch("B has $(size(B, 1)) rows, A is $n×$n"))
            uplo = F isa LowerTriangular ? Cchar('L') : Cchar('U')
            with_handle(h -> _check(@gla($chs, h, uplo, n, size(B, 2), pointer(A), ld(A),
                                         pointer(B), ld(B), C_NULL), "chol_solve!"))
            return B
        end
    end
end

# ldlt!(A::Hermitian, blocksize) and ldiv!/lmul! with its LDLt —
# src/ldlt.jl:150-157,225-234, for a Hermitian view of a device matrix.
for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    fac = Symbol("gla_ldlt_", sfx)
    sol = Symbol("gla_ldlt_solve_", sfx)
    mul = Symbol("gla_ldlt_lmul_", sfx)
    @eval begin
        function LinearAlgebra.ldlt!(A::Hermitian{$T,<:HIPMatrix{$T}},
                                     blocksize::Int = max(1, 128 ÷ sizeof($T)))
            D = A.data
            n = _square(D)
            blocksize >= 1 || throw(ArgumentError("blocksize must be positive"))
            with_handle(h -> _check(@gla($fac, h, Cchar(A.uplo), n, pointer(D), ld(D),
                                         Int64(blocksize), C_NULL), "ldlt!"))
            return LDLt(A)
        end
        function LinearAlgebra.ldiv!(F::LDLt{$T,<:Hermitian{$T,<:HIPMatrix{$T}}}, X::HIPMatrix{$T})
            D = F.data.data
            n = _square(D)
            size(X, 1) == n || throw(DimensionMismatch("X has $(size(X, 1)) rows, F is $n×$n"))
            with_handle(h -> _check(@gla($sol, h, Cchar(F.data.uplo), n, size(X, 2), pointer(D),
                                         ld(D), pointer(X), ld(X), C_NULL), "ldiv!"))
            return X
        end
        function LinearAlgebra.lmul!(F::LDLt{$T,<:Hermitian{$T,<:HIPMatrix{$T}}}, X::HIPMatrix{$T})
            D = F.data.data
            n = _square(D)
            size(X, 1) == n || throw(DimensionMismatch("X has $(size(X, 1)) rows, F is $n×$n"))
            with_handle(h -> _check(@gla($mul, h, Cchar(F.data.uplo), n, size(X, 2), pointer(D),
                                         ld(D), pointer(X), ld(X), C_NULL), "lmul!"))
            return X
        end
    end
end

# symtri!(A::Hermitian) — src/eigenSelfAdjoint.jl:446-564, real element
# types: the reflectors stay on the device (EigenQ), the tridiagonal is
# downloaded (SymTridiagonal), as the reference returns them.
for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    fn = Symbol("gla_symtri_", sfx)
    @eval function GenericLinearAlgebra.symtri!(A::Union{Hermitian{$T,<:HIPMatrix{$T}},
                                                         Symmetric{$T,<:HIPMatrix{$T}}})
        D = A.data
        n = _square(D)
        τ = hipzeros($T, max(n - 1, 0))
        with_handle(h -> _check(@gla($fn, h, Cchar(A.uplo), n, pointer(D), ld(D), pointer(τ),
                                     C_NULL), "symtri!"))
        H = Array(D)
        d = diag(H)
        e = A.uplo == 'L' ? diag(H, -1) : diag(H, 1)
        return GenericLinearAlgebra.SymmetricTridiagonalFactorization(
            GenericLinearAlgebra.EigenQ(A.uplo, D, τ), SymTridiagonal(d, e))
    end
end

# bidiagonalize!(A) — src/svd.jl:328-381, real element types: reflectors and
# τ vectors stay on the device, the bidiagonal is downloaded.
for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    fn = Symbol("gla_bidiag_", sfx)
    @eval function GenericLinearAlgebra.bidiagonalize!(A::HIPMatrix{$T})
        m, n = size(A)
        k = min(m, n)
        nl, nr = m >= n ? (k, max(k - 1, 0)) : (max(k - 1, 0), k)
        τl = hipzeros($T, max(nl, 1))
        τr = hipzeros($T, max(nr, 1))
        with_handle(h -> _check(@gla($fn, h, m, n, pointer(A), ld(A), pointer(τl), pointer(τr),
                                     C_NULL), "bidiagonalize!"))
        H = Array(A)
        bd = m >= n ? Bidiagonal(diag(H), diag(H, 1), :U) : Bidiagonal(diag(H), diag(H, -1), :L)
        return GenericLinearAlgebra.BidiagonalFactorization{$T,Vector{$T},typeof(A),typeof(τl)}(
            bd, A, hipview(τl, 1:nl), hipview(τr, 1:nr))
    end
end

# Complex QR (ComplexF64/ComplexF32): qrUnblocked!, qrBlocked!, the
# compact-WY block and lmul! with it — src/qr.jl:64-146,
# src/householder.jl:82-157 (the workspace stays internal).
for (T, sfx) in ((ComplexF64, "z"), (ComplexF32, "c"))
    unb = Symbol("gla_qr_unblocked_", sfx)
    blk = Symbol("gla_qr_blocked_", sfx)
    tbl = Symbol("gla_qr_tblock_", sfx)
    wya = Symbol("gla_wy_apply_", sfx)
    @eval begin
        function qrUnblocked!(A::HIPMatrix{$T},
                              τ::HIPVector{$T} = hipzeros($T, min(size(A)...)))
            m, n = size(A)
            with_handle(h -> _check(@gla($unb, h, m, n, pointer(A), ld(A), pointer(τ), C_NULL),
                                    "qrUnblocked!"))
            return QR2{$T,typeof(A),typeof(τ)}(A, τ)
        end
        function qrBlocked!(A::HIPMatrix{$T}, blocksize::Integer = 12,
                            τ::HIPVector{$T} = hipzeros($T, min(size(A)...)))
            m, n = size(A)
            blocksize >= 1 || throw(ArgumentError("blocksize must be positive"))
            with_handle(h -> _check(@gla($blk, h, m, n, pointer(A), ld(A), Int64(blocksize),
                                         pointer(τ), C_NULL), "qrBlocked!"))
            return QR2{$T,typeof(A),typeof(τ)}(A, τ)
        end
        function Base.getindex(F::QR2{$T,<:HIPMatrix{$T}}, ::Type{Tuple{:QBlocked}})
            m, n = size(F.factors)
            k = min(m, n)
            Tm = hipzeros($T, k, k)
            with_handle(h -> _check(@gla($tbl, h, m, n, pointer(F.factors), ld(F.factors),
                                         pointer(F.τ), pointer(Tm), ld(Tm), C_NULL),
                                    "QR2[QBlocked]"))
            return HIPHouseholderBlock{$T}(F.factors, UpperTriangular(Tm))
        end
        function LinearAlgebra.lmul!(H::HIPHouseholderBlock{$T}, A::HIPMatrix{$T})
            V, Tm = H.V, H.T.data
            with_handle(h -> _check(@gla($wya, h, Cchar('N'), size(A, 1), size(A, 2), pointer(V),
                                         ld(V), size(V, 2), pointer(Tm), ld(Tm), pointer(A),
                                         ld(A), C_NULL), "lmul!"))
            return A
        end
        function LinearAlgebra.lmul!(Ha::Adjoint{$T,HIPHouseholderBlock{$T}}, A::HIPMatrix{$T})
            H = parent(Ha)
            V, Tm = H.V, H.T.data
            with_handle(h -> _check(@gla($wya, h, Cchar('C'), size(A, 1), size(A, 2), pointer(V),
                                         ld(V), size(V, 2), pointer(Tm), ld(Tm), pointer(A),
                                         ld(A), C_NULL), "lmul!"))
            return A
        end
    end
end

@doc """
    qr_solve!(F::QR2, B) -> X

Least squares through a device QR (m ≥ n): B ← Qᵀ B (lmul!(Hᴴ, B),
src/householder.jl:119-157), then its first n rows ← R⁻¹ of themselves;
returns that n-row view.
""" qr_solve!

@doc """
    chol_solve!(F, B) -> B

B ← A⁻¹ B for A = L Lᵀ (F::LowerTriangular, as cholRecursive! returns it) or
A = Uᵀ U (F::UpperTriangular) — Julia's ldiv! with a Cholesky.
""" chol_solve!

Base.:*(H::HIPHouseholderBlock, A::HIPMatrix) = lmul!(H, copy(A))
Base.:*(Ha::Adjoint{<:Any,<:HIPHouseholderBlock}, A::HIPMatrix) = lmul!(Ha, copy(A))

# F[Tuple{:R}] — src/qr.jl:55-62
function Base.getindex(F::QR2{T,<:HIPMatrix{T}}, ::Type{Tuple{:R}}) where {T}
    m, n = size(F.factors)
    m >= n || throw(ArgumentError("R matrix is trapezoid and cannot be extracted with indexing"))
    return UpperTriangular(hipview(F.factors, 1:n, 1:n))
end

LinearAlgebra.ldiv!(F::QR2{T,<:HIPMatrix{T}}, B::HIPMatrix{T}) where {T<:Floats} = qr_solve!(F, B)

# ------------------------------------------- mixed precision + refinement --
"""
    mixed_solve(A, B; precision = 'a', tol = 1e-12, max_iter = 30) -> (X, info)

X = A⁻¹B for SPD Float64 A (lower triangle read) factored in `precision`
('d' FP64, 's' FP32, 'b' BF16 trailing updates, 'h' FP16, 'a' FP32 then
FP64) and refined in Float64 to a normwise backward error ≤ tol.
"""
function mixed_solve(A::HIPMatrix{Float64}, B::HIPMatrix{Float64}; precision::Char = 'a',
                     tol::Real = 1e-12, max_iter::Integer = 30)
    n = _square(A)
    size(B, 1) == n || throw(DimensionMismatch("B has $(size(B, 1)) rows, A is $n×$n"))
    X = hipzeros(Float64, n, size(B, 2))
    iters = Ref{Int32}(0)
    berr = Ref{Float64}(0)
    info = Ref{Int64}(0)
    used = Ref{Cchar}(0)
    with_handle(h -> _check(@gla(gla_chol_solve_mixed_d, h, Cchar(precision), n, pointer(A),
                                 ld(A), size(B, 2), pointer(B), ld(B), pointer(X), ld(X),
                                 Float64(tol), Int32(max_iter), iters, berr, info, used, C_NULL),
                            "mixed_solve"))
    _domain(info[])
    return X, (iters = Int(iters[]), berr = berr[], precision = Char(used[]),
               converged = berr[] <= tol)
end

# ------------------------------------------------------- batched solves --
"""
    BatchPlan(sizes; T = Float64, wave_max = 256, grid_mult = 4)

The sizes of a batch of SPD matrices sorted into the batched solver's bins
(gla_batch_plan_create); reuse it for every call with those sizes.
"""
mutable struct BatchPlan
    ptr::Ptr{Cvoid}
    batch::Int
    function BatchPlan(sizes::AbstractVector{<:Integer}; wave_max::Integer = 256,
                       grid_mult::Integer = 4)
        ns = Int32.(collect(sizes))
        r = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve ns with_handle(h -> _check(@gla(gla_batch_plan_create, h, length(ns),
                                                     pointer(ns), Int32(wave_max),
                                                     Int32(grid_mult), r),
                                                "BatchPlan"))
        p = new(r[], length(ns))
        finalizer(p) do x
            x.ptr == C_NULL || ccall((:gla_batch_plan_destroy, libgla), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
        return p
    end
end

for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    fn = Symbol("gla_chol_solve_batched_", sfx)
    @eval function chol_solve_batched!(plan::BatchPlan, A::HIPVector{Ptr{$T}},
                                       lda::HIPVector{Int32}, b::HIPVector{Ptr{$T}},
                                       info::HIPVector{Int32})
        with_handle(h -> _check(@gla($fn, h, plan.ptr, pointer(A), pointer(lda), pointer(b),
                                     pointer(info), C_NULL), "chol_solve_batched!"))
        return info
    end
end

@doc """
    chol_solve_batched!(plan, A, lda, b, info) -> info

Factors every A_i = L_i L_iᵀ in place (cholUnblocked!'s result,
src/cholesky.jl:3-15) and overwrites b_i with A_i⁻¹ b_i, in one launch.
`A`, `b`: device vectors of pointers; `lda`, `info`: device Int32 vectors.
""" chol_solve_batched!

# ---------------------------------------------------------------- adjoints --
for (T, sfx) in ((Float64, "d"), (Float32, "s"))
    cad = Symbol("gla_chol_adjoint_", sfx)
    cab = Symbol("gla_chol_adjoint_batched_", sfx)
    qad = Symbol("gla_qr_adjoint_", sfx)
    @eval begin
        function chol_adjoint(L::HIPMatrix{$T}, Lbar::HIPMatrix{$T}; conv::Char = 'S')
            n = _square(L)
            size(Lbar) == (n, n) || throw(DimensionMismatch("L̄ must be $n×$n"))
            Abar = hipzeros($T, n, n)
            with_handle(h -> _check(@gla($cad, h, Cchar(conv), n, pointer(L), ld(L),
                                         pointer(Lbar), ld(Lbar), pointer(Abar), ld(Abar),
                                         C_NULL), "chol_adjoint"))
            return Abar
        end
        function chol_adjoint_batched!(Abar::HIPArray{$T,3}, A::HIPArray{$T,3},
                                       Lbar::HIPArray{$T,3}; conv::Char = 'S')
            n, n2, b = size(A)
            n == n2 || throw(DimensionMismatch("A must be n×n×batch"))
            size(Lbar) == size(A) == size(Abar) || throw(DimensionMismatch("sizes differ"))
            info = hipzeros(Int32, max(b, 1))
            s(X) = Int64(ld(X) * n)
            with_handle(h -> _check(@gla($cab, h, Cchar(conv), n, b, pointer(A), ld(A), s(A),
                                         pointer(Lbar), ld(Lbar), s(Lbar), pointer(Abar),
                                         ld(Abar), s(Abar), pointer(info), C_NULL),
                                    "chol_adjoint_batched!"))
            bad = findfirst(!iszero, Array(info)[1:b])
            bad === nothing ||
                throw(DomainError(bad, "matrix $bad is not positive definite"))
            return Abar
        end
        function qr_adjoint(F::QR2{$T,<:HIPMatrix{$T}}, Rbar::HIPMatrix{$T})
            m, n = size(F.factors)
            m >= n || throw(ArgumentError("the QR adjoint needs m ≥ n"))
            size(Rbar) == (n, n) || throw(DimensionMismatch("R̄ must be $n×$n"))
            Abar = hipzeros($T, m, n)
            with_handle(h -> _check(@gla($qad, h, m, n, pointer(F.factors), ld(F.factors),
                                         pointer(F.τ), pointer(Rbar), ld(Rbar), pointer(Abar),
                                         ld(Abar), C_NULL), "qr_adjoint"))
            return Abar
        end
    end
end

@doc """
    chol_adjoint(L, L̄; conv = 'S') -> Ā

Ā = ½ L⁻ᵀ(P + Pᵀ)L⁻¹, P = Φ(Lᵀ L̄) (Φ: lower triangle, halved
diagonal) for the factor L of cholRecursive!; 'L' returns the
derivative with respect to the lower triangle the factorisation reads.
""" chol_adjoint

@doc """
    chol_adjoint_batched!(Ā, A, L̄; conv = 'S') -> Ā

The same for n×n×batch device arrays (n ≤ 128), each factor
recomputed from A's lower triangle.
""" chol_adjoint_batched!

@doc """
    qr_adjoint(F::QR2, R̄) -> Ā

Ā = Q[copyltu(R R̄ᵀ) R⁻ᵀ; 0] for A = QR (m ≥ n) stored by
qrBlocked!/qrUnblocked! and the cotangent R̄ (upper part read), Q̄ = 0.
""" qr_adjoint

_tangent(x::HIPMatrix) = x
_tangent(x::Union{LowerTriangular,UpperTriangular}) = x.data
_tangent(x) = throw(ArgumentError("unsupported tangent type $(typeof(x)) for a device factor"))

# ChainRulesCore rules.  The mutating methods return the factor that
# overwrote A, so their pullbacks read L from A itself.
for f in (:cholUnblocked!, :cholRecursive!)
    @eval function ChainRulesCore.rrule(::typeof($f), A::HIPMatrix{T}, ::Type{Val{:L}},
                                        args::Integer...) where {T<:Floats}
        Y = $f(A, Val{:L}, args...)
        function pullback(Ybar)
            Abar = chol_adjoint(A, _tangent(unthunk(Ybar)); conv = 'S')
            return (NoTangent(), Abar, NoTangent(), map(_ -> NoTangent(), args)...)
        end
        return Y, pullback
    end
end
function ChainRulesCore.rrule(::typeof(cholBlocked!), A::HIPMatrix{T}, ::Type{Val{:L}},
                              blocksize::Integer) where {T<:Floats}
    Y = cholBlocked!(A, Val{:L}, blocksize)
    pullback(Ybar) = (NoTangent(), chol_adjoint(A, _tangent(unthunk(Ybar)); conv = 'S'),
                      NoTangent(), NoTangent())
    return Y, pullback
end

"""`cholRecursive!(copy(A), Val{:L})` without touching A, with an rrule."""
chol_lower(A::HIPMatrix) = cholRecursive!(copy(A), Val{:L})
function ChainRulesCore.rrule(::typeof(chol_lower), A::HIPMatrix{T}) where {T<:Floats}
    L = chol_lower(A)
    pullback(Lbar) = (NoTangent(), chol_adjoint(L.data, _tangent(unthunk(Lbar)); conv = 'S'))
    return L, pullback
end

"""R of `qrBlocked!(copy(A), blocksize)` (m ≥ n), with an rrule (R̄ only)."""
qr_R(A::HIPMatrix, blocksize::Integer = 128) = qrBlocked!(copy(A), blocksize)[Tuple{:R}]
function ChainRulesCore.rrule(::typeof(qr_R), A::HIPMatrix{T},
                              blocksize::Integer = 128) where {T<:Floats}
    F = qrBlocked!(copy(A), blocksize)
    pullback(Rbar) = (NoTangent(), qr_adjoint(F, _tangent(unthunk(Rbar))), NoTangent())
    return F[Tuple{:R}], pullback
end

end # module
