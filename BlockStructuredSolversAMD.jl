# BlockStructuredSolversAMD.jl -- the reference's Julia API (initialize_batched /
# initialize_seq / factorize! / solve!, src/BlockStructuredSolvers.jl:9-11) on libbss, the
# MI355X HIP library (C ABI: include/bss.h), behind a device array type `HipArray`.
#
# With this module loaded after BlockStructuredSolvers, a caller written for ROCArray
# (test/test_rocm.jl:11-47,97-107; benchmark/benchmark_gpu.jl:124-153) changes only
# `M = ROCArray` to `M = HipArray`: HipArray supports what those callers do with M --
# `M(host_array)`, `M(zeros(...))`, `copyto!` host<->device and device<->device, block
# reads `A[:, :, i]`, block views (`@views`, `data.d_list[i]`), `.=` broadcasts into a
# block, `*`, `+`, `-`, `'` on blocks, `norm`, `Array(x)` -- and every solver call is one
# ccall into libbss.  Arithmetic on HipArrays outside the solver is evaluated on the host
# (a copy down, the operation, a copy up): it is test-data plumbing, as @allowscalar code
# on a ROCArray is; the solve itself never leaves the device.
#
# Julia is not part of this build's toolchain, so this file has not been run here; the
# Python ctypes mirror (blockstructuredsolvers_amd/solver.py) makes the same calls in the
# same order and is what the GPU tests exercise.
module BlockStructuredSolversAMD

using LinearAlgebra
import BlockStructuredSolvers: initialize_batched, initialize_seq, factorize!, solve!

export HipArray, factorize_solve!, info, backward_error, synchronize, iterations

const libbss = get(ENV, "LIBBSS_PATH",
                   joinpath(@__DIR__, "blockstructuredsolvers_amd", "libbss.so"))
const libhip = "libamdhip64.so"

const H2D = Cint(1)
const D2H = Cint(2)
const D2D = Cint(3)

hipcheck(rc::Cint, what::AbstractString) =
    rc == 0 || error("HIP error $rc in $what")

function memcpy!(dst::Ptr, src::Ptr, nbytes::Integer, kind::Cint)
    nbytes == 0 && return nothing
    hipcheck(ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
                   dst, src, nbytes, kind), "hipMemcpy")
    nothing
end

# ----------------------------------------------------------------------------- HipArray

"""
    HipArray{T,N}

Column-major device array.  Either it owns a hipMalloc'd buffer (freed by a finalizer),
or it aliases part of another HipArray (`view(A, :, :, i)` and the per-block lists of
the solver data), whose memory it keeps alive through `parent`.
"""
mutable struct HipArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    parent::Any  # nothing: owns ptr; otherwise the array whose memory this aliases
end

function HipArray{T,N}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
    nbytes = prod(dims) * sizeof(T)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    if nbytes > 0
        hipcheck(ccall((:hipMalloc, libhip), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), p, nbytes),
                 "hipMalloc")
        hipcheck(ccall((:hipMemset, libhip), Cint, (Ptr{Cvoid}, Cint, Csize_t), p[], 0,
                       nbytes), "hipMemset")
    end
    a = HipArray{T,N}(Ptr{T}(p[]), dims, nothing)
    finalizer(a) do x
        x.ptr != C_NULL && ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), x.ptr)
    end
    return a
end
HipArray{T,N}(::UndefInitializer, dims::Vararg{Int,N}) where {T,N} =
    HipArray{T,N}(undef, dims)
HipArray{T}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N} = HipArray{T,N}(undef, dims)
HipArray{T}(::UndefInitializer, dims::Vararg{Int,N}) where {T,N} = HipArray{T,N}(undef, dims)

# M(host_array): a device copy (test/test_rocm.jl:14-15,34,38)
function HipArray(a::AbstractArray{T,N}) where {T,N}
    h = HipArray{T,N}(undef, size(a))
    copyto!(h, Array{T,N}(a))
    return h
end
HipArray{T}(a::AbstractArray{S,N}) where {T,S,N} = HipArray(Array{T,N}(a))
HipArray(a::HipArray) = copy(a)

Base.size(a::HipArray) = a.dims
Base.pointer(a::HipArray) = a.ptr
Base.unsafe_convert(::Type{Ptr{T}}, a::HipArray{T}) where {T} = a.ptr
Base.IndexStyle(::Type{<:HipArray}) = IndexLinear()
Base.similar(::HipArray, ::Type{S}, dims::Dims{M}) where {S,M} = HipArray{S,M}(undef, dims)

function Base.copyto!(dst::HipArray{T}, src::Array{T}) where {T}       # host -> device
    length(src) <= length(dst) || throw(DimensionMismatch("source has more elements"))
    GC.@preserve src memcpy!(dst.ptr, pointer(src), length(src) * sizeof(T), H2D)
    return dst
end
Base.copyto!(dst::HipArray{T}, src::Array{S}) where {T,S} = copyto!(dst, Array{T}(src))

function Base.copyto!(dst::Array{T}, src::HipArray{T}) where {T}       # device -> host
    length(src) <= length(dst) || throw(DimensionMismatch("source has more elements"))
    GC.@preserve dst memcpy!(pointer(dst), src.ptr, length(src) * sizeof(T), D2H)
    return dst
end

# device -> device (copyto!(data.A_tensor, A_tensor), test/test_rocm.jl:29-30,47)
function Base.copyto!(dst::HipArray{T}, src::HipArray{T}) where {T}
    length(src) <= length(dst) || throw(DimensionMismatch("source has more elements"))
    memcpy!(dst.ptr, src.ptr, length(src) * sizeof(T), D2D)
    return dst
end

Base.Array(a::HipArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, a.dims), a)
Base.collect(a::HipArray) = Array(a)
Base.copy(a::HipArray{T,N}) where {T,N} = copyto!(HipArray{T,N}(undef, a.dims), a)

# scalar access (what @allowscalar permits on a ROCArray): one small copy per element
function Base.getindex(a::HipArray{T}, i::Int) where {T}
    @boundscheck checkbounds(a, i)
    v = Vector{T}(undef, 1)
    GC.@preserve v memcpy!(pointer(v), a.ptr + (i - 1) * sizeof(T), sizeof(T), D2H)
    return v[1]
end
function Base.setindex!(a::HipArray{T}, x, i::Int) where {T}
    @boundscheck checkbounds(a, i)
    v = T[convert(T, x)]
    GC.@preserve v memcpy!(a.ptr + (i - 1) * sizeof(T), pointer(v), sizeof(T), H2D)
    return a
end

# block i of a 3-d array is contiguous (column-major): its view is an alias, and
# `@views ... A[:, :, i]`, `A[:, :, i] .= x` and the solver's d_list use it
function Base.view(a::HipArray{T,3}, ::Colon, ::Colon, i::Int) where {T}
    1 <= i <= a.dims[3] || throw(BoundsError(a, (:, :, i)))
    return HipArray{T,2}(a.ptr + (i - 1) * a.dims[1] * a.dims[2] * sizeof(T),
                         (a.dims[1], a.dims[2]), a)
end
Base.getindex(a::HipArray{T,3}, ::Colon, ::Colon, i::Int) where {T} = copy(view(a, :, :, i))
function Base.setindex!(a::HipArray{T,3}, v::HipArray, ::Colon, ::Colon, i::Int) where {T}
    blk = view(a, :, :, i)
    length(v) == length(blk) || throw(DimensionMismatch("block of $(size(blk))"))
    copyto!(blk, eltype(v) === T ? v : HipArray{T}(Array(v)))
    return a
end
function Base.setindex!(a::HipArray{T,3}, v::AbstractArray, ::Colon, ::Colon, i::Int) where {T}
    blk = view(a, :, :, i)
    length(v) == length(blk) || throw(DimensionMismatch("block of $(size(blk))"))
    copyto!(blk, Array{T}(v))
    return a
end

# arithmetic on blocks (right-hand-side construction in the callers): on the host
Base.:*(a::HipArray{T,2}, b::HipArray{T,2}) where {T} = HipArray(Array(a) * Array(b))
Base.:*(a::Adjoint{T,HipArray{T,2}}, b::HipArray{T,2}) where {T} =
    HipArray(Array(parent(a))' * Array(b))
Base.:+(a::HipArray{T,N}, b::HipArray{T,N}) where {T,N} = HipArray(Array(a) + Array(b))
Base.:-(a::HipArray{T,N}, b::HipArray{T,N}) where {T,N} = HipArray(Array(a) - Array(b))
LinearAlgebra.norm(a::HipArray, p::Real=2) = norm(Array(a), p)

# broadcasting (dest .= src, a .+ b): evaluated on the host, the result copied up
struct HipStyle{N} <: Broadcast.AbstractArrayStyle{N} end
HipStyle{M}(::Val{N}) where {M,N} = HipStyle{N}()
Base.BroadcastStyle(::Type{<:HipArray{T,N}}) where {T,N} = HipStyle{N}()
hostarg(x::HipArray) = Array(x)
hostarg(x::Broadcast.Broadcasted) = Broadcast.broadcasted(x.f, map(hostarg, x.args)...)
hostarg(x) = x
function Base.similar(bc::Broadcast.Broadcasted{HipStyle{N}}, ::Type{T}) where {N,T}
    return HipArray{T}(undef, map(length, axes(bc)))
end
function Base.copyto!(dest::HipArray{T}, bc::Broadcast.Broadcasted{HipStyle{N}}) where {T,N}
    h = Array{T}(undef, size(dest))
    h .= Broadcast.materialize(hostarg(bc))
    return copyto!(dest, h)
end

# ------------------------------------------------------------------------ libbss calls

const BSS_F32 = Cint(0)
const BSS_F64 = Cint(1)
const BSS_Q_N_USE = Cint(0)
const BSS_Q_LAST_ITERS = Cint(112)
const BSS_OPT_TOL = Cint(0)
const BSS_OPT_BASE_MAX = Cint(4)
const BSS_NOT_CONVERGED = Cint(1)

"""Raise on a negative libbss status (argument or HIP error); return a non-negative one
(0, or BSS_NOT_CONVERGED from a mixed-mode solve -- like the reference, a numerical
outcome is not an exception)."""
function check(rc::Cint)
    rc >= 0 && return rc
    buf = Vector{UInt8}(undef, 1024)
    GC.@preserve buf ccall((:bss_last_error, libbss), Cint, (Ptr{UInt8}, Csize_t), buf,
                           length(buf))
    error("libbss status $rc: ", GC.@preserve buf unsafe_string(pointer(buf)))
end

dtype_code(::Type{Float64}) = BSS_F64
dtype_code(::Type{Float32}) = BSS_F32

"""
Solver data of one system on the device, with the reference's fields
(BlockTriDiagData_batched, src/gpu_batched.jl:1-71): `N` (padded size N_use), `n`,
`A_tensor` (n, n, N), `B_tensor` (n, n, N-1), `d_tensor` (n, 1, N), and the per-block
views `A_list`, `B_list`, `d_list` (aliases of the tensors' blocks).  `plan` is the
libbss plan that holds the schedule and the per-level workspace.
"""
mutable struct BlockTriDiagData_bss{T}
    plan::Ptr{Cvoid}
    N::Int
    n::Int
    A_tensor::HipArray{T,3}
    B_tensor::HipArray{T,3}
    d_tensor::HipArray{T,3}
    A_list::Vector{HipArray{T,2}}
    B_list::Vector{HipArray{T,2}}
    d_list::Vector{HipArray{T,2}}
    mixed::Bool
end

function make_data(plan::Ptr{Cvoid}, N::Int, n::Int, ::Type{T}, mixed::Bool) where {T}
    nuse = Ref{Int64}(0)
    check(ccall((:bss_plan_get, libbss), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), plan,
                BSS_Q_N_USE, nuse))
    Nu = Int(nuse[])
    A = HipArray{T,3}(undef, (n, n, Nu))
    B = HipArray{T,3}(undef, (n, n, max(Nu - 1, 0)))
    d = HipArray{T,3}(undef, (n, 1, Nu))
    if Nu > N  # identity padding of blocks N..N_use (src/gpu_batched.jl:177-179)
        pad = zeros(T, n, n, Nu - N + 1)
        for i in axes(pad, 3)
            pad[:, :, i] .= Matrix{T}(I, n, n)
        end
        GC.@preserve pad memcpy!(A.ptr + (N - 1) * n * n * sizeof(T), pointer(pad),
                                 length(pad) * sizeof(T), H2D)
    end
    check(ccall((:bss_plan_bind, libbss), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}), plan,
                A.ptr, Nu > 1 ? B.ptr : Ptr{T}(C_NULL), d.ptr))
    data = BlockTriDiagData_bss{T}(plan, Nu, n, A, B, d,
                                   [view(A, :, :, i) for i in 1:Nu],
                                   [view(B, :, :, i) for i in 1:max(Nu - 1, 0)],
                                   [view(d, :, :, i) for i in 1:Nu], mixed)
    finalizer(x -> ccall((:bss_plan_destroy, libbss), Cint, (Ptr{Cvoid},), x.plan), data)
    return data
end

"""
    initialize_batched(N, n, T, HipArray; mixed=false, reference_schedule=false, tol=1e-12)

The reference's `initialize_batched(N, n, T, M)` (src/gpu_batched.jl:165-182) for
`M = HipArray`: the schedule search and all workspace allocation happen in libbss
(`bss_plan_create`).  `reference_schedule=true` reproduces the reference's level
schedule and its identity padding to N_use; `mixed=true` (T = Float64) factors in FP32
and refines in FP64 to the normwise backward error `tol`.
"""
function initialize_batched(N::Int, n::Int, ::Type{T}, ::Type{HipArray};
                            mixed::Bool=false, reference_schedule::Bool=false,
                            tol::Real=1e-12) where {T}
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bss_plan_create, libbss), Cint,
                (Cint, Cint, Cint, Cint, Cint, Cint, Cint, Ptr{Ptr{Cvoid}}),
                N, n, 1, 1, dtype_code(T), mixed ? 1 : 0, reference_schedule ? 1 : 0, plan))
    if mixed
        check(ccall((:bss_set_option, libbss), Cint, (Ptr{Cvoid}, Cint, Cdouble), plan[],
                    BSS_OPT_TOL, Float64(tol)))
    end
    return make_data(plan[], N, n, T, mixed)
end

"""
    initialize_seq(N, n, T, HipArray)

The sequential solver (src/gpu_seq.jl:45-84): the same plan with a single-chain
schedule (`BSS_OPT_BASE_MAX = N`), so one workgroup eliminates the N blocks in order.
"""
function initialize_seq(N::Int, n::Int, ::Type{T}, ::Type{HipArray}) where {T}
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bss_plan_create, libbss), Cint,
                (Cint, Cint, Cint, Cint, Cint, Cint, Cint, Ptr{Ptr{Cvoid}}),
                N, n, 1, 1, dtype_code(T), 0, 0, plan))
    check(ccall((:bss_set_option, libbss), Cint, (Ptr{Cvoid}, Cint, Cdouble), plan[],
                BSS_OPT_BASE_MAX, Float64(N)))
    return make_data(plan[], N, n, T, false)
end

"""`factorize!(data)` (src/gpu_batched.jl:184-227): enqueued on the null stream."""
factorize!(data::BlockTriDiagData_bss) =
    check(ccall((:bss_factorize, libbss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), data.plan, C_NULL))

"""`solve!(data)` (src/gpu_batched.jl:229-259): the solution replaces d_tensor.  Returns
0, or BSS_NOT_CONVERGED (mixed mode, target not reached even by the FP64 fallback)."""
solve!(data::BlockTriDiagData_bss) =
    check(ccall((:bss_solve, libbss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), data.plan, C_NULL))

"""`factorize!(data); solve!(data)` in one call, for a d that is final before the
factorisation (the first solve's forward sweeps overlap the factorisation)."""
factorize_solve!(data::BlockTriDiagData_bss) =
    check(ccall((:bss_factorize_solve, libbss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), data.plan,
                C_NULL))

"""Wait for the data's device work (what `AMDGPU.@sync` does around the reference's
calls, test/test_rocm.jl:97,101)."""
synchronize(data::BlockTriDiagData_bss) =
    check(ccall((:bss_sync, libbss), Cint, (Ptr{Cvoid},), data.plan))

"""LAPACK-order status of the last factorisation: 0, or the 1-based block where the
sequential block Cholesky meets its first non-positive pivot (src/backbone_cpu.jl:5-15).
The reference passes potrf an info buffer and never reads it (src/backbone_rocm.jl:64-80)."""
function info(data::BlockTriDiagData_bss)
    v = Ref{Cint}(0)
    check(ccall((:bss_info, libbss), Cint, (Ptr{Cvoid}, Ptr{Cint}), data.plan, v))
    return Int(v[])
end

"""Mixed mode: the device-measured normwise backward error of the last solve."""
function backward_error(data::BlockTriDiagData_bss)
    v = Ref{Cdouble}(0.0)
    check(ccall((:bss_last_backward_error, libbss), Cint, (Ptr{Cvoid}, Ptr{Cdouble}),
                data.plan, v))
    return v[]
end

"""Mixed mode: refinement iterations of the last solve."""
function iterations(data::BlockTriDiagData_bss)
    v = Ref{Int64}(0)
    check(ccall((:bss_plan_get, libbss), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), data.plan,
                BSS_Q_LAST_ITERS, v))
    return Int(v[])
end

end # module
