# Julia side of the drop-in: methods of Muscle's Tensor functions for tensors
# whose data live in libtenet_hip.so device memory.  Tenet.jl's own code
# (canonize.jl, simple_update.jl, overlap.jl, evolve.jl, DMRG.jl) calls
# Muscle.binary_einsum / tensor_qr_thin / tensor_svd_thin / simple_update
# unchanged; dispatch on the array type routes those calls here.
#
# Not compiled in this repository's image (it has no Julia); the C side it
# calls is the header include/tenet_hip.h, exercised by the Python host
# package through the same symbols.
module TenetHIP

using Muscle: Muscle, Tensor, Index, inds
using LinearAlgebra

const libtenet = joinpath(@__DIR__, "..", "lib", "libtenet_hip.so")
const MAX_RANK = 8

struct TenetError <: Exception
    status::Cint
    msg::String
end

function check(st::Cint)
    st == 0 && return nothing
    throw(TenetError(st, unsafe_string(ccall((:tenet_last_error, libtenet), Cstring, ()))))
end

# ---- context ---------------------------------------------------------------
mutable struct Context
    h::Ptr{Cvoid}
end
const CTX = Ref{Union{Nothing,Context}}(nothing)
function context()
    if CTX[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tenet_create, libtenet), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h))
        CTX[] = Context(h[])
        atexit(() -> ccall((:tenet_destroy, libtenet), Cint, (Ptr{Cvoid},), h[]))
    end
    return CTX[]::Context
end

# ---- device arrays (column-major, like Array) --------------------------------
dtype_code(::Type{ComplexF64}) = Cint(0)
dtype_code(::Type{ComplexF32}) = Cint(1)

mutable struct HIPArray{T<:Union{ComplexF64,ComplexF32},N} <: DenseArray{T,N}
    ptr::Ptr{Cvoid}
    dims::NTuple{N,Int}
    owner::Any  # the array whose allocation a view shares (nothing: owns its block)
    function HIPArray{T,N}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tenet_malloc, libtenet), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                    context().h, max(prod(dims; init=1) * sizeof(T), 1), p))
        a = new{T,N}(p[], dims, nothing)
        finalizer(x -> ccall((:tenet_free, libtenet), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), context().h, x.ptr), a)
        return a
    end
    # a view of `owner`'s memory: no allocation, no finalizer; keeps owner alive
    HIPArray{T,N}(ptr::Ptr{Cvoid}, dims::NTuple{N,Int}, owner) where {T,N} = new{T,N}(ptr, dims, owner)
end
Base.size(a::HIPArray) = a.dims
Base.pointer(a::HIPArray) = a.ptr

function HIPArray(x::Array{T,N}) where {T,N}
    a = HIPArray{T,N}(undef, size(x))
    GC.@preserve x check(ccall((:tenet_memcpy_h2d, libtenet), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                               context().h, a.ptr, pointer(x), sizeof(x), C_NULL))
    return a
end
function Base.Array(a::HIPArray{T,N}) where {T,N}
    x = Array{T,N}(undef, a.dims)
    GC.@preserve x check(ccall((:tenet_memcpy_d2h, libtenet), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                               context().h, pointer(x), a.ptr, sizeof(x), C_NULL))
    return x
end

# ---- tenet_tdesc ------------------------------------------------------------
struct TDesc
    rank::Int32
    reserved::Int32
    extent::NTuple{MAX_RANK,Int64}
    stride::NTuple{MAX_RANK,Int64}
    label::NTuple{MAX_RANK,Int64}
end
pad(v) = ntuple(i -> i <= length(v) ? Int64(v[i]) : Int64(0), MAX_RANK)
function tdesc(dims, labels, strides=cumprod([1; collect(dims)])[1:end-1])
    TDesc(Int32(length(dims)), Int32(0), pad(dims), pad(strides), pad(labels))
end

# Index labels -> int64 codes, fresh per call
codes() = (d = Dict{Index,Int64}(); i -> get!(d, i, length(d) + 1))

const HIPTensor{T,N} = Tensor{T,N,<:HIPArray{T,N}}

# ---- Muscle.binary_einsum (canonize.jl:44-61, overlap.jl:26-46, evolve.jl:29-119)
function Muscle.binary_einsum(a::HIPTensor{T}, b::HIPTensor{T}; dims=intersect(inds(a), inds(b)),
                              out=nothing) where {T}
    ia, ib = inds(a), inds(b)
    oc = out === nothing ? vcat(setdiff(ia, dims), setdiff(ib, ia)) : out
    sz = Dict(zip(ia, size(a)))
    merge!(sz, Dict(zip(ib, size(b))))
    c = HIPArray{T,length(oc)}(undef, Tuple(sz[i] for i in oc))
    code = codes()
    da, db, dc = tdesc(size(a), code.(ia)), tdesc(size(b), code.(ib)), tdesc(size(c), code.(oc))
    check(ccall((:tenet_einsum, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ref{TDesc}, Ptr{Cvoid}, Ref{TDesc}, Ptr{Cvoid}, Ref{TDesc},
                 Ptr{Cvoid}, Cint, Cint),
                context().h, C_NULL, dtype_code(T), Cint(0), da, parent(a).ptr, db, parent(b).ptr, dc, c.ptr, 0, 0))
    return Tensor(c, oc)
end

# permute a device tensor into (rows..., cols...) order and return the matrix view
function matricize(A::HIPTensor{T}, rows, cols) where {T}
    order = vcat(rows, cols)
    P = Muscle.binary_einsum(A, Tensor(HIPArray(ones(T)), Index[]); dims=Index[], out=order)
    m = prod(size(A, i) for i in rows; init=1)
    n = prod(size(A, i) for i in cols; init=1)
    return parent(P), m, n
end

# ---- Muscle.tensor_qr_thin (canonize.jl:58) -----------------------------------
function Muscle.tensor_qr_thin(A::HIPTensor{T}; inds_q, inds_r, ind_virtual) where {T}
    M, m, n = matricize(A, inds_q, inds_r)
    k = min(m, n)
    Q = HIPArray{T,2}(undef, (m, k))
    R = HIPArray{T,2}(undef, (k, n))
    check(ccall((:tenet_qr_thin, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64),
                context().h, C_NULL, dtype_code(T), m, n, M.ptr, m, Q.ptr, m, R.ptr, k))
    dq = Tuple(size(A, i) for i in inds_q)
    dr = Tuple(size(A, i) for i in inds_r)
    return Tensor(reshape_view(Q, (dq..., k)), [inds_q..., ind_virtual]),
           Tensor(reshape_view(R, (k, dr...)), [ind_virtual, inds_r...])
end

# ---- Muscle.tensor_svd_thin (canonize.jl:41,97; evolve.jl:62) -----------------
function Muscle.tensor_svd_thin(A::HIPTensor{T}; inds_u, inds_v=setdiff(inds(A), inds_u), ind_s,
                                maxdim=0, threshold=0.0) where {T}
    M, m, n = matricize(A, inds_u, inds_v)
    kmax = min(m, n)
    U = HIPArray{T,2}(undef, (m, kmax))
    V = HIPArray{T,2}(undef, (kmax, n))
    S = HIPArray{ComplexF64,1}(undef, (cld(kmax, 2),))  # kmax doubles
    kd = HIPArray{ComplexF64,1}(undef, (1,))            # k (int64) + discarded weight
    check(ccall((:tenet_svd_thin, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid},
                 Ptr{Cvoid}, Int64, Int64, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}),
                context().h, C_NULL, dtype_code(T), m, n, M.ptr, m, U.ptr, m, S.ptr, V.ptr, kmax, maxdim, threshold,
                kd.ptr, kd.ptr + 8))
    s = reinterpret(Float64, Array(S))[1:kmax]
    # the kept count of compress.jl:46-59 (maxdim, then the absolute
    # threshold), computed on the device
    k = (maxdim > 0 || threshold > 0) ? Int(reinterpret(Int64, Array(kd))[1]) : kmax
    du = Tuple(size(A, i) for i in inds_u)
    dv = Tuple(size(A, i) for i in inds_v)
    # U's first k columns are a prefix of its column-major block: a view;
    # V's first k rows are strided (leading dimension kmax): one gather
    Vk = k == kmax ? V : leading_rows(V, k)
    return Tensor(reshape_view(U, (du..., k)), [inds_u..., ind_s]), Tensor(s[1:k], [ind_s]),
           Tensor(reshape_view(Vk, (k, dv...)), [ind_s, inds_v...])
end

# contiguous copy of the first k rows of a column-major m x n device matrix
function leading_rows(M::HIPArray{T,2}, k::Int) where {T}
    m, n = size(M)
    out = HIPArray{T,2}(undef, (k, n))
    one = HIPArray(ones(T))
    check(ccall((:tenet_einsum, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ref{TDesc}, Ptr{Cvoid}, Ref{TDesc}, Ptr{Cvoid}, Ref{TDesc},
                 Ptr{Cvoid}, Cint, Cint),
                context().h, C_NULL, dtype_code(T), Cint(0), tdesc((k, n), (1, 2), (1, m)), M.ptr,
                tdesc((), ()), one.ptr, tdesc((k, n), (1, 2)), out.ptr, 0, 0))
    return out
end

# ---- one canonize move in one launch (generic_canonize_site!, canonize.jl:58-70) ----
# tensor_qr_thin(A) followed by binary_einsum(R, B) across `ind_bond`, both
# matricised dimensions of A at most 64; returns (Q, R B) on the bond's label
# (the reference renames the virtual index back to it), or `nothing` when the
# shapes are out of range and the two calls above are the route.
function qr_absorb(A::HIPTensor{T}, B::HIPTensor{T}, ind_bond) where {T}
    inds_a = filter(!=(ind_bond), inds(A))
    inds_b = filter(!=(ind_bond), inds(B))
    m = prod(size(A, i) for i in inds_a; init=1)
    n = size(A, ind_bond)
    (m <= 64 && n <= 64) || return nothing
    M, = matricize(A, inds_a, [ind_bond])
    Bm, _, nb = matricize(B, [ind_bond], inds_b)
    k = min(m, n)
    Q = HIPArray{T,2}(undef, (m, k))
    Bo = HIPArray{T,2}(undef, (k, nb))
    check(ccall((:tenet_qr_absorb, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Cint, Int64,
                 Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64),
                context().h, C_NULL, dtype_code(T), Cint(0), m, n, M.ptr, m, Q.ptr, m, Cint(0), nb, Bm.ptr, n,
                Bo.ptr, k))
    da = Tuple(size(A, i) for i in inds_a)
    db = Tuple(size(B, i) for i in inds_b)
    return Tensor(reshape_view(Q, (da..., k)), [inds_a..., ind_bond]),
           Tensor(reshape_view(Bo, (k, db...)), [ind_bond, inds_b...])
end

# ---- Muscle.simple_update (simple_update.jl:53-64) ----------------------------
function Muscle.simple_update(A::HIPTensor{T}, ind_a, B::HIPTensor{T}, ind_b, ind_bond, G::HIPTensor{T}, ind_op_a,
                              ind_op_b; maxdim=0, threshold=0.0, absorb=Muscle.AbsorbEqually()) where {T}
    # A as (chil, d1, chi), B as (chi, d2, chir), G as G[oa, ob, ia, ib]
    la = setdiff(inds(A), [ind_a, ind_bond])
    lb = setdiff(inds(B), [ind_b, ind_bond])
    Am, = matricize(A, vcat(la, [ind_a]), [ind_bond])
    Bm, = matricize(B, [ind_bond], vcat([ind_b], lb))
    Gm, = matricize(G, [ind_op_a, ind_op_b], [ind_a, ind_b])
    chil = prod(size(A, i) for i in la; init=1)
    chir = prod(size(B, i) for i in lb; init=1)
    d1, d2, chi = size(A, ind_a), size(B, ind_b), size(A, ind_bond)
    kmax = min(chil * d1, d2 * chir)
    Ao = HIPArray{T,1}(undef, (chil * d1 * kmax,))
    Bo = HIPArray{T,1}(undef, (kmax * d2 * chir,))
    mode = absorb isa Muscle.AbsorbEqually ? 1 : 0
    k = Ref{Int64}(0)
    check(ccall((:tenet_simple_update, libtenet), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, Int64, Int64, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cdouble, Cint, Ptr{Cvoid}, Ptr{Cvoid},
                 Ref{Int64}),
                context().h, C_NULL, dtype_code(T), 0, chil, d1, chi, d2, chir, Am.ptr, Bm.ptr, Gm.ptr, Ao.ptr,
                Bo.ptr, C_NULL, maxdim, threshold, mode, C_NULL, C_NULL, k))
    kk = k[]
    dla = Tuple(size(A, i) for i in la)
    dlb = Tuple(size(B, i) for i in lb)
    # the gate outputs come back on the input labels (simple_update.jl:66-69 renames them anyway)
    return Tensor(reshape_view(Ao, (dla..., d1, kk)), [la..., ind_a, ind_bond]),
           Tensor(reshape_view(Bo, (kk, d2, dlb...)), [ind_bond, ind_b, lb...])
end

# a column-major view of the leading prod(dims) elements of `a` (no copy;
# the view keeps `a`, or the array `a` itself views, alive)
function reshape_view(a::HIPArray{T}, dims::NTuple{N,Int}) where {T,N}
    prod(dims; init=1) <= prod(size(a); init=1) || throw(DimensionMismatch("view larger than its parent"))
    return HIPArray{T,N}(a.ptr, dims, a.owner === nothing ? a : a.owner)
end

end # module
