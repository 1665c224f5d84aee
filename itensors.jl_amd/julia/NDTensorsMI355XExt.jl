# NDTensorsMI355XExt -- package-extension glue that routes NDTensors'
# contraction to libitg.so (the MI355X backend) through `ccall`.
#
# Loaded as an NDTensors extension (NDTensors/Project.toml [weakdeps] +
# [extensions] pattern, as the cuTENSOR/AMDGPU extensions are). Storage on
# the device is `ItgVector{T}`, a vector of device memory owned by libitg
# (itg_malloc / itg_free); tensors are moved there with `mi355x(T)` and back
# with `cpu(T)`. No AMDGPU.jl or CUDA.jl is involved.
#
# Seams (ITensors.jl v0.9.27 / NDTensors v0.4.27):
#   dense        NDTensors.contract!(::Exposed{<:ItgVector,<:DenseTensor}, ...)
#                -- the method NDTensors/ext/NDTensorscuTENSORExt/contract.jl:15-24
#                overrides for CuArray
#   block-sparse NDTensors.contract(::BlockSparseTensor{..,<:ItgVector}, ...)
#                -- NDTensors/src/blocksparse/contract.jl:3-17, whole plan in one
#                grouped launch
#   AD           ChainRulesCore.rrule(::typeof(contract), ::ITensor, ::ITensor)
#                -- ext/ITensorsChainRulesCoreExt/itensor.jl:94-103, conj in-kernel
# On ITG_ERR_UNSUPPORTED every method falls back to the stock NDTensors path
# on host copies of the operands and copies the result back, as the cuTENSOR
# extension falls back to the generic path on CUTENSORError
# (NDTensorscuTENSORExt/contract.jl:45-53). Mixed element types are promoted
# on the device first (itg_convert), as NDTensorscuTENSORExt/contract.jl:34-41
# and dense/tensoralgebra/contract.jl:196-211 do.
module NDTensorsMI355XExt

using NDTensors
using NDTensors: DenseTensor, BlockSparseTensor, BlockOffsets, Block, Dense, BlockSparse,
    blockoffsets, inds, nnzblocks, nblocks, blockdim, data, tensor, contract_labels
using NDTensors: DiagBlockSparseTensor
using NDTensors.Expose: Exposed, unexpose
using Dictionaries: Dictionary

const libitg = get(ENV, "ITG_LIBRARY", "libitg")

# ------------------------------------------------------------------ handle
const ITG_OK = Cint(0)
const ITG_ERR_UNSUPPORTED = Cint(2)
const ITG_F32 = Cint(0)
const ITG_F64 = Cint(1)
const ITG_C64 = Cint(2)
const ITG_C128 = Cint(3)
const PRECISION = Ref{Cint}(0)  # 0 FP64 MFMA, 2 split (int8 slices)
const HANDLE = Ref{Ptr{Cvoid}}(C_NULL)

struct ItgError <: Exception
    status::Cint
    msg::String
end

function handle()
    if HANDLE[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:itg_init, libitg), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        st == ITG_OK || throw(ItgError(st, "itg_init failed"))
        HANDLE[] = h[]
    end
    return HANDLE[]
end

function check(st::Cint)
    st == ITG_OK && return nothing
    msg = unsafe_string(ccall((:itg_last_error, libitg), Cstring, (Ptr{Cvoid},), handle()))
    throw(ItgError(st, msg))
end

using_mi355x_split_precision(on::Bool = true) = (PRECISION[] = on ? Cint(2) : Cint(0))

# NDTensors.gemm_backend[] (:Auto / :BLAS / :Generic, NDTensors/src/abstractarray/
# tensoralgebra/contract.jl:15-24) selects the device kernel family of the plans
# made from here on: by shape, the grouped MFMA GEMM, or the strided kernel.
const BACKEND_CODE = Dict(:Auto => Cint(0), :BLAS => Cint(1), :Generic => Cint(2))
# The threaded block-sparse switch selects the block-pair order of the plans
# (contract_blockoffsets, NDTensors/src/blocksparse/contract.jl:42-55): the
# reference takes its threaded algorithm when the switch is on and Julia runs
# more than one thread, which fixes the output's block order.
const ITG_BS_ORDER_SEQUENTIAL = Cint(0)
const ITG_BS_ORDER_THREADED = Cint(1)
function sync_backend()
    code = get(BACKEND_CODE, NDTensors.gemm_backend[], Cint(0))
    check(ccall((:itg_set_gemm_backend, libitg), Cint, (Ptr{Cvoid}, Cint), handle(), code))
    order = NDTensors.using_threaded_blocksparse() && Threads.nthreads() > 1 ?
            ITG_BS_ORDER_THREADED : ITG_BS_ORDER_SEQUENTIAL
    check(ccall((:itg_set_blocksparse_order, libitg), Cint, (Ptr{Cvoid}, Cint), handle(), order))
    return nothing
end

eltcode(::Type{Float32}) = ITG_F32
eltcode(::Type{Float64}) = ITG_F64
eltcode(::Type{ComplexF32}) = ITG_C64
eltcode(::Type{ComplexF64}) = ITG_C128
eltcode(T) = nothing

# --------------------------------------------------------- device storage
mutable struct ItgVector{T} <: AbstractVector{T}
    ptr::Ptr{T}
    len::Int
    function ItgVector{T}(::UndefInitializer, n::Integer) where {T}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:itg_malloc, libitg), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                    handle(), n * sizeof(T), p))
        v = new{T}(Ptr{T}(p[]), n)
        finalizer(v) do x
            ccall((:itg_free, libitg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), handle(), x.ptr)
        end
        return v
    end
end
Base.size(v::ItgVector) = (v.len,)
Base.similar(::ItgVector{T}, ::Type{S}, dims::Dims{1}) where {T, S} = ItgVector{S}(undef, dims[1])
function Base.getindex(v::ItgVector{T}, i::Int) where {T}  # scalar access: slow, host copy
    r = Ref{T}()
    check(ccall((:itg_memcpy_dtoh, libitg), Cint, (Ptr{Cvoid}, Ref{T}, Ptr{T}, Csize_t),
                handle(), r, v.ptr + (i - 1) * sizeof(T), sizeof(T)))
    return r[]
end
function Base.Array(v::ItgVector{T}) where {T}
    a = Vector{T}(undef, v.len)
    check(ccall((:itg_memcpy_dtoh, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                handle(), a, v.ptr, v.len * sizeof(T)))
    return a
end
function ItgVector(a::Vector{T}) where {T}
    v = ItgVector{T}(undef, length(a))
    copyto!(v, a)
    return v
end
function Base.setindex!(v::ItgVector{T}, x, i::Int) where {T}  # scalar access: slow
    r = Ref{T}(convert(T, x))
    check(ccall((:itg_memcpy_htod, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ref{T}, Csize_t),
                handle(), v.ptr + (i - 1) * sizeof(T), r, sizeof(T)))
    return v
end
function Base.copyto!(v::ItgVector{T}, a::Vector{T}) where {T}  # host -> device
    length(a) <= v.len || throw(DimensionMismatch("source longer than destination"))
    check(ccall((:itg_memcpy_htod, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                handle(), v.ptr, a, length(a) * sizeof(T)))
    return v
end
function Base.copyto!(a::Vector{T}, v::ItgVector{T}) where {T}  # device -> host
    v.len <= length(a) || throw(DimensionMismatch("source longer than destination"))
    check(ccall((:itg_memcpy_dtoh, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                handle(), a, v.ptr, v.len * sizeof(T)))
    return a
end
function Base.copyto!(w::ItgVector{T}, v::ItgVector{T}) where {T}  # device -> device
    v.len <= w.len || throw(DimensionMismatch("source longer than destination"))
    check(ccall((:itg_memcpy_dtod, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                handle(), w.ptr, v.ptr, v.len * sizeof(T)))
    return w
end

# A widened device copy (Float32 -> Float64, real -> complex): the promotion
# by allocation of dense/tensoralgebra/contract.jl:196-211, on the device.
function promote_data(v::ItgVector{S}, ::Type{T}) where {S, T}
    S === T && return v
    w = ItgVector{T}(undef, v.len)
    check(ccall((:itg_convert, libitg), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64),
                handle(), eltcode(S), v.ptr, eltcode(T), w.ptr, v.len))
    return w
end
promote_tensor(T::DenseTensor, ::Type{E}) where {E} =
    eltype(T) === E ? T : tensor(Dense(promote_data(data(T), E)), inds(T))
promote_tensor(T::BlockSparseTensor, ::Type{E}) where {E} =
    eltype(T) === E ? T : tensor(BlockSparse(promote_data(data(T), E), blockoffsets(T)), inds(T))

mi355x(T::DenseTensor) = tensor(Dense(ItgVector(Vector(data(T)))), inds(T))
mi355x(T::BlockSparseTensor) = tensor(BlockSparse(ItgVector(Vector(data(T))), blockoffsets(T)), inds(T))
# a stored diagonal moves to the device; a uniform one (delta) is a number and stays
mi355x(T::DiagBlockSparseTensor) = data(T) isa Number ? T :
    tensor(NDTensors.DiagBlockSparse(ItgVector(Vector(data(T))), NDTensors.diagblockoffsets(T)), inds(T))
cpu(T::DenseTensor{<:Any, <:Any, <:Dense{<:Any, <:ItgVector}}) = tensor(Dense(Array(data(T))), inds(T))
cpu(T::DenseTensor) = T
cpu(T::BlockSparseTensor) = tensor(BlockSparse(Array(data(T)), blockoffsets(T)), inds(T))
cpu(T::DiagBlockSparseTensor) = data(T) isa ItgVector ?
    tensor(NDTensors.DiagBlockSparse(Array(data(T)), NDTensors.diagblockoffsets(T)), inds(T)) : T

# --------------------------------------------------------------- dense
function NDTensors.contract!(
        exposedR::Exposed{<:ItgVector, <:DenseTensor}, labelsR,
        exposedT1::Exposed{<:ItgVector, <:DenseTensor}, labelsT1,
        exposedT2::Exposed{<:ItgVector, <:DenseTensor}, labelsT2,
        α::Number = one(Bool), β::Number = zero(Bool)
    )
    R, T1, T2 = unexpose.((exposedR, exposedT1, exposedT2))
    elt = eltype(R)
    # promote the inputs to R's element type (NDTensorscuTENSORExt/contract.jl:34-41)
    if promote_type(elt, eltype(T1), eltype(T2)) !== elt
        error("In MI355X contraction, input tensors have element types `$(eltype(T1))` and " *
              "`$(eltype(T2))` while the output has element type `$(elt)`.")
    end
    code = eltcode(elt)
    st = ITG_ERR_UNSUPPORTED
    if code !== nothing
        T1 = promote_tensor(T1, elt)
        T2 = promote_tensor(T2, elt)
        sync_backend()
        st = ccall((:itg_contract_dense, libitg), Cint,
                   (Ptr{Cvoid}, Cint,
                    Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid},
                    Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid},
                    Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid},
                    Ref{ComplexF64}, Ref{ComplexF64}, Cint),
                   handle(), code,
                   ndims(R), collect(Int64, size(R)), collect(Int64, labelsR), data(R).ptr,
                   ndims(T1), collect(Int64, size(T1)), collect(Int64, labelsT1), data(T1).ptr,
                   ndims(T2), collect(Int64, size(T2)), collect(Int64, labelsT2), data(T2).ptr,
                   ComplexF64(α), ComplexF64(β), PRECISION[])
    end
    if st == ITG_ERR_UNSUPPORTED
        # the stock path on host copies (the cuTENSOR extension's fallback,
        # NDTensorscuTENSORExt/contract.jl:45-53), then the result back
        Rh = cpu(R)
        NDTensors.contract!(Rh, labelsR, cpu(T1), labelsT1, cpu(T2), labelsT2, α, β)
        copyto!(data(R), data(Rh))
        return R
    end
    check(st)
    return R
end

# -------------------------------------------------------- block sparse
function blockstructure(T::BlockSparseTensor)
    N = ndims(T)
    bo = blockoffsets(T)
    blocks = Int64[]
    offs = Int64[]
    for (b, o) in pairs(bo)
        append!(blocks, Int64.(Tuple(b)))
        push!(offs, o)
    end
    nsec = Int64[nblocks(inds(T), i) for i in 1:N]
    secd = Int64[blockdim(inds(T)[i], j) for i in 1:N for j in 1:nsec[i]]
    return length(offs), blocks, offs, nsec, secd
end

function plan_output(plan, labelsR, indsR, elt)
    nblk, nnz, npairs = Ref{Int64}(), Ref{Int64}(), Ref{Int64}()
    check(ccall((:itg_bs_plan_sizes, libitg), Cint,
                (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}), plan, nblk, nnz, npairs))
    NR = length(labelsR)
    blocksR = Vector{Int64}(undef, NR * nblk[])
    offsR = Vector{Int64}(undef, nblk[])
    check(ccall((:itg_bs_plan_blocks, libitg), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}),
                plan, blocksR, offsR))
    boffsR = BlockOffsets{NR}()
    for k in 1:nblk[]
        blk = Block(ntuple(i -> UInt(blocksR[(k - 1) * NR + i]), NR))
        insert!(boffsR, blk, Int(offsR[k]))
    end
    R = tensor(BlockSparse(ItgVector{elt}(undef, nnz[]), boffsR), indsR)
    return R, npairs[]
end

# the stock block-sparse path on host copies, result moved back
fallback_contract(T1, labels1, T2, labels2, labelsR) =
    mi355x(NDTensors.contract(cpu(T1), labels1, cpu(T2), labels2, labelsR))

function NDTensors.contract(
        T1::BlockSparseTensor{<:Any, N1, <:Any, <:BlockSparse{<:Any, <:ItgVector}}, labels1,
        T2::BlockSparseTensor{<:Any, N2, <:Any, <:BlockSparse{<:Any, <:ItgVector}}, labels2,
        labelsR = contract_labels(labels1, labels2)
    ) where {N1, N2}
    elt = promote_type(eltype(T1), eltype(T2))
    eltcode(elt) === nothing && return fallback_contract(T1, labels1, T2, labels2, labelsR)
    T1 = promote_tensor(T1, elt)
    T2 = promote_tensor(T2, elt)
    n1, b1, o1, ns1, sd1 = blockstructure(T1)
    n2, b2, o2, ns2, sd2 = blockstructure(T2)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    sync_backend()
    st = ccall((:itg_bs_plan_create, libitg), Cint,
               (Ptr{Cvoid}, Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                Cint, Ptr{Int64}, Ref{Ptr{Cvoid}}),
               handle(), N1, n1, b1, o1, ns1, sd1, collect(Int64, labels1),
               N2, n2, b2, o2, ns2, sd2, collect(Int64, labels2),
               length(labelsR), collect(Int64, labelsR), plan)
    st == ITG_ERR_UNSUPPORTED && return fallback_contract(T1, labels1, T2, labels2, labelsR)
    check(st)
    try
        indsR = NDTensors.contract_inds(inds(T1), labels1, inds(T2), labels2, labelsR)
        R, npairs = plan_output(plan[], labelsR, indsR, elt)
        if npairs > 0
            st = ccall((:itg_bs_execute, libitg), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                        Ref{ComplexF64}, Ptr{Cvoid}, Cint, Cuint),
                       handle(), plan[], eltcode(elt), data(R).ptr, data(T1).ptr, data(T2).ptr,
                       ComplexF64(1), C_NULL, PRECISION[], Cuint(0))
            st == ITG_ERR_UNSUPPORTED &&
                return fallback_contract(T1, labels1, T2, labels2, labelsR)
            check(st)
        end
        return R
    finally
        ccall((:itg_bs_plan_destroy, libitg), Cint, (Ptr{Cvoid},), plan[])
    end
end

# --------------------------------------------------------- diagonal
# BlockSparse x DiagBlockSparse (NDTensors/src/blocksparse/diagblocksparse.jl:
# 596-675): the same block-pair plan with the diagonal tensor's blocks and
# diagonal offsets, one grouped launch; a uniform diagonal (delta) passes its
# value instead of data.
function diagstructure(D::DiagBlockSparseTensor)
    N = ndims(D)
    blocks = Int64[]
    offs = Int64[]
    for (b, o) in pairs(NDTensors.diagblockoffsets(D))
        append!(blocks, Int64.(Tuple(b)))
        push!(offs, o)
    end
    nsec = Int64[nblocks(inds(D), i) for i in 1:N]
    secd = Int64[blockdim(inds(D)[i], j) for i in 1:N for j in 1:nsec[i]]
    return length(offs), blocks, offs, nsec, secd
end

function NDTensors.contract(
        T1::BlockSparseTensor{<:Any, N1, <:Any, <:BlockSparse{<:Any, <:ItgVector}}, labels1,
        T2::DiagBlockSparseTensor{<:Any, N2}, labels2,
        labelsR = contract_labels(labels1, labels2)
    ) where {N1, N2}
    uniform = data(T2) isa Number
    # a stored diagonal must be on the device (mi355x); it is promoted with T1
    elt = promote_type(eltype(T1), eltype(T2))
    if eltcode(elt) === nothing || !(uniform || data(T2) isa ItgVector)
        return fallback_contract(T1, labels1, T2, labels2, labelsR)
    end
    T1 = promote_tensor(T1, elt)
    D = uniform ? T2 : tensor(NDTensors.DiagBlockSparse(promote_data(data(T2), elt),
                                                        NDTensors.diagblockoffsets(T2)), inds(T2))
    n1, b1, o1, ns1, sd1 = blockstructure(T1)
    n2, b2, o2, ns2, sd2 = diagstructure(D)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:itg_bs_diag_plan_create, libitg), Cint,
               (Ptr{Cvoid}, Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                Cint, Ptr{Int64}, Ref{Ptr{Cvoid}}),
               handle(), N1, n1, b1, o1, ns1, sd1, collect(Int64, labels1),
               N2, n2, b2, o2, ns2, sd2, collect(Int64, labels2),
               length(labelsR), collect(Int64, labelsR), plan)
    st == ITG_ERR_UNSUPPORTED && return fallback_contract(T1, labels1, T2, labels2, labelsR)
    check(st)
    try
        indsR = NDTensors.contract_inds(inds(T1), labels1, inds(D), labels2, labelsR)
        R, npairs = plan_output(plan[], labelsR, indsR, elt)
        if npairs > 0
            st = ccall((:itg_bs_diag_execute, libitg), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                        Ptr{ComplexF64}, Ref{ComplexF64}),
                       handle(), plan[], eltcode(elt), data(R).ptr, data(T1).ptr,
                       uniform ? C_NULL : data(D).ptr,
                       uniform ? Ref(ComplexF64(data(D))) : C_NULL, ComplexF64(1))
            st == ITG_ERR_UNSUPPORTED &&
                return fallback_contract(T1, labels1, T2, labels2, labelsR)
            check(st)
        end
        return R
    finally
        ccall((:itg_bs_plan_destroy, libitg), Cint, (Ptr{Cvoid},), plan[])
    end
end

# the diagonal tensor first: the other operand's labels come first
# (diagblocksparse.jl:626-634)
function NDTensors.contract(
        T1::DiagBlockSparseTensor, labels1,
        T2::BlockSparseTensor{<:Any, N2, <:Any, <:BlockSparse{<:Any, <:ItgVector}}, labels2,
        labelsR = contract_labels(labels2, labels1)
    ) where {N2}
    return NDTensors.contract(T2, labels2, T1, labels1, labelsR)
end

# --------------------------------------------------------------- combiner
# Block-sparse combiner contraction (NDTensors/src/blocksparse/combiner.jl):
# NDTensors' own bookkeeping chooses the output blocks; each block's
# permutedims! into (or copyto! out of) a range of a merged block becomes one
# descriptor of a single grouped strided copy on the device (itg_bs_copy,
# bounds-checked on the host by the library).
function itg_copy!(R::ItgVector{T}, S::ItgVector{T}, descs) where {T}
    isempty(descs) && return R
    nm = maximum(d -> length(d[1]), descs)
    pad(v, x) = vcat(collect(Int64, v), fill(Int64(x), nm - length(v)))
    dims = reduce(vcat, [pad(d[1], 1) for d in descs])
    ss = reduce(vcat, [pad(d[3], 0) for d in descs])
    ds = reduce(vcat, [pad(d[5], 0) for d in descs])
    so = Int64[d[2] for d in descs]
    dof = Int64[d[4] for d in descs]
    check(ccall((:itg_bs_copy, libitg), Cint,
                (Ptr{Cvoid}, Cint, Int64, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64),
                handle(), eltcode(T), length(descs), nm, dims, so, ss, dof, ds,
                S.ptr, length(S), R.ptr, length(R)))
    return R
end

colmajor_strides(d) = Int64[prod(d[1:(k - 1)]; init = 1) for k in 1:length(d)]

# permutedims_combine (NDTensors/src/blocksparse/blocksparsetensor.jl:571-630)
# for device storage: the same per-block offsets, one copy launch.
function NDTensors.permutedims_combine(
        T::BlockSparseTensor{ElT, N, <:Any, <:BlockSparse{<:Any, <:ItgVector}}, is,
        perm::NTuple{N, Int}, combdims::NTuple{NC, Int}, blockperm::Vector{Int},
        blockcomb::Vector{Int}
    ) where {ElT, N, NC}
    Rh = NDTensors.permutedims_combine_output(cpu(T), is, perm, combdims, blockperm, blockcomb)
    R = tensor(BlockSparse(ItgVector(zeros(ElT, length(data(Rh)))), blockoffsets(Rh)), inds(Rh))
    inds_perm = NDTensors.permute(inds(T), perm)
    combdims_perm = NDTensors.TupleTools.sort(NDTensors._permute_combdims(combdims, perm))
    cloc = minimum(combdims_perm)
    ind_comb = NDTensors.permuteblocks(NDTensors.outer(NDTensors.getindices(inds_perm, combdims_perm)...), blockperm)
    iperm = invperm(perm)
    descs = Any[]
    for (b, off) in pairs(blockoffsets(T))
        bc = NDTensors.perm_block(NDTensors.combine_dims(NDTensors.permute(b, perm), inds_perm, combdims_perm), cloc, blockperm)
        bin = bc[cloc]
        nb = blockcomb[bin]
        o, k = 0, 1
        while bin - k > 0 && blockcomb[bin - k] == nb
            o += blockdim(ind_comb, bin - k)
            k += 1
        end
        bnew = NDTensors.setindex(bc, nb, cloc)
        dT = NDTensors.blockdims(T, b)
        dR = NDTensors.blockdims(R, bnew)
        sR = colmajor_strides(dR)
        # permuted T mode j sits at R mode j (before cloc), inside the
        # combined range (combdims_perm), or at R mode j - NC + 1 (after)
        dperm = NDTensors.permute(dT, perm)
        dstr = zeros(Int64, N)
        acc = 1
        for j in 1:N
            if j in combdims_perm
                dstr[perm[j]] = acc
                acc *= dperm[j]
            else
                dstr[perm[j]] = sR[j < cloc ? j : j - NC + 1]
            end
        end
        push!(descs, (collect(Int64, dT), Int64(off), colmajor_strides(dT),
                      Int64(NDTensors.offset(R, bnew) + o), dstr))
    end
    itg_copy!(data(R), data(T), descs)
    return R
end

end # module
