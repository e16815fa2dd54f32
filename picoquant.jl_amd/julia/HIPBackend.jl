# HIPBackend: PicoQuant's AbstractBackend on libpq_hip.so (MI355X / gfx950).
#
# Drop-in for InteractiveBackend{Array{ComplexF64}} (src/backends/interactive.jl):
# the nine backend verbs keep their signatures and semantics, tensors live on
# the GPU behind integer handles, and every arithmetic call goes through the
# C ABI declared in include/pq_hip.h.  contract_pair!, contract_network!,
# full_wavefunction_contraction! and the slicing helpers (src/layer2.jl,
# src/layer2/slicing.jl) run unchanged on top of it.
#
# Usage (inside the PicoQuant module, after src/backends.jl):
#     include("backends/HIPBackend.jl")
#     tn = TensorNetworkCircuit(12, HIPBackend{ComplexF64}())
#
# This file is the binding a maintainer adds to PicoQuant; it is not run in
# this repository's tests (no Julia toolchain here) -- the same calls are
# exercised through the Python mirror in picoquant.jl_amd/picoquant_hip.

const LIBPQ = get(ENV, "PQ_HIP_LIB", joinpath(@__DIR__, "..", "picoquant_hip", "libpq_hip.so"))

const PQ_OK = Cint(0)
const PQ_ERR_NOT_FOUND = Cint(3)
const PQ_C128 = Cint(0)
const PQ_C64 = Cint(1)
const PQ_MAX_RANK = 64

pq_code(::Type{ComplexF64}) = PQ_C128
pq_code(::Type{ComplexF32}) = PQ_C64

"Backend with device tensors; T = ComplexF64 (the plain method) or ComplexF32."
mutable struct HIPBackend{T} <: AbstractBackend
    ctx::Ptr{Cvoid}
    tensors::Dict{Symbol, Int64}
    metrics::Metrics

    function HIPBackend{T}(device::Integer=0) where {T}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:pq_ctx_create, LIBPQ), Cint, (Cint, Cint, Ref{Ptr{Cvoid}}), device, pq_code(T), p)
        rc == PQ_OK || error("pq_ctx_create failed ($rc): no HIP device?")
        b = new{T}(p[], Dict{Symbol, Int64}(), Metrics())
        finalizer(b) do x
            x.ctx == C_NULL || ccall((:pq_ctx_destroy, LIBPQ), Cint, (Ptr{Cvoid},), x.ctx)
            x.ctx = C_NULL
        end
        b
    end
end
HIPBackend() = HIPBackend{ComplexF64}()

function pq_check(b::HIPBackend, rc::Cint)
    rc == PQ_OK && return nothing
    msg = unsafe_string(ccall((:pq_last_error, LIBPQ), Cstring, (Ptr{Cvoid},), b.ctx))
    rc == PQ_ERR_NOT_FOUND ? throw(KeyError(msg)) : error("libpq_hip error $rc: $msg")
end

function pq_shape(b::HIPBackend, h::Int64)
    r = Ref{Cint}(0)
    dims = zeros(Int64, PQ_MAX_RANK)
    pq_check(b, ccall((:pq_tensor_shape, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Ref{Cint}, Ptr{Int64}),
                      b.ctx, h, r, dims))
    Tuple(dims[1:r[]])
end

function pq_set!(b::HIPBackend, label::Symbol, h::Int64)
    old = get(b.tensors, label, nothing)
    b.tensors[label] = h
    old === nothing || pq_check(b, ccall((:pq_tensor_free, LIBPQ), Cint, (Ptr{Cvoid}, Int64), b.ctx, old))
    nothing
end

"save_tensor_data (src/backends/interactive.jl:32-36): store a converted copy."
function save_tensor_data(b::HIPBackend{T}, label::Symbol, data::AbstractArray) where {T}
    host = Array{ComplexF64}(data)                  # the library converts to T on upload
    dims = Int64[size(host)...]
    h = Ref{Int64}(0)
    pq_check(b, ccall((:pq_tensor_upload, LIBPQ), Cint,
                      (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{ComplexF64}, Cint, Ref{Int64}),
                      b.ctx, length(dims), dims, host, PQ_C128, h))
    pq_set!(b, label, h[])
end

"load_tensor_data (src/backends/interactive.jl:44-49): the tensor, or nothing."
function load_tensor_data(b::HIPBackend{T}, label::Symbol) where {T}
    h = get(b.tensors, label, nothing)
    h === nothing && return nothing
    out = Array{T}(undef, pq_shape(b, h))
    pq_check(b, ccall((:pq_tensor_download, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Ptr{T}), b.ctx, h, out))
    out
end

"contract_tensors (src/backends/interactive.jl:60-75): C from A and B; A and B deleted."
function contract_tensors(b::HIPBackend, A::Symbol, ia::Array{Int, 1}, B::Symbol, ib::Array{Int, 1}, C::Symbol)
    ha, hb = b.tensors[A], b.tensors[B]
    hc = Ref{Int64}(0)
    la, lb = Int32.(ia), Int32.(ib)
    pq_check(b, ccall((:pq_contract, LIBPQ), Cint,
                      (Ptr{Cvoid}, Int64, Cint, Ptr{Int32}, Int64, Cint, Ptr{Int32}, Ref{Int64}),
                      b.ctx, ha, length(la), la, hb, length(lb), lb, hc))
    pq_set!(b, C, hc[])
    delete_tensor!(b, A)
    delete_tensor!(b, B)
end

"save_output (src/backends/interactive.jl:84-88): `name` refers to the node's tensor."
function save_output(b::HIPBackend, node::Symbol, name::String="result")
    h = b.tensors[node]
    n = length(pq_shape(b, h))
    perm = Int32.(collect(1:n))
    out = Ref{Int64}(0)
    pq_check(b, ccall((:pq_permute, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}, Ref{Int64}),
                      b.ctx, h, perm, out))
    pq_set!(b, Symbol(name), out[])
end

"reshape_tensor (src/backends/interactive.jl:97-102)."
function reshape_tensor(b::HIPBackend, tensor::Symbol, groups::Array{Array{Int, 1}, 1})
    h = b.tensors[tensor]
    d = pq_shape(b, h)
    dims = Int64[prod(d[g]) for g in groups]
    pq_check(b, ccall((:pq_reshape, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Cint, Ptr{Int64}),
                      b.ctx, h, length(dims), dims))
end

"permute_tensor (src/backends/interactive.jl:111-115): zero-copy strided view."
function permute_tensor(b::HIPBackend, tensor::Symbol, axes::Array{Int, 1})
    out = Ref{Int64}(0)
    pq_check(b, ccall((:pq_permute, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}, Ref{Int64}),
                      b.ctx, b.tensors[tensor], Int32.(axes), out))
    pq_set!(b, tensor, out[])
end

"decompose_tensor! (src/backends/interactive.jl:130-152): SVD split on the GPU (pq_decompose)."
function decompose_tensor!(b::HIPBackend{T}, tensor::Symbol, left_positions::Array{Int, 1},
                           right_positions::Array{Int, 1}; threshold::AbstractFloat=1e-13,
                           max_rank::Int=0, left_label::Symbol, right_label::Symbol) where {T}
    hb, hc, chi = Ref{Int64}(0), Ref{Int64}(0), Ref{Cint}(0)
    pq_check(b, ccall((:pq_decompose, LIBPQ), Cint,
                      (Ptr{Cvoid}, Int64, Cint, Ptr{Int32}, Cint, Ptr{Int32}, Cdouble, Cint,
                       Ref{Int64}, Ref{Int64}, Ref{Cint}),
                      b.ctx, b.tensors[tensor], length(left_positions), Int32.(left_positions),
                      length(right_positions), Int32.(right_positions), threshold, max_rank, hb, hc, chi))
    pq_set!(b, left_label, hb[])
    pq_set!(b, right_label, hc[])
    tensor in (left_label, right_label) || delete_tensor!(b, tensor)
    Int(chi[])
end

"delete_tensor! (src/backends/interactive.jl:159-161)."
function delete_tensor!(b::HIPBackend, label::Symbol)
    h = pop!(b.tensors, label, nothing)
    h === nothing || pq_check(b, ccall((:pq_tensor_free, LIBPQ), Cint, (Ptr{Cvoid}, Int64), b.ctx, h))
    nothing
end

"view_tensor! (src/backends/interactive.jl:169-172): zero-copy view of bond_idx at bond_range."
function view_tensor!(b::HIPBackend, view_node, node, bond_idx, bond_range)
    out = Ref{Int64}(0)
    pq_check(b, ccall((:pq_view, LIBPQ), Cint, (Ptr{Cvoid}, Int64, Cint, Int64, Int64, Ref{Int64}),
                      b.ctx, b.tensors[node], bond_idx, first(bond_range), last(bond_range), out))
    pq_set!(b, Symbol(view_node), out[])
end

# Batched path (no reference counterpart): a compiled plan (pq_plan_create,
# built by the host-side lowering) runs every slice of a sliced network and a
# batch of output bitstrings in one graph replay.

"""plan_run!(b, plan, bitstrings) -> Vector{ComplexF64}: bitstring words
(bit q = closed output qubit q's value, as add_output! takes them per
bitstring, src/layer3.jl:292-313), one run, the FP64 amplitudes of every
bitstring (out_elems each, pq_plan_stats)."""
function plan_run!(b::HIPBackend, plan::Ptr{Cvoid}, bitstrings::Vector{UInt64})
    n_out = Ref{Int64}(0)
    pq_check(b, ccall((:pq_plan_stats, LIBPQ), Cint,
                      (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}),
                      plan, C_NULL, C_NULL, C_NULL, C_NULL, n_out))
    out = Vector{Float64}(undef, 2 * n_out[])
    pq_check(b, ccall((:pq_plan_run, LIBPQ), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{UInt64}, Ptr{Float64}),
                      b.ctx, plan, length(bitstrings), bitstrings, out))
    [complex(out[2i - 1], out[2i]) for i in 1:n_out[]]
end

"""plan_cost(b, plan) -> (macs, bytes): exact multiply-adds of one run, as
the Metrics ledger counts them (src/backends.jl:38-55), and algorithmic bytes."""
function plan_cost(b::HIPBackend, plan::Ptr{Cvoid})
    c = Ref{NTuple{4, Int64}}((0, 0, 0, 0))   # pq_cost: two doubles, two int64
    pq_check(b, ccall((:pq_plan_cost, LIBPQ), Cint, (Ptr{Cvoid}, Ref{NTuple{4, Int64}}, Ptr{Float64}, Ptr{Float64}),
                      plan, c, C_NULL, C_NULL))
    (reinterpret(Float64, c[][1]), reinterpret(Float64, c[][2]))
end

export HIPBackend
