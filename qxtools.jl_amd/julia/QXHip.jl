"""
    QXHip

Julia binding of libqxhip.so (C ABI in include/qxhip.h): the GPU executor for the
compute graphs QXTools builds (`build_compute_graph`, src/compute_graph/compute_graph.jl:15-98)
and for `single_amplitude` (src/simulation.jl:86-91).

Source only in this repository (no Julia toolchain in the build image); every entry
point below is a thin `ccall` over the C ABI that the Python harness exercises.
"""
module QXHip

using ChainRulesCore
using JLD2
using QXTools: TensorNetworkCircuit, build_compute_graph, generate_dsl_files

const libqxhip = get(ENV, "QXHIP_LIB", joinpath(@__DIR__, "..", "qxhip", "libqxhip.so"))

const C128 = Cint(0)
const C64_4M = Cint(1)
const C64_3M = Cint(2)
const C64_BF16X3 = Cint(3)

struct QxError <: Exception
    status::Int32
    msg::String
end
Base.showerror(io::IO, e::QxError) = print(io, "qxhip error ", e.status, ": ", e.msg)

function check(status::Int32)
    if status != 0
        msg = unsafe_string(ccall((:qx_last_error, libqxhip), Cstring, ()))
        throw(QxError(status, msg))
    end
    nothing
end

"""One device and one HIP stream; calls on a context must be serialised."""
mutable struct Context
    handle::Ptr{Cvoid}
    lock::ReentrantLock
    function Context(device::Integer=0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qx_ctx_create, libqxhip), Int32, (Cint, Ptr{Ptr{Cvoid}}), device, h))
        ctx = new(h[], ReentrantLock())
        finalizer(c -> ccall((:qx_ctx_destroy, libqxhip), Cvoid, (Ptr{Cvoid},), c.handle), ctx)
        ctx
    end
end

"""A .qx program with its tensors resident on the device."""
mutable struct Plan
    handle::Ptr{Cvoid}
    ctx::Context
    num_chars::Int
    result_elems::Int
    num_slices::Int
    keys::Vector{Tuple{Symbol,Int}}     # data keys and element counts, gradient layout
end

"""
    Plan(ctx, dsl::String, data::Dict{Symbol,<:AbstractArray}; precision=C128, batch=256)

`dsl` is the text of a .qx file (QXContexts' DSL, docs/src/users_guide.md:37-170);
`data` maps each `load` key to its tensor (the dictionary `convert(Dict, tc)` of a
TensorCache, src/compute_graph/tensor_cache.jl:109-111).
"""
function Plan(ctx::Context, dsl::String, data::AbstractDict; precision::Cint=C128, batch::Integer=256,
              arena_mb::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    lock(ctx.lock) do
        check(ccall((:qx_plan_from_dsl, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Csize_t, Cint, Ptr{Ptr{Cvoid}}),
                    ctx.handle, dsl, sizeof(dsl), precision, h))
        check(ccall((:qx_plan_set_option, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Int64), h[], "batch", batch))
        arena_mb > 0 && check(ccall((:qx_plan_set_option, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Int64),
                                    h[], "arena_mb", arena_mb))
        for (k, v) in data
            z = convert(Vector{ComplexF64}, vec(v))        # column-major, interleaved re/im
            check(ccall((:qx_plan_set_tensor, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Ptr{ComplexF64}, Int64),
                        h[], String(k), z, length(z)))
        end
        check(ccall((:qx_plan_finalize, libqxhip), Int32, (Ptr{Cvoid},), h[]))
    end
    nc, re, ns = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    ks = lock(ctx.lock) do
        check(ccall((:qx_plan_info, libqxhip), Int32,
                    (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                    h[], nc, re, ns, C_NULL, C_NULL))
        query_keys(h[])
    end
    p = Plan(h[], ctx, nc[], re[], ns[], ks)
    finalizer(q -> ccall((:qx_plan_destroy, libqxhip), Cvoid, (Ptr{Cvoid},), q.handle), p)
    p
end

"""
    Plan(ctx, cg; precision=C128, batch=256)

A plan from the compute graph QXTools builds (`build_compute_graph`,
src/compute_graph/compute_graph.jl:15-98): the graph is written in its saved form (the
.qx DSL text and the .jld2 tensors, src/simulation.jl:70-76) and read back, so the
device runs exactly what QXContexts would load.
"""
function Plan(ctx::Context, cg; precision::Cint=C128, batch::Integer=256, arena_mb::Integer=0)
    mktempdir() do dir
        prefix = joinpath(dir, "plan")
        generate_dsl_files(cg, prefix; force=true)
        dsl = read(prefix * ".qx", String)
        data = jldopen(prefix * ".jld2", "r") do f
            Dict{String,Any}(k => f[k] for k in keys(f))
        end
        Plan(ctx, dsl, data; precision=precision, batch=batch, arena_mb=arena_mb)
    end
end

# data keys and sizes of a plan handle, read once at construction (under the lock)
function query_keys(h::Ptr{Cvoid})
    nk = Ref{Int64}(0)
    check(ccall((:qx_plan_info, libqxhip), Int32,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                h, C_NULL, C_NULL, C_NULL, nk, C_NULL))
    buf = Vector{UInt8}(undef, 512)
    n = Ref{Int64}(0)
    out = Tuple{Symbol,Int}[]
    for i in 0:nk[]-1
        check(ccall((:qx_plan_key, libqxhip), Int32, (Ptr{Cvoid}, Int64, Ptr{UInt8}, Csize_t, Ptr{Int64}),
                    h, i, buf, length(buf), n))
        push!(out, (Symbol(unsafe_string(pointer(buf))), n[]))
    end
    out
end

"""
    update_tensor!(p, key, value)

New values of one data key after the plan is built (`qx_plan_update_tensor`): same
device buffer, so the lowering, the tuned kernels and captured graphs are kept.
"""
function update_tensor!(p::Plan, key, value::AbstractArray)
    z = convert(Vector{ComplexF64}, vec(value))               # column-major, interleaved re/im
    lock(p.ctx.lock) do
        check(ccall((:qx_plan_update_tensor, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Ptr{ComplexF64}, Int64),
                    p.handle, String(key), z, length(z)))
    end
    nothing
end

"""Upload every entry of `data` (key => tensor) into the plan's device buffers."""
function update_tensors!(p::Plan, data::AbstractDict)
    for (k, v) in data
        update_tensor!(p, k, v)
    end
    nothing
end

function packbits(p::Plan, bitstrings::AbstractVector{<:AbstractString})
    all(length(b) >= p.num_chars for b in bitstrings) || throw(ArgumentError("bitstring too short"))
    join(first(b, p.num_chars) for b in bitstrings)
end

"""
    amplitudes(p, bitstrings; slices=1:p.num_slices)

Amplitudes of many bitstrings summed over a slice range: one call replaces the
per-bitstring loop of `run_simulation` (src/simulation.jl:129-131) and the
bitstring x slice loop of `QXContexts.execute` (bin/qxrun.jl:83-87).
"""
function amplitudes(p::Plan, bitstrings::AbstractVector{<:AbstractString}; slices::UnitRange=1:p.num_slices)
    nb = length(bitstrings)
    out = Vector{ComplexF64}(undef, nb * p.result_elems)
    bits = packbits(p, bitstrings)
    lock(p.ctx.lock) do
        check(ccall((:qx_amplitudes, libqxhip), Int32, (Ptr{Cvoid}, Cstring, Int64, Int64, Int64, Ptr{ComplexF64}),
                    p.handle, bits, nb, first(slices) - 1, last(slices), out))
    end
    p.result_elems == 1 ? out : reshape(out, p.result_elems, nb)
end

"""Keys of the plan's data in the order qx_amplitudes_vjp writes gradients."""
keys_and_sizes(p::Plan) = p.keys

"""grad[key] = sum_b cot_b conj(d amp_b / d data[key]) for every data key."""
function amplitudes_vjp(p::Plan, bitstrings::AbstractVector{<:AbstractString}, cot::AbstractVector{<:Complex})
    ks = keys_and_sizes(p)
    g = Vector{ComplexF64}(undef, sum(last, ks))
    c = convert(Vector{ComplexF64}, cot)
    bits = packbits(p, bitstrings)
    lock(p.ctx.lock) do
        check(ccall((:qx_amplitudes_vjp, libqxhip), Int32,
                    (Ptr{Cvoid}, Cstring, Int64, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    p.handle, bits, length(bitstrings), c, g))
    end
    out = Dict{Symbol,Vector{ComplexF64}}()
    o = 0
    for (k, n) in ks
        out[k] = g[o+1:o+n]
        o += n
    end
    out
end

"""
    norm_loss_and_gradient(p, bitstrings) -> (loss, amplitudes, Dict(key => gradient))

The loss sum |amp|^2 over the bitstrings, the amplitudes and the loss's gradient with
respect to every data key in one pass (`qx_norm_vjp`: the cotangent 2 amp is formed on
the device from the same pass's forward).  Throws when slice variables would loop;
use `amplitudes` + `amplitudes_vjp(p, bitstrings, 2 .* amps)` there.
"""
function norm_loss_and_gradient(p::Plan, bitstrings::AbstractVector{<:AbstractString})
    ks = keys_and_sizes(p)
    g = Vector{ComplexF64}(undef, sum(last, ks))
    nb = length(bitstrings)
    amps = Vector{ComplexF64}(undef, nb * p.result_elems)
    bits = packbits(p, bitstrings)
    lock(p.ctx.lock) do
        check(ccall((:qx_norm_vjp, libqxhip), Int32,
                    (Ptr{Cvoid}, Cstring, Int64, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    p.handle, bits, nb, amps, g))
    end
    out = Dict{Symbol,Vector{ComplexF64}}()
    o = 0
    for (k, n) in ks
        out[k] = g[o+1:o+n]
        o += n
    end
    sum(abs2, amps), (p.result_elems == 1 ? amps : reshape(amps, p.result_elems, nb)), out
end

"""
    gpu_amplitudes(data, p, bitstrings)

Amplitudes as a function of the tensor data, differentiable through ChainRulesCore.
`data` (key => tensor) is uploaded into the plan's buffers first
(`qx_plan_update_tensor`), so every call contracts the values it is given -- the
parameters of a variational loop change between calls without rebuilding the plan.
The pullback returns, per key, the gradient of the same contraction.
"""
function gpu_amplitudes(data::AbstractDict, p::Plan, bitstrings)
    update_tensors!(p, data)
    amplitudes(p, bitstrings)
end

function ChainRulesCore.rrule(::typeof(gpu_amplitudes), data::AbstractDict, p::Plan, bitstrings)
    y = gpu_amplitudes(data, p, bitstrings)
    function pullback(ybar)
        # the device holds `data` unless another call uploaded different values since
        update_tensors!(p, data)
        g = amplitudes_vjp(p, bitstrings, vec(unthunk(ybar)))
        dd = Dict(k => reshape(g[Symbol(k)], size(v)) for (k, v) in data)
        return NoTangent(), Tangent{typeof(data)}(dd), NoTangent(), NoTangent()
    end
    return y, pullback
end

# plans built at the reference's seam, per (network, plan) object
const SEAM_PLANS = IdDict{Any,Plan}()

"""
    single_amplitude_gpu(ctx, tnc, plan, amplitude=nothing) -> ComplexF64

Drop-in for `single_amplitude(tnc, plan, amplitude)` (src/simulation.jl:86-91): the
network and plan are lowered once with `build_compute_graph` (no sliced bonds) and
kept on the device for later calls with the same objects; `amplitude` fixes the
outputs as `add_output!` does (`nothing`: the all-zero string, the first element of
the open network's result).
"""
function single_amplitude_gpu(ctx::Context, tnc::TensorNetworkCircuit, plan::Vector{NTuple{3,Symbol}},
                              amplitude::Union{String,Nothing}=nothing)
    key = (tnc, plan)
    p = get!(SEAM_PLANS, key) do
        Plan(ctx, build_compute_graph(tnc, plan, Vector{Vector{Symbol}}()); batch=1)
    end
    s = amplitude === nothing ? "0"^p.num_chars : amplitude
    amplitudes(p, [s])[1]
end

"""
    amplitudes_gpu(ctx, cg, bitstrings)

Amplitudes of many bitstrings from a `ComputeGraph` (src/compute_graph/compute_graph.jl:97),
summed over all slices, in one batched device call.
"""
amplitudes_gpu(ctx::Context, cg, bitstrings::AbstractVector{<:AbstractString}) =
    amplitudes(Plan(ctx, cg), bitstrings)

end # module
