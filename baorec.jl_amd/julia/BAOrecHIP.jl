"""
    BAOrecHIP

Julia binding of libbaorec_hip.so (MI355X / gfx950): the reconstruction entry
points of BAOrec.jl (src/recon.jl) with the same names, arguments and results,
computed on the GPU through the C ABI of include/baorec.h.

Drop-in use: `using BAOrec, BAOrecHIP`, then call `BAOrecHIP.run!`,
`BAOrecHIP.reconstructed_positions`, ... with the recon structs of BAOrec
(or `import BAOrecHIP: run!, read_shifts, ...`).  Host `Array{Float32}`
arguments are copied to and from the GPU by the library; device arrays (e.g.
AMDGPU.jl `ROCArray{Float32}`) are passed as device pointers and used in place.
"""
module BAOrecHIP

import BAOrec
using BAOrec: IterativeRecon, MultigridRecon, AbstractRecon
using Libdl

const libbaorec = get(ENV, "BAOREC_HIP_LIB", joinpath(@__DIR__, "..", "libbaorec_hip.so"))

const MEM_HOST = Cint(0)
const MEM_DEVICE = Cint(1)
const ITERATIVE = Cint(0)
const MULTIGRID = Cint(1)
const FIELD = Dict(:disp => Cint(0), :rsd => Cint(1), :sum => Cint(2))
const ERR_OUT_OF_BOX = Cint(-5)

# Mirrors struct baorec_params (same field order and C layout).
mutable struct Params
    bias::Float32
    f::Float32
    smoothing_radius::Float32
    box_size::NTuple{3,Float32}
    box_min::NTuple{3,Float32}
    has_los::Int32
    los::NTuple{3,Float32}
    n_iter::Int32
    beta::Float32
    jacobi_damping_factor::Float32
    jacobi_niterations::Int32
    vcycle_niterations::Int32
    mas::Int32
    mg_schedule::Int32
    tolerance::Float64
    max_vcycles::Int32
    rb_sweeps::Int32
    fp64_coarse_max::Int32
    box_pad::Float32
    ran_min::Float32
    mutate_inputs::Int32
    vcycles_done::Int32
    final_residual::Float64
    final_change::Float64
    Params() = new()
end

mutable struct Context
    ptr::Ptr{Cvoid}
end

function Context(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:baorec_create, libbaorec), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device)
    st == 0 || error("baorec_create: " * unsafe_string(ccall((:baorec_last_error, libbaorec), Cstring,
                                                             (Ptr{Cvoid},), C_NULL)))
    ctx = Context(r[])
    finalizer(c -> ccall((:baorec_destroy, libbaorec), Cint, (Ptr{Cvoid},), c.ptr), ctx)
end

const _default = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(_default[], (_default[] = Context(0)))

struct BaorecError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::BaorecError) = print(io, "BAOrecHIP error ", e.code, ": ", e.msg)

function check(ctx::Context, st::Integer)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:baorec_last_error, libbaorec), Cstring, (Ptr{Cvoid},), ctx.ptr))
    st == ERR_OUT_OF_BOX && throw(BoundsError(msg))   # the reference throws BoundsError here
    throw(BaorecError(st, msg))
end

# Where an array lives: Julia Arrays are host memory, anything else exposing a
# raw pointer (ROCArray) is taken as device memory on the context's GPU.
_mem(::Array) = MEM_HOST
_mem(::Nothing) = MEM_HOST
_mem(_) = MEM_DEVICE
_ptr(::Nothing) = Ptr{Float32}(C_NULL)
_ptr(a) = Ptr{Float32}(UInt(pointer(a)))
function _mem(arrays...)
    ms = unique(_mem(a) for a in arrays if a !== nothing)
    length(ms) <= 1 || throw(ArgumentError("mix of host and device arrays"))
    isempty(ms) ? MEM_HOST : ms[1]
end

_t3(v) = (Float32(v[1]), Float32(v[2]), Float32(v[3]))

function params(recon::AbstractRecon)
    p = Params()
    ccall((:baorec_params_default, libbaorec), Cvoid, (Ref{Params}, Cfloat, Cfloat, Cfloat), p,
          recon.bias, recon.f, recon.smoothing_radius)
    p.beta = Float32(recon.β)
    if recon.box_size !== nothing
        p.box_size = _t3(recon.box_size)
        p.box_min = _t3(recon.box_min)
    end
    if recon.los !== nothing
        p.has_los = 1
        p.los = _t3(recon.los)
    end
    if recon isa IterativeRecon
        p.n_iter = recon.n_iter
    else
        p.jacobi_damping_factor = recon.jacobi_damping_factor
        p.jacobi_niterations = recon.jacobi_niterations
        p.vcycle_niterations = recon.vcycle_niterations
    end
    p
end

_grid(gs) = Int32[gs[1], gs[2], gs[3]]
_kind(::IterativeRecon) = ITERATIVE
_kind(::MultigridRecon) = MULTIGRID
_similar_grid(x::Array, gs) = zeros(Float32, gs...)
_similar_grid(x, gs) = fill!(similar(x, Float32, gs...), 0)

const RUNARGS = (Ptr{Cvoid}, Ref{Params}, Ptr{Int32}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                 Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int64)

# run!(recon, grid_size, x,y,z,w[, rx,ry,rz,rw]): src/recon.jl:134-180, 215-261
function run!(recon::AbstractRecon, grid_size::Tuple{Int,Int,Int}, x, y, z, w,
              rx = nothing, ry = nothing, rz = nothing, rw = nothing; ctx::Context = default_context())
    p = params(recon)
    mem = _mem(x, y, z, w, rx, ry, rz, rw)
    nr = rx === nothing ? 0 : length(rx)
    g = _grid(grid_size)
    if recon isa IterativeRecon
        st = ccall((:baorec_run_iterative, libbaorec), Cint, RUNARGS, ctx.ptr, p, g, mem, _ptr(x), _ptr(y),
                   _ptr(z), _ptr(w), length(x), _ptr(rx), _ptr(ry), _ptr(rz), _ptr(rw), nr)
    else
        st = ccall((:baorec_run_multigrid, libbaorec), Cint, RUNARGS, ctx.ptr, p, g, mem, _ptr(x), _ptr(y),
                   _ptr(z), _ptr(w), length(x), _ptr(rx), _ptr(ry), _ptr(rz), _ptr(rw), nr)
    end
    check(ctx, st)
    if nr > 0   # setup_box(randoms, pad) overwrote the box (src/recon.jl:172, 253)
        recon.box_size = collect(p.box_size)
        recon.box_min = collect(p.box_min)
    end
    out = _similar_grid(x, grid_size)
    check(ctx, ccall((:baorec_result_copy, libbaorec), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}), ctx.ptr,
                     _mem(out), _ptr(out)))
    recon.result_cache = out
end

const GRIDARGS = (Ptr{Cvoid}, Ref{Params}, Ptr{Int32}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                  Ptr{Float32}, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                  Ptr{Float32}, Int64)

for (jl, c) in ((:setup_overdensity!, :baorec_setup_overdensity),
                (:reconstructed_overdensity!, :baorec_reconstructed_overdensity),
                (:reconstructed_potential!, :baorec_reconstructed_potential))
    @eval function $jl(field::AbstractArray{Float32,3}, recon::AbstractRecon, x, y, z, w,
                       rx = nothing, ry = nothing, rz = nothing, rw = nothing;
                       ctx::Context = default_context())
        p = params(recon)
        nr = rx === nothing ? 0 : length(rx)
        st = ccall(($(QuoteNode(c)), libbaorec), Cint, GRIDARGS, ctx.ptr, p, _grid(size(field)),
                   _mem(field, x, y, z, w, rx, ry, rz, rw), _ptr(field), _ptr(x), _ptr(y), _ptr(z), _ptr(w),
                   length(x), _ptr(rx), _ptr(ry), _ptr(rz), _ptr(rw), nr)
        check(ctx, st)
        field
    end
end

const SHIFTARGS = (Ptr{Cvoid}, Ref{Params}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int64,
                   Ptr{Float32}, Ptr{Int32}, Cint, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32})

function _shifts(cfun::Symbol, recon, x, y, z, field3d, field::Symbol, ctx)
    p = params(recon)
    out = (similar(x), similar(y), similar(z))
    # field3d === nothing: the context's cached result (its displacement grids
    # are kept, so repeated read-outs of one result only gather)
    grid = field3d === nothing ? size(recon.result_cache) : size(field3d)
    mem = _mem(x, y, z, field3d, out...)
    st = ccall(Libdl.dlsym(Libdl.dlopen(libbaorec), cfun), Cint, SHIFTARGS, ctx.ptr, p, mem, _ptr(x),
               _ptr(y), _ptr(z), length(x), _ptr(field3d), _grid(grid), _kind(recon), FIELD[field],
               _ptr(out[1]), _ptr(out[2]), _ptr(out[3]))
    check(ctx, st)
    out
end

# read_shifts(recon, x,y,z, δ_r; field): src/recon.jl:264-364
read_shifts(recon::AbstractRecon, x, y, z, field3d; field = :disp, ctx::Context = default_context()) =
    _shifts(:baorec_read_shifts, recon, x, y, z, field3d, field, ctx)

# reconstructed_positions(recon, x,y,z[, δ_r]; field): src/recon.jl:366-388.
# The 4-argument form reads the result of the last run! on the context.
reconstructed_positions(recon::AbstractRecon, x, y, z, field3d; field = :disp,
                        ctx::Context = default_context()) =
    _shifts(:baorec_reconstructed_positions, recon, x, y, z, field3d, field, ctx)
reconstructed_positions(recon::AbstractRecon, x, y, z; field = :disp, ctx::Context = default_context()) =
    _shifts(:baorec_reconstructed_positions, recon, x, y, z, nothing, field, ctx)

end # module
