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
    mg_warm_start::Int32
    mg_seed_steps::Int32
    vcycles_done::Int32
    final_residual::Float64
    final_change::Float64
    seed_residual::Float64
    rhs_shift::Float64
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
        o = MG_OPTIONS[]
        if o !== nothing   # extension: RBGS solve and its warm starts
            p.mg_schedule = o.schedule === :rbgs ? 1 : 0
            p.tolerance = o.tolerance
            p.mg_warm_start = Dict(:none => 0, :spectral => 1, :learned => 2, :refined => 3)[o.warm_start]
            p.mg_seed_steps = o.seed_steps
        end
    end
    p
end

# Extension (no reference counterpart): MultigridRecon solves through the RBGS
# schedule to `tolerance`, started cold (:none), from the spectral seed, from
# the refined spectral seed or from the learned seed (the network loaded with
# warmstart_load).  `nothing` restores the reference's fmg schedule.
const MG_OPTIONS = Ref{Any}(nothing)
multigrid_solver!(; schedule::Symbol = :rbgs, tolerance::Float64 = 1e-6, warm_start::Symbol = :none,
                  seed_steps::Integer = 3) =
    (MG_OPTIONS[] = (schedule = schedule, tolerance = tolerance, warm_start = warm_start,
                     seed_steps = Int32(seed_steps)))
reference_multigrid!() = (MG_OPTIONS[] = nothing)

# The warm-start network of the context (baorec_warmstart_load / _save).
warmstart_load(path::AbstractString; ctx::Context = default_context()) =
    check(ctx, ccall((:baorec_warmstart_load, libbaorec), Cint, (Ptr{Cvoid}, Cstring), ctx.ptr, path))
warmstart_save(path::AbstractString; ctx::Context = default_context()) =
    check(ctx, ccall((:baorec_warmstart_save, libbaorec), Cint, (Ptr{Cvoid}, Cstring), ctx.ptr, path))

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
    gen = Ref{UInt64}(0)
    check(ctx, ccall((:baorec_result_info, libbaorec), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ref{Int32}, Ref{UInt64}),
                     ctx.ptr, zeros(Int32, 3), Ref{Int32}(0), gen))
    check(ctx, ccall((:baorec_result_copy, libbaorec), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, UInt64),
                     ctx.ptr, _mem(out), _ptr(out), gen[]))
    recon.result_cache = out
    RESULT_GEN[objectid(recon)] = (ctx.ptr, gen[])
    out
end

# The context keeps the result of its last run! only.  Each recon remembers the
# generation of its own; a read-out of a recon whose result a later run! on the
# same context replaced passes the recon's own result_cache as the field, as the
# reference reads recon.result_cache (src/recon.jl:382-388).
const RESULT_GEN = Dict{UInt64,Tuple{Ptr{Cvoid},UInt64}}()

function _current_result(recon, ctx)
    key = get(RESULT_GEN, objectid(recon), nothing)
    key === nothing && return false
    gen = Ref{UInt64}(0)
    st = ccall((:baorec_result_info, libbaorec), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ref{Int32}, Ref{UInt64}),
               ctx.ptr, zeros(Int32, 3), Ref{Int32}(0), gen)
    st == 0 && key == (ctx.ptr, gen[])
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
    # are kept, so repeated read-outs of one result only gather), or the
    # recon's own result_cache when a later run! replaced the context's
    if field3d === nothing && !_current_result(recon, ctx)
        recon.result_cache === nothing && error("no cached result (run! first)")
        field3d = recon.result_cache
    end
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

# ------------------------------------------------------------ batches

# A driver loop of run!(recon, grid, x,y,z,w) + reconstructed_positions(...;
# field = :sum) over periodic-box catalogs laid end to end: catalog c is
# x[offsets[c]+1 : offsets[c+1]] (offsets has ncat + 1 entries, offsets[1] == 0).
# Returns the reconstructed positions (bitwise those of the loop); x, y, z are
# wrapped in place as run! wraps them.
function run_batch!(recon::IterativeRecon, grid_size::Tuple{Int,Int,Int}, x, y, z, w,
                    offsets::Vector{Int64}; ctx::Context = default_context())
    out = (similar(x), similar(y), similar(z))
    st = ccall((:baorec_run_batch, libbaorec), Cint,
               (Ptr{Cvoid}, Ref{Params}, Ptr{Int32}, Cint, Cint, Ptr{Int64}, Ptr{Float32}, Ptr{Float32},
                Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}),
               ctx.ptr, params(recon), _grid(grid_size), _mem(x, y, z, w, out...), length(offsets) - 1,
               offsets, _ptr(x), _ptr(y), _ptr(z), _ptr(w), _ptr(out[1]), _ptr(out[2]), _ptr(out[3]))
    check(ctx, st)
    out
end

# Times the scatter tiles and strided-pass kernels on this catalog and keeps
# the fastest in the context (extension; results unchanged).
function autotune!(recon::AbstractRecon, grid_size::Tuple{Int,Int,Int}, x, y, z, w;
                   ctx::Context = default_context())
    check(ctx, ccall((:baorec_autotune, libbaorec), Cint,
                     (Ptr{Cvoid}, Ref{Params}, Ptr{Int32}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                      Ptr{Float32}, Int64),
                     ctx.ptr, params(recon), _grid(grid_size), _mem(x, y, z, w), _ptr(x), _ptr(y), _ptr(z),
                     _ptr(w), length(x)))
end
tune_save(path::AbstractString; ctx::Context = default_context()) =
    check(ctx, ccall((:baorec_tune_save, libbaorec), Cint, (Ptr{Cvoid}, Cstring), ctx.ptr, path))
tune_load(path::AbstractString; ctx::Context = default_context()) =
    check(ctx, ccall((:baorec_tune_load, libbaorec), Cint, (Ptr{Cvoid}, Cstring), ctx.ptr, path))

# ------------------------------------------------------------ cosmology and sky

# Mirrors struct baorec_cosmology; Cosmology(BAOrec.Cosmology) copies the
# reference's fields (src/cosmo.jl:22-62).
mutable struct CosmologyC
    omega_b::Float32
    omega_c::Float32
    h::Float32
    Neff::Float32
    Omega_k0::Float32
    T_cmb::Float32
    w0::Float32
    wa::Float32
    z_tab_min::Float64
    z_tab_max::Float64
    z_tab_num::Int32
    CosmologyC() = new()
end

# The parameters of a BAOrec.Cosmology (src/cosmo.jl:22-62; the derived
# densities are recomputed by the library in the same precision)
function cosmology(c::BAOrec.Cosmology)
    k = CosmologyC()
    ccall((:baorec_cosmology_default, libbaorec), Cvoid, (Ref{CosmologyC},), k)
    k.omega_b, k.omega_c, k.h, k.Neff = c.ω_b, c.ω_c, c.h, c.Neff
    k.Omega_k0, k.T_cmb, k.w0, k.wa = c.Ω_k₀, c.T_cmb, c.w0, c.wa
    k.z_tab_min, k.z_tab_max, k.z_tab_num = c.z_tab_min, c.z_tab_max, c.z_tab_num
    k
end

# comoving_distance_interp(c)(z), src/cosmo.jl:86-93 (Mpc)
function comoving_distance(z::Vector{Float64}, c::CosmologyC; ctx::Context = default_context())
    r = similar(z)
    check(ctx, ccall((:baorec_comoving_distance, libbaorec), Cint,
                     (Ptr{Cvoid}, Ref{CosmologyC}, Cint, Ptr{Float64}, Int64, Ptr{Float64}),
                     ctx.ptr, c, MEM_HOST, z, length(z), r))
    r
end

function _sky(cfun::Symbol, a, b, d, c::CosmologyC, ctx)
    out = (similar(a), similar(b), similar(d))
    st = ccall(Libdl.dlsym(Libdl.dlopen(libbaorec), cfun), Cint,
               (Ptr{Cvoid}, Ref{CosmologyC}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int64,
                Ptr{Float32}, Ptr{Float32}, Ptr{Float32}),
               ctx.ptr, c, _mem(a, b, d, out...), _ptr(a), _ptr(b), _ptr(d), length(a), _ptr(out[1]),
               _ptr(out[2]), _ptr(out[3]))
    check(ctx, st)
    out
end
# the drivers' sky_to_cartesian / cartesian_to_sky (examples/lightcone.jl:30-80)
sky_to_cartesian(ra, dec, z, c::CosmologyC; ctx::Context = default_context()) =
    _sky(:baorec_sky_to_cartesian, ra, dec, z, c, ctx)
cartesian_to_sky(x, y, z, c::CosmologyC; ctx::Context = default_context()) =
    _sky(:baorec_cartesian_to_sky, x, y, z, c, ctx)

# ------------------------------------------------------------ catalog I/O (host)

function _io_check(st)
    st == 0 && return nothing
    throw(BaorecError(st, unsafe_string(ccall((:baorec_io_last_error, libbaorec), Cstring, ()))))
end

# CSV.File(path; delim = " ", types = Float32) of the drivers
# (test/benchmarks_cpu.jl:22-23): the selected columns (1-based) as Vector{Float32}s
function read_catalog(path::AbstractString, columns = 1:5)
    nrows = Ref{Int64}(0)
    _io_check(ccall((:baorec_catalog_rows, libbaorec), Cint, (Cstring, Ref{Int64}, Ptr{Int32}), path,
                    nrows, C_NULL))
    out = [Vector{Float32}(undef, nrows[]) for _ in columns]
    ptrs = [pointer(o) for o in out]
    got = Ref{Int64}(0)
    GC.@preserve out _io_check(ccall((:baorec_catalog_read, libbaorec), Cint,
                                     (Cstring, Int32, Ptr{Int32}, Ptr{Ptr{Float32}}, Int64, Ref{Int64}), path,
                                     length(columns), Int32[c - 1 for c in columns], ptrs, nrows[], got))
    out
end

# npzwrite(path, hcat(cols...)) of the drivers (examples/lightcone.jl:125-127)
function write_npy(path::AbstractString, cols::Vector{Float32}...)
    n = length(cols[1])
    all(length(c) == n for c in cols) || throw(ArgumentError("columns differ in length"))
    ptrs = [pointer(c) for c in cols]
    GC.@preserve cols _io_check(ccall((:baorec_write_npy, libbaorec), Cint,
                                      (Cstring, Int32, Ptr{Ptr{Float32}}, Int64), path, length(cols), ptrs, n))
end

# ------------------------------------------------------------ power spectrum

mutable struct PkParams
    box_size::NTuple{3,Float32}
    box_min::NTuple{3,Float32}
    los::NTuple{3,Float32}
    mas::Int32
    interlace::Int32
    kmin::Float64
    dk::Float64
    nbins::Int32
    PkParams() = new()
end

# P_0, P_2, P_4 of a periodic box catalog with the configuration of the
# reference's validation (test_helpers/powspec_auto.conf); extension
function power_multipoles(x, y, z, w, grid_size::Tuple{Int,Int,Int}; box_size = 1000.0, box_min = 0.0,
                          dk = 0.005, nbins = 300, ctx::Context = default_context())
    pk = PkParams()
    ccall((:baorec_pk_params_default, libbaorec), Cvoid, (Ref{PkParams},), pk)
    pk.box_size = _t3(fill(box_size, 3))
    pk.box_min = _t3(fill(box_min, 3))
    pk.dk = dk
    pk.nbins = nbins
    k, nm, p0, p2, p4 = (zeros(Float64, nbins) for _ in 1:5)
    check(ctx, ccall((:baorec_power_multipoles, libbaorec), Cint,
                     (Ptr{Cvoid}, Ref{PkParams}, Ptr{Int32}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                      Ptr{Float32}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.ptr, pk, _grid(grid_size), _mem(x, y, z, w), _ptr(x), _ptr(y), _ptr(z), _ptr(w),
                     length(x), k, nm, p0, p2, p4))
    (k = k, nmodes = nm, p0 = p0, p2 = p2, p4 = p4)
end

end # module
