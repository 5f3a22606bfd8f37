"""
    FDFDAMD

GPU (MI355X) backend for the driven-solve and eigenfrequency path of FDFD.jl,
calling `libfdfd` through `ccall`.  The problem description stays FDFD.jl's
own: `Grid`, `Device`, `setup_ϵᵣ!`, `setup_src!`; results are FDFD.jl's
`FieldTM` / `FieldTE` built with the `(grid, ω, data::Array{<:Complex,3})`
constructors of src/data.jl:56-58,71-73 (the library writes the Nx×Ny×3
column-major array those constructors take, so no copy is made).

    using FDFD, FDFDAMD
    f = FDFDAMD.solve(dev, TM)               # replaces FDFD.solve (driven.jl:4)
    ω, fs = FDFDAMD.eigenfrequency(dev, TM, 8)  # replaces eigen.jl:69
    β, v = FDFDAMD.eigenmode(dev1d, TM, 3.5, 1)  # replaces eigen.jl:6 (mode sources)
    FDFDAMD.install!()   # make FDFD.solve / FDFD.eigenfrequency use the GPU
"""
module FDFDAMD

using FDFD
using Libdl
using ChainRulesCore

const libfdfd = get(ENV, "FDFD_LIB",
                    normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libfdfd.so")))

const FDFD_FP64 = Cint(0)
const FDFD_FP32_IR = Cint(1)

# mirrors fdfd_stats in include/fdfd.h
struct Stats
    iterations::Cint
    restarts::Cint
    residual::Cdouble
    t_factor_ms::Cdouble
    t_solve_ms::Cdouble
    pivot_growth::Cdouble
    singular_line::Cint
    nconv::Cint
    pivot_fallback::Cint
end
Stats() = Stats(0, 0, 0, 0, 0, 0, -1, 0, 0)

mutable struct Handle
    ptr::Ptr{Cvoid}
end

function Handle(device::Integer=0)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:fdfd_create, libfdfd), Cint, (Cint, Ref{Ptr{Cvoid}}), device, ref)
    rc == 0 || error("fdfd_create failed with status $rc (no GPU visible?)")
    h = Handle(ref[])
    finalizer(h) do x
        x.ptr == C_NULL || ccall((:fdfd_destroy, libfdfd), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return h
end

const _handle = Ref{Union{Nothing,Handle}}(nothing)
default_handle() = (_handle[] === nothing && (_handle[] = Handle(0)); _handle[])

function check(h::Handle, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:fdfd_last_error, libfdfd), Cstring, (Ptr{Cvoid},), h.ptr))
    error("libfdfd status $rc: $msg")
end

precision_code(p::Symbol) = p === :fp64 ? FDFD_FP64 :
                            p in (:mixed, :fp32) ? FDFD_FP32_IR :
                            throw(ArgumentError("precision must be :fp64 or :mixed"))

function set_problem!(h::Handle, d::FDFD.Device)
    g = d.grid
    (Nx, Ny) = size(g)
    check(h, ccall((:fdfd_set_grid, libfdfd), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Cint, Cdouble),
                   h.ptr, Nx, Ny, FDFD.dx(g), FDFD.dy(g), g.Npml[1], g.Npml[2], g.L₀))
    eps = Array{ComplexF64}(d.ϵᵣ)  # Device.ϵᵣ is Array{Complex}: make it concrete
    GC.@preserve eps check(h, ccall((:fdfd_set_eps, libfdfd), Cint,
                                    (Ptr{Cvoid}, Ptr{ComplexF64}), h.ptr, eps))
    return (Nx, Ny)
end

# mirrors fdfd_mode_problem in include/fdfd.h
struct ModeProblem
    n::Cint
    pol::Cint
    omega::Cdouble
    dh::Cdouble
    neff::ComplexF64
    eps::Ptr{ComplexF64}
end

"""
    eigenmode_slices(slices, dh, pol, ω, neff, nev; L₀, handle)

1-D eigenmodes of several ε slices in one `fdfd_eigenmode` call (one GPU
workgroup per slice).  Returns a vector of `(β, vectors)` per slice.
"""
function eigenmode_slices(slices::Vector{Vector{ComplexF64}}, dh::Vector{Float64},
                          pol::Vector{Int}, ω::Vector{Float64}, neff::Vector{ComplexF64},
                          nev::Int; L₀::Real=1e-6, ncv::Int=0, tol::Real=0.0, maxiter::Int=0,
                          handle::Handle=default_handle())
    h = handle
    k = length(slices)
    probs = [ModeProblem(length(slices[p]), pol[p], ω[p], dh[p], neff[p], pointer(slices[p]))
             for p in 1:k]
    ntot = sum(length, slices)
    β = Vector{ComplexF64}(undef, k * nev)
    vecs = Vector{ComplexF64}(undef, ntot * nev)
    nconv = Vector{Cint}(undef, k)
    st = Ref(Stats())
    GC.@preserve slices probs β vecs nconv check(h, ccall((:fdfd_eigenmode, libfdfd), Cint,
        (Ptr{Cvoid}, Cint, Ptr{ModeProblem}, Cint, Cint, Cdouble, Cdouble, Cint,
         Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Cint}, Ref{Stats}),
        h.ptr, k, probs, nev, ncv, L₀, tol, maxiter, β, vecs, nconv, st))
    out = Tuple{Vector{ComplexF64},Matrix{ComplexF64}}[]
    o = 0
    for p in 1:k
        n = length(slices[p])
        push!(out, (β[(p - 1) * nev + 1:p * nev], reshape(vecs[o + 1:o + n * nev], n, nev)))
        o += n * nev
    end
    return out
end

"""
    eigenmode(d::AbstractDevice{1}, pol, neff, neigenvalues)

Same contract as FDFD.eigenmode (src/solver/eigen.jl:6-29): the
`neigenvalues` propagation constants β whose β² are nearest (ω/c₀·neff)², and
their eigenvectors (columns; unit 2-norm, largest entry real and positive).
"""
function eigenmode(d::FDFD.AbstractDevice{1}, pol::FDFD.Polarization, neff::Number,
                   neigenvalues::Int; handle::Handle=default_handle())
    eps = Vector{ComplexF64}(d.ϵᵣ[:])
    (β, v) = eigenmode_slices([eps], [Float64(FDFD.dx(d.grid))], [Int(pol)], [Float64(d.ω[1])],
                              [ComplexF64(neff)], neigenvalues; L₀=d.grid.L₀, handle=handle)[1]
    return (β, v)
end

"""
    eigenmode(d::AbstractDevice{2}, neff, neigenvalues; precision=:fp64)

Same contract as FDFD.eigenmode for a 2-D cross-section (src/solver/eigen.jl:31-66):
the propagation constants β nearest ω/c₀·neff and their `FieldAll` fields.
"""
function eigenmode(d::FDFD.AbstractDevice{2}, neff::Number, neigenvalues::Int;
                   precision::Symbol=:fp64, tol::Real=1e-12, ncv::Int=0, maxiter::Int=300,
                   handle::Handle=default_handle())
    h = handle
    (Nx, Ny) = set_problem!(h, d)
    β = Vector{ComplexF64}(undef, neigenvalues)
    data = Array{ComplexF64}(undef, Nx, Ny, 6, neigenvalues)
    st = Ref(Stats())
    GC.@preserve β data check(h, ccall((:fdfd_vector_eigenmode, libfdfd), Cint,
        (Ptr{Cvoid}, Cdouble, ComplexF64, Cint, Cint, Cint, Cdouble, Cint, Ptr{ComplexF64},
         Ptr{ComplexF64}, Ref{Stats}),
        h.ptr, d.ω[1], ComplexF64(neff), neigenvalues, ncv, precision_code(precision), tol,
        maxiter, β, data, st))
    fields = [FDFD.FieldAll(d.grid, d.ω[1], data[:, :, :, i]) for i in 1:neigenvalues]
    return (β, fields)
end

# src/device.jl:125-142: indices of the odd-length slice through pt and its cell size
function slice_indices(g::FDFD.Grid, pt, slicenormal, slicewidth::Number)
    (indx, indy) = FDFD.coord2ind(g, pt)
    srcpoints = slicenormal == FDFD.x̂ ? Int(round(slicewidth / FDFD.dy(g))) :
                                        Int(round(slicewidth / FDFD.dx(g)))
    srcpoints % 2 == 0 && (srcpoints += 1)
    M = (srcpoints - 1) ÷ 2
    if slicenormal == FDFD.x̂
        return (indx:indx, indy .+ (-M:M), FDFD.dy(g))
    else
        return (indx .+ (-M:M), indy:indy, FDFD.dx(g))
    end
end

"""
    get_modes(d, pol, ω, neff, nmodes, pt, slicenormal, slicewidth)

Same contract as FDFD.get_modes (src/device.jl:124-149): (β, vectors, indx, indy).
"""
function get_modes(d::FDFD.AbstractDevice, pol::FDFD.Polarization, ω::Real, neff::Number,
                   nmodes::Int, pt, slicenormal, slicewidth::Number;
                   handle::Handle=default_handle())
    (ix, iy, dh) = slice_indices(d.grid, pt, slicenormal, slicewidth)
    eps = Vector{ComplexF64}(d.ϵᵣ[ix, iy][:])
    (β, v) = eigenmode_slices([eps], [Float64(dh)], [Int(pol)], [Float64(ω)],
                              [ComplexF64(neff)], nmodes; L₀=d.grid.L₀, handle=handle)[1]
    return (β, v, length(ix) == 1 ? ix[1] : ix, length(iy) == 1 ? iy[1] : iy)
end

"""
    setup_mode!(d, pol, ω, neff, pt, srcnormal, srcwidth)

Same contract as FDFD.setup_mode! (src/device.jl:118-121): adds
normalize(abs.(v)) of the mode nearest the guess to `d.src` on the slice.
"""
function setup_mode!(d::FDFD.AbstractDevice, pol::FDFD.Polarization, ω::Real, neff::Number,
                     pt, srcnormal, srcwidth::Number; handle::Handle=default_handle())
    (_, v, indx, indy) = get_modes(d, pol, ω, neff, 1, pt, srcnormal, srcwidth; handle=handle)
    a = abs.(v[:, 1])
    d.src[indx, indy] += a ./ sqrt(sum(abs2, a))
    return nothing
end

# driven.jl:15-19: the source of one frequency from the device's modes (TM,
# as the reference passes), all slices in one batched call
function mode_source!(d::FDFD.Device, ω::Real, handle::Handle)
    d.src = zeros(Complex, size(d.grid))
    isempty(d.modes) && return nothing
    sl = [slice_indices(d.grid, m.pt, m.dir, m.width) for m in d.modes]
    res = eigenmode_slices([Vector{ComplexF64}(d.ϵᵣ[ix, iy][:]) for (ix, iy, _) in sl],
                           [Float64(dh) for (_, _, dh) in sl], fill(Int(FDFD.TM), length(sl)),
                           fill(Float64(ω), length(sl)), [ComplexF64(m.neff) for m in d.modes],
                           1; L₀=d.grid.L₀, handle=handle)
    for ((ix, iy, _), (_, v)) in zip(sl, res)
        a = abs.(v[:, 1])
        d.src[ix, iy] .+= reshape(a ./ sqrt(sum(abs2, a)), length(ix), length(iy))
    end
    return nothing
end

"""
    solve(d::Device, pol::Polarization=TM; precision=:fp64, tol=1e-10)

Same contract as FDFD.solve (src/solver/driven.jl:4-59): one factorization
per frequency in `d.ω`, returns a `FieldTM`/`FieldTE` or a vector of them.
With mode sources (`d.modes`) the source is rebuilt from the TM modes at every
frequency (driven.jl:15-19).
"""
function solve(d::FDFD.Device, pol::FDFD.Polarization=FDFD.TM; precision::Symbol=:fp64,
               tol::Real=1e-10, handle::Handle=default_handle())
    h = handle
    (Nx, Ny) = set_problem!(h, d)
    fields = pol == FDFD.TM ? FDFD.FieldTM[] : FDFD.FieldTE[]
    for (i, ω) in enumerate(d.ω)
        @info "Frequency: $i/$(length(d.ω))"
        isempty(d.modes) || mode_source!(d, ω, h)
        src = Array{ComplexF64}(d.src)
        fid = Ref{Cint}(0)
        st = Ref(Stats())
        check(h, ccall((:fdfd_factor, libfdfd), Cint,
                       (Ptr{Cvoid}, Cint, Cdouble, Cint, Ref{Cint}, Ref{Stats}),
                       h.ptr, Int(pol), ω, precision_code(precision), fid, st))
        out = Array{ComplexF64}(undef, Nx, Ny, 3)
        rc = GC.@preserve src out ccall((:fdfd_solve, libfdfd), Cint,
                                         (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64},
                                          Ptr{ComplexF64}, Cdouble, Ref{Stats}),
                                         h.ptr, fid[], 1, src, out, tol, st)
        ccall((:fdfd_release_factor, libfdfd), Cint, (Ptr{Cvoid}, Cint), h.ptr, fid[])
        check(h, rc)
        @info "Unknowns: $(Nx*Ny)  factor $(st[].t_factor_ms) ms  solve $(st[].t_solve_ms) ms"
        push!(fields, pol == FDFD.TM ? FDFD.FieldTM(d.grid, ω, out) : FDFD.FieldTE(d.grid, ω, out))
    end
    return length(fields) == 1 ? fields[1] : fields
end

"""
    solve_sources(d::Device, srcs::AbstractArray{<:Number,3}, pol=TM;
                  precision=:mixed, tol=1e-10)

Several sources per frequency on one factorization: `fdfd_solve` with
`nrhs = size(srcs, 3)` runs a lockstep multi-source FGMRES (16 sources per
pass over the factor blocks).  Returns, per frequency of `d.ω`, a vector of
`FieldTM`/`FieldTE`, one per source.  The reference solves one source per
`solve` call (src/solver/driven.jl:4-59).
"""
function solve_sources(d::FDFD.Device, srcs::AbstractArray{<:Number,3},
                       pol::FDFD.Polarization=FDFD.TM; precision::Symbol=:mixed, tol::Real=1e-10,
                       handle::Handle=default_handle())
    h = handle
    (Nx, Ny) = set_problem!(h, d)
    (size(srcs, 1), size(srcs, 2)) == (Nx, Ny) ||
        throw(DimensionMismatch("srcs must be Nx × Ny × nsrc"))
    ns = size(srcs, 3)
    src = Array{ComplexF64,3}(srcs)
    F = pol == FDFD.TM ? FDFD.FieldTM : FDFD.FieldTE
    fields = Vector{Vector{F}}()
    for ω in d.ω
        fid = Ref{Cint}(0)
        st = Ref(Stats())
        check(h, ccall((:fdfd_factor, libfdfd), Cint,
                       (Ptr{Cvoid}, Cint, Cdouble, Cint, Ref{Cint}, Ref{Stats}),
                       h.ptr, Int(pol), ω, precision_code(precision), fid, st))
        out = Array{ComplexF64}(undef, Nx, Ny, 3, ns)
        rc = GC.@preserve src out ccall((:fdfd_solve, libfdfd), Cint,
                                         (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64},
                                          Ptr{ComplexF64}, Cdouble, Ref{Stats}),
                                         h.ptr, fid[], ns, src, out, tol, st)
        ccall((:fdfd_release_factor, libfdfd), Cint, (Ptr{Cvoid}, Cint), h.ptr, fid[])
        check(h, rc)
        push!(fields, [F(d.grid, ω, out[:, :, :, r]) for r in 1:ns])
    end
    return fields
end

"""
    solve_batch(jobs; precision=:mixed, tol=1e-10) -> (results, stats)

Many devices × frequencies × sources in one `fdfd_batch_run`.  `jobs` is a
vector of `(d::Device, pol, srcs::AbstractArray{<:Number,3})`: each device's
grid, `ϵᵣ` and frequency list `d.ω`, with `size(srcs, 3)` sources solved at
every frequency.  Grids may differ.  `results[j][iω][isrc]` is a
`FieldTM`/`FieldTE`.  The handle's own factors are released and its grid is
reset (`set_problem!` before the next `solve`).  Replaces a loop of
`solve` calls over devices, frequencies and sources (driven.jl:11).
"""
function solve_batch(jobs; precision::Symbol=:mixed, tol::Real=1e-10,
                     handle::Handle=default_handle())
    h = handle
    bref = Ref{Ptr{Cvoid}}(C_NULL)
    check(h, ccall((:fdfd_batch_create, libfdfd), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                   h.ptr, bref))
    b = bref[]
    try
        ids = Cint[]
        for (d, pol, srcs) in jobs
            g = d.grid
            (Nx, Ny) = size(g)
            eps = Array{ComplexF64}(d.ϵᵣ)
            om = Float64.(collect(d.ω))
            src = Array{ComplexF64,3}(srcs)
            jid = Ref{Cint}(0)
            GC.@preserve eps om src check(h, ccall((:fdfd_batch_add, libfdfd), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Cint, Cdouble, Cint,
                 Ptr{ComplexF64}, Cint, Ptr{Cdouble}, Cint, Ptr{ComplexF64}, Ref{Cint}),
                b, Nx, Ny, FDFD.dx(g), FDFD.dy(g), g.Npml[1], g.Npml[2], g.L₀, Int(pol),
                eps, length(om), om, size(src, 3), src, jid))
            push!(ids, jid[])
        end
        st = Ref(Stats())
        check(h, ccall((:fdfd_batch_run, libfdfd), Cint, (Ptr{Cvoid}, Cint, Cdouble, Ref{Stats}),
                       b, precision_code(precision), tol, st))
        results = Any[]
        for ((d, pol, srcs), jid) in zip(jobs, ids)
            (Nx, Ny) = size(d.grid)
            F = pol == FDFD.TM ? FDFD.FieldTM : FDFD.FieldTE
            perω = Vector{Vector{F}}()
            for (io, ω) in enumerate(d.ω)
                fs = F[]
                for r in 1:size(srcs, 3)
                    out = Array{ComplexF64}(undef, Nx, Ny, 3)
                    GC.@preserve out check(h, ccall((:fdfd_batch_result, libfdfd), Cint,
                        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{ComplexF64}), b, jid, io - 1, r - 1, out))
                    push!(fs, F(d.grid, ω, out))
                end
                push!(perω, fs)
            end
            push!(results, perω)
        end
        return results, st[]
    finally
        ccall((:fdfd_batch_destroy, libfdfd), Cint, (Ptr{Cvoid},), b)
    end
end

"""
    autotune!(d::Device, pol=TM; precision=:mixed) -> NamedTuple

Times every variant of the factorization kernels on `d`'s grid and `ϵᵣ` at
`d.ω[1]` (`fdfd_autotune`) and keeps the fastest on the handle; with
`ENV["FDFD_TUNE_CACHE"]` naming a file the choice is recorded there and read
by later handles.  No counterpart in the reference.
"""
function autotune!(d::FDFD.Device, pol::FDFD.Polarization=FDFD.TM; precision::Symbol=:mixed,
                   handle::Handle=default_handle())
    h = handle
    set_problem!(h, d)
    st = Ref(Stats())
    check(h, ccall((:fdfd_autotune, libfdfd), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint, Ref{Stats}),
                   h.ptr, Int(pol), Float64(d.ω[1]), precision_code(precision), st))
    knobs = Vector{Cint}(undef, 5)
    check(h, ccall((:fdfd_get_tuning, libfdfd), Cint, (Ptr{Cvoid}, Ptr{Cint}, Cint),
                   h.ptr, knobs, 5))
    return (step_ms=st[].t_factor_ms, gemm_fp16x3=knobs[1] == 1, trailing_fp16x3=knobs[2] == 1,
            fused_panels=knobs[3] == 1, gemm_waves=Int(knobs[4]), gemm3m=knobs[5] == 1)
end

# ---- reverse-mode derivative of solve with respect to ϵᵣ (ChainRulesCore) --
# The reference has no derivative.  The forward pass is `solve`, keeping every
# frequency's factor alive until the pullback has run; the pullback is one
# `fdfd_solve_pullback` per frequency: a transposed FGMRES solve on the same
# factor (A^T λ = conj(ē), ē the cotangent reaching Ez through the H recovery
# of driven.jl:40-41) and the gradient kernel dL/dϵᵣ[k] = -Re(λ_k ω² ϵ₀' Ez_k)
# (∂A/∂ϵᵣ[k] = ω² ϵ₀' e_k e_kᵀ, driven.jl:21,35).  TM only; the source is
# held fixed (mode sources whose slice crosses the design are not followed).

mutable struct FactorSet
    handle::Handle
    ids::Vector{Cint}
end

function release!(fs::FactorSet)
    for id in fs.ids
        ccall((:fdfd_release_factor, libfdfd), Cint, (Ptr{Cvoid}, Cint), fs.handle.ptr, id)
    end
    empty!(fs.ids)
    return nothing
end

function solve_keep(d::FDFD.Device, precision::Symbol, tol::Real, h::Handle)
    isempty(d.modes) || throw(ArgumentError("the solve rrule needs an explicit source (d.modes is not empty)"))
    (Nx, Ny) = set_problem!(h, d)
    src = Array{ComplexF64}(d.src)
    fs = FactorSet(h, Cint[])
    finalizer(release!, fs)
    fields = FDFD.FieldTM[]
    for ω in d.ω
        fid = Ref{Cint}(0)
        st = Ref(Stats())
        check(h, ccall((:fdfd_factor, libfdfd), Cint,
                       (Ptr{Cvoid}, Cint, Cdouble, Cint, Ref{Cint}, Ref{Stats}),
                       h.ptr, Int(FDFD.TM), ω, precision_code(precision), fid, st))
        push!(fs.ids, fid[])
        out = Array{ComplexF64}(undef, Nx, Ny, 3)
        GC.@preserve src out check(h, ccall((:fdfd_solve, libfdfd), Cint,
            (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}, Cdouble, Ref{Stats}),
            h.ptr, fid[], 1, src, out, tol, st))
        push!(fields, FDFD.FieldTM(d.grid, ω, out))
    end
    return fields, fs
end

# cotangent of a FieldTM: a Tangent with a `data` field, the data array
# itself, or nothing / ZeroTangent
_field_bar(Δ, Nx, Ny) = Δ isa AbstractZero || Δ === nothing ? nothing :
                        Δ isa AbstractArray ? Array{ComplexF64}(Δ) :
                        _field_bar(unthunk(Δ).data, Nx, Ny)

function pullback_eps(fields, fs::FactorSet, Δs, tol::Real)
    h = fs.handle
    (Nx, Ny) = size(fields[1].grid)
    total = zeros(Float64, Nx, Ny)
    g = Array{Float64}(undef, Nx, Ny)
    for (f, id, Δ) in zip(fields, fs.ids, Δs)
        bar = _field_bar(Δ, Nx, Ny)
        bar === nothing && continue
        if size(bar) == (Nx, Ny)      # cotangent of Ez alone
            full = zeros(ComplexF64, Nx, Ny, 3)
            full[:, :, 1] .= bar
            bar = full
        end
        ez = Array{ComplexF64}(f.data[:, :, 1])
        st = Ref(Stats())
        GC.@preserve ez bar g check(h, ccall((:fdfd_solve_pullback, libfdfd), Cint,
            (Ptr{Cvoid}, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}, Cdouble, Ptr{Float64},
             Ref{Stats}),
            h.ptr, id, ez, bar, tol, g, st))
        total .+= g
    end
    return total
end

"""
    solve_eps(ϵᵣ, d::Device; precision=:fp64, tol=1e-12)

The TM driven solve of `d` with permittivity `ϵᵣ` (d.ϵᵣ is not modified): the
differentiable form of `solve(d, TM)` for `Zygote` / `ChainRulesCore`, whose
`rrule` returns dL/dϵᵣ as a real array.
"""
function solve_eps(ϵᵣ::AbstractMatrix, d::FDFD.Device; precision::Symbol=:fp64,
                   tol::Real=1e-12, handle::Handle=default_handle())
    d2 = deepcopy(d)
    d2.ϵᵣ = Array{Complex}(ϵᵣ)
    return solve(d2, FDFD.TM; precision=precision, tol=tol, handle=handle)
end

function ChainRulesCore.rrule(::typeof(solve_eps), ϵᵣ::AbstractMatrix, d::FDFD.Device;
                              precision::Symbol=:fp64, tol::Real=1e-12,
                              handle::Handle=default_handle())
    d2 = deepcopy(d)
    d2.ϵᵣ = Array{Complex}(ϵᵣ)
    fields, fs = solve_keep(d2, precision, tol, handle)
    y = length(fields) == 1 ? fields[1] : fields
    function solve_eps_pullback(Δ)
        Δs = length(fields) == 1 ? [Δ] : collect(unthunk(Δ))
        g = pullback_eps(fields, fs, Δs, tol)
        release!(fs)
        return (NoTangent(), g, NoTangent())
    end
    return y, solve_eps_pullback
end

# rrule of solve(d::Device, TM) itself: the tangent of the Device carries ϵᵣ
function ChainRulesCore.rrule(::typeof(solve), d::FDFD.Device,
                              pol::FDFD.Polarization=FDFD.TM; precision::Symbol=:fp64,
                              tol::Real=1e-12, handle::Handle=default_handle())
    pol == FDFD.TM || throw(ArgumentError("the solve rrule is implemented for TM"))
    fields, fs = solve_keep(d, precision, tol, handle)
    y = length(fields) == 1 ? fields[1] : fields
    function solve_pullback(Δ)
        Δs = length(fields) == 1 ? [Δ] : collect(unthunk(Δ))
        g = pullback_eps(fields, fs, Δs, tol)
        release!(fs)
        return (NoTangent(), Tangent{typeof(d)}(; ϵᵣ=g), NoTangent())
    end
    return y, solve_pullback
end

"""
    eigenfrequency(d, pol, nev; which=:LM, precision=:fp64, tol=1e-12, ncv=0)

Same contract as FDFD.eigenfrequency (src/solver/eigen.jl:69-115): complex
eigenfrequencies nearest ω₀ = d.ω[1] and their fields.
"""
function eigenfrequency(d::FDFD.AbstractDevice, pol::FDFD.Polarization, nev::Int;
                        which::Symbol=:LM, precision::Symbol=:fp64, tol::Real=1e-12,
                        ncv::Int=0, maxiter::Int=300, handle::Handle=default_handle())
    h = handle
    (Nx, Ny) = set_problem!(h, d)
    ω = Vector{ComplexF64}(undef, nev)
    data = Array{ComplexF64}(undef, Nx, Ny, 3, nev)
    st = Ref(Stats())
    wh = which === :LR ? Cint(1) : Cint(0)
    GC.@preserve ω data check(h, ccall((:fdfd_eigenfrequency, libfdfd), Cint,
        (Ptr{Cvoid}, Cint, Cdouble, Cint, Cint, Cint, Cint, Cdouble, Cint,
         Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{Stats}),
        h.ptr, Int(pol), d.ω[1], nev, ncv, wh, precision_code(precision), tol, maxiter,
        ω, data, st))
    fields = [pol == FDFD.TM ? FDFD.FieldTM(d.grid, ω[i], data[:, :, :, i]) :
                               FDFD.FieldTE(d.grid, ω[i], data[:, :, :, i]) for i in 1:nev]
    return (ω, fields)
end

"""
    solve(d::ModulatedDevice; precision=:fp64, tol=1e-10)

Same contract as FDFD.solve(d::ModulatedDevice) (src/solver/modulation.jl:35-119):
a `length(d.ω) × (2K+1)` array of `FieldTM`, sideband j at ωₙ = ω + (j-K-1)Ω.
Each carrier is one `fdfd_modulated_solve` call: the 2K+1 sideband operators
are factored on the GPU and FP64 FGMRES on the coupled operator, preconditioned
by them, reaches `tol`.  Mode sources are rebuilt per carrier (TM).
"""
function solve(d::FDFD.ModulatedDevice; precision::Symbol=:fp64, tol::Real=1e-10,
               handle::Handle=default_handle())
    h = handle
    g = d.grid
    (Nx, Ny) = size(g)
    check(h, ccall((:fdfd_set_grid, libfdfd), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Cint, Cdouble),
                   h.ptr, Nx, Ny, FDFD.dx(g), FDFD.dy(g), g.Npml[1], g.Npml[2], g.L₀))
    eps = Array{ComplexF64}(d.ϵᵣ)
    GC.@preserve eps check(h, ccall((:fdfd_set_eps, libfdfd), Cint,
                                    (Ptr{Cvoid}, Ptr{ComplexF64}), h.ptr, eps))
    K = d.nsidebands
    nf = 2K + 1
    deps = Array{ComplexF64}(d.Δϵᵣ)
    fields = Array{FDFD.FieldTM}(undef, length(d.ω), nf)
    for (i, ω) in enumerate(d.ω)
        @info "Frequency: $i/$(length(d.ω))"
        if !isempty(d.modes)
            d.src = zeros(Complex, size(g))
            for m in d.modes
                setup_mode!(d, FDFD.TM, ω, m.neff, m.pt, m.dir, m.width; handle=h)
            end
        end
        src = Array{ComplexF64}(d.src)
        out = Array{ComplexF64}(undef, Nx, Ny, 3, nf)
        st = Ref(Stats())
        GC.@preserve deps src out check(h, ccall((:fdfd_modulated_solve, libfdfd), Cint,
            (Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}, Cint,
             Cdouble, Ptr{ComplexF64}, Ref{Stats}),
            h.ptr, ω, d.Ω, K, d.sharedpml ? 1 : 0, deps, src, precision_code(precision), tol,
            out, st))
        for j in 1:nf
            fields[i, j] = FDFD.FieldTM(g, ω + d.Ω * (j - K - 1), out[:, :, :, j])
        end
    end
    return fields
end

# mirrors fdfd_flux_plane in include/fdfd.h (0-based indices)
struct FluxPlane
    normal::Cint
    index::Cint
    lo::Cint
    hi::Cint
end

_flat(f::FDFD.Field) = Array{ComplexF64}(f.data)   # Nx x Ny x 3, column-major

"""
    poynting(field)

Same contract as FDFD.poynting (src/flux.jl:15-30), with the TE branch
completed (the reference's reads an undefined field): returns the
reference's `Flux(grid, ω, Sx, Sy)`.
"""
function poynting(f::FDFD.Field; handle::Handle=default_handle())
    (Nx, Ny) = size(f.grid)
    data = _flat(f)
    S = Array{Float64}(undef, Nx, Ny, 2)
    pol = isa(f, FDFD.FieldTM) ? 1 : 2
    GC.@preserve data S check(handle, ccall((:fdfd_poynting, libfdfd), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{ComplexF64}, Ptr{Float64}),
        handle.ptr, pol, Nx, Ny, data, S))
    # the 3-D constructor: the (Sx, Sy) one uses the pre-1.0 cat(3, ...) (data.jl:108)
    return FDFD.Flux(f.grid, f.ω, S)
end

"""
    flux_surface_integral(field, center, width, normal)

Same contract as FDFD.flux_surface_integral (src/flux.jl:37-66); the TE and
ŷ-normal branches (broken or unimplemented in the reference) are completed.
"""
function flux_surface_integral(f::FDFD.Field, center, width::Real, normal;
                               handle::Handle=default_handle())
    g = f.grid
    (Nx, Ny) = size(g)
    xc, yc = FDFD.xc(g), FDFD.yc(g)
    if normal == FDFD.x̂
        i = findfirst(x -> abs(x - center.x) <= FDFD.dx(g) / 2, xc)
        line = findall(y -> center.y - width <= y <= center.y + width, yc)
    else
        i = findfirst(y -> abs(y - center.y) <= FDFD.dy(g) / 2, yc)
        line = findall(x -> center.x - width <= x <= center.x + width, xc)
    end
    (i === nothing || isempty(line)) && error("flux plane outside the grid")
    planes = [FluxPlane(Int(normal), i - 1, first(line) - 1, last(line) - 1)]
    data = _flat(f)
    out = Vector{Float64}(undef, 1)
    pol = isa(f, FDFD.FieldTM) ? 1 : 2
    GC.@preserve data planes out check(handle, ccall((:fdfd_flux, libfdfd), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Cdouble, Ptr{ComplexF64}, Cint,
         Ptr{FluxPlane}, Ptr{Float64}),
        handle.ptr, pol, Nx, Ny, FDFD.dx(g), FDFD.dy(g), data, 1, planes, out))
    return out[1]
end

"""
    install!()

Route `FDFD.solve(::Device, ::Polarization)`, `FDFD.solve(::ModulatedDevice)`,
`FDFD.eigenfrequency` and the 1-D `FDFD.eigenmode` (so `setup_mode!` /
`get_modes` as well) to this backend (method replacement in the FDFD module).
"""
function install!()
    @eval FDFD solve(d::Device, pol::Polarization=TM) = $(solve)(d, pol)
    @eval FDFD solve(d::ModulatedDevice) = $(solve)(d)
    @eval FDFD poynting(f::Field) = $(poynting)(f)
    @eval FDFD flux_surface_integral(f::Field, c::Point, w::Real, n::Direction) =
        $(flux_surface_integral)(f, c, w, n)
    @eval FDFD eigenmode(d::AbstractDevice{1}, pol::Polarization, neff::Number, n::Int) =
        $(eigenmode)(d, pol, neff, n)
    @eval FDFD eigenmode(d::AbstractDevice{2}, neff::Number, n::Int) = $(eigenmode)(d, neff, n)
    @eval FDFD eigenfrequency(d::AbstractDevice, pol::Polarization, n::Int; which::Symbol=:LM) =
        $(eigenfrequency)(d, pol, n; which=which)
    return nothing
end

end # module
