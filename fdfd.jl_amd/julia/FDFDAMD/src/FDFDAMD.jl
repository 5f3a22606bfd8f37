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
    FDFDAMD.install!()   # make FDFD.solve / FDFD.eigenfrequency use the GPU
"""
module FDFDAMD

using FDFD
using Libdl

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

"""
    solve(d::Device, pol::Polarization=TM; precision=:fp64, tol=1e-10)

Same contract as FDFD.solve (src/solver/driven.jl:4-59): one factorization
per frequency in `d.ω`, returns a `FieldTM`/`FieldTE` or a vector of them.
Mode sources (`d.modes`) are not supported by this backend.
"""
function solve(d::FDFD.Device, pol::FDFD.Polarization=FDFD.TM; precision::Symbol=:fp64,
               tol::Real=1e-10, handle::Handle=default_handle())
    isempty(d.modes) || throw(ArgumentError("mode sources are not supported by FDFDAMD"))
    h = handle
    (Nx, Ny) = set_problem!(h, d)
    src = Array{ComplexF64}(d.src)
    fields = pol == FDFD.TM ? FDFD.FieldTM[] : FDFD.FieldTE[]
    for (i, ω) in enumerate(d.ω)
        @info "Frequency: $i/$(length(d.ω))"
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
    install!()

Route `FDFD.solve(::Device, ::Polarization)` and `FDFD.eigenfrequency` to
this backend (method replacement in the FDFD module).
"""
function install!()
    @eval FDFD solve(d::Device, pol::Polarization=TM) = $(solve)(d, pol)
    @eval FDFD eigenfrequency(d::AbstractDevice, pol::Polarization, n::Int; which::Symbol=:LM) =
        $(eigenfrequency)(d, pol, n; which=which)
    return nothing
end

end # module
