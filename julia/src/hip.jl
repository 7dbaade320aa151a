# hip.jl — the libmcrt_hip.so binding of MonteCarloRadiativeTransfer.jl.
#
# Include from src/MonteCarloRadiativeTransfer.jl after geometries/ and
# scatterers/ (it uses Config, PlaneGeometry, SphereGeometry, Layer,
# RayleighScatterer and MieScatterer as they are defined there).  The GPU
# branch of `simulate` (MonteCarloRadiativeTransfer.jl:36-56) then becomes
#
#     use_gpu ? simulate_hip(cfg, g, I₀) : <the CPU path>
#
# and `run_rt`'s table preparation (:97-104 sphere, :194-199 plane) becomes
# `pin_tables(s.scatterer) do ... end` around its Stokes-column loop.
#
# The structs mirror include/mcrt.h field for field (the library's own test
# suite checks the C layout); every pointer a call reads is kept alive by
# GC.@preserve for the duration of the ccall.

const libmcrt = get(ENV, "MCRT_LIB", joinpath(@__DIR__, "..", "deps", "libmcrt_hip.so"))

# ---------------------------------------------------------------- status codes
const MCRT_OK = Int32(0)
const MCRT_ERR_INVALID = Int32(-1)      # the reference's ArgumentError
const MCRT_ERR_UNSUPPORTED = Int32(-2)
const MCRT_ERR_NOMEM = Int32(-3)
const MCRT_ERR_HIP = Int32(-4)
const MCRT_ERR_CANCELLED = Int32(-5)
const MCRT_ERR_NOT_READY = Int32(-6)
const MCRT_ERR_NO_DEVICE = Int32(-7)

const PRECISION = Dict(:fp64_plain => Int32(0), :fp64 => Int32(1), :mixed => Int32(2))

# ---------------------------------------------------------------- C structs
struct McrtScatterer          # mcrt_scatterer
    kind::Int32               # 0 Rayleigh, 1 Mie, 2 Spline
    n_theta::Int32
    random_points::Int32
    reserved::Int32
    omega::Float64
    kl::Float64
    csrn::Ptr{Float64}
    F11::Ptr{Float64}
    F12::Ptr{Float64}
    F33::Ptr{Float64}
    F34::Ptr{Float64}
    S11::Ptr{Float64}         # ComplexF64 vectors seen as interleaved (re, im)
    S22::Ptr{Float64}
    F22::Ptr{Float64}         # Spline only
    F44::Ptr{Float64}
end

struct McrtLayer              # mcrt_layer
    d::Float64
    n::Float64
    tau::Float64
    omega::Float64
    scatterer::Ptr{McrtScatterer}
end

struct McrtGeometry           # mcrt_geometry
    kind::Int32               # 0 plane, 1 sphere
    n_layers::Int32
    theta::Float64
    phi::Float64
    layers::Ptr{McrtLayer}
    R::Float64
    n::Float64
    tau::Float64
    omega::Float64
    scatterer::Ptr{McrtScatterer}
end

struct McrtConfig             # mcrt_config
    number_of_rays::Int64
    minimum_intensity::Float64
    maximum_scattering_times::Int64
    maximum_threads::Int64
    compute_cb::Int32
    reserved::Int32
    tau0::Float64
    n_phi::Int64
    phi::Ptr{Float64}
    cos_phi::Ptr{Float64}
    sin_phi::Ptr{Float64}
    n_theta::Int64
    cos_theta::Ptr{Float64}
    sin_theta::Ptr{Float64}
    inorm::Ptr{Float64}
    n_theta_b::Int64
    theta_b::Ptr{Float64}
    cos_theta_b::Ptr{Float64}
    sin_theta_b::Ptr{Float64}
    n_phi_b::Int64
    phi_b::Ptr{Float64}
    cos_phi_b::Ptr{Float64}
    sin_phi_b::Ptr{Float64}
end

struct McrtStats              # mcrt_stats
    rays::UInt64
    events::UInt64
    rt_evals::UInt64
    cb_evals::UInt64
    cb_evals_single::UInt64
    rejections::UInt64
    nonfinite::UInt64
    kernel_ms::Float64
    total_ms::Float64
    batches::UInt64
    chunks::UInt64
    queue_ms::Float64
    wait_ms::Float64
end
McrtStats() = McrtStats(0, 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0, 0, 0.0, 0.0)

struct McrtRequest            # mcrt_request
    cfg::Ptr{McrtConfig}
    geom::Ptr{McrtGeometry}
    I0::NTuple{4, Float64}
    seed::UInt64
    ray_begin::UInt64
    instance::UInt32
    precision::Int32
    chunk_rays::Int64
    slots::Int32              # launch shape of the solve's chunks (0: the defaults)
    stage_s::Int32
end

# ---------------------------------------------------------------- errors
"""The library's message for the last failed call (on `ctx`, or on the
calling thread when `ctx` is C_NULL)."""
last_error(ctx::Ptr{Cvoid} = C_NULL) =
    unsafe_string(ccall((:mcrt_last_error, libmcrt), Cstring, (Ptr{Cvoid},), ctx))

function check(rc::Integer, ctx::Ptr{Cvoid} = C_NULL)
    rc == MCRT_OK && return nothing
    msg = last_error(ctx)
    rc == MCRT_ERR_INVALID && throw(ArgumentError(msg))   # as the reference's validate
    error("libmcrt_hip ($rc): $msg")
end

# ---------------------------------------------------------------- context
const CTX = Ref{Ptr{Cvoid}}(C_NULL)

"""The process-wide device context (device 0), created on first use."""
function hip_context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:mcrt_context_create, libmcrt), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h))
        CTX[] = h[]
        atexit() do
            CTX[] == C_NULL || ccall((:mcrt_context_destroy, libmcrt), Cvoid, (Ptr{Cvoid},), CTX[])
            CTX[] = C_NULL
        end
    end
    return CTX[]
end

# ---------------------------------------------------------------- marshalling
# Each helper returns the C struct and a vector of the Julia objects whose
# memory it points into; keep that vector alive (GC.@preserve) across the call.

_f64(v) = v isa Vector{Float64} ? v : collect(Float64, v)

"""`mcrt_scatterer(s)` → (McrtScatterer, keep).  RayleighScatterer
(rayleigh.jl:3-6) has no tables; MieScatterer (mie.jl:3-15) passes its
F11..F34 and csrn vectors and its S11/S22 complex vectors as interleaved
(re, im) Float64; SplineScatterer (spline.jl:3-15) passes F22 and F44 too, with
Nθ = length(F11) and random_points = length(csrn) - 1."""
function mcrt_scatterer(s::RayleighScatterer)
    return McrtScatterer(0, 0, 0, 0, s.ω, s.kl, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL,
                         C_NULL, C_NULL, C_NULL), Any[]
end

function mcrt_scatterer(s::MieScatterer)
    csrn, F11, F12, F33, F34 = _f64(s.csrn), _f64(s.F11), _f64(s.F12), _f64(s.F33), _f64(s.F34)
    S11 = collect(ComplexF64, s.S11)
    S22 = collect(ComplexF64, s.S22)
    length(csrn) == s.random_points + 1 ||
        throw(ArgumentError("csrn must hold random_points + 1 values"))
    all(length(v) == s.Nθ for v in (F11, F12, F33, F34, S11, S22)) ||
        throw(ArgumentError("the Mie tables must hold Nθ values"))
    p(v) = Ptr{Float64}(pointer(v))
    c = McrtScatterer(1, s.Nθ, s.random_points, 0, s.ω, s.kl, p(csrn), p(F11), p(F12), p(F33),
                      p(F34), p(S11), p(S22), C_NULL, C_NULL)
    return c, Any[csrn, F11, F12, F33, F34, S11, S22]
end

function mcrt_scatterer(s::SplineScatterer)
    csrn, F11, F12, F22 = _f64(s.csrn), _f64(s.F11), _f64(s.F12), _f64(s.F22)
    F33, F34, F44 = _f64(s.F33), _f64(s.F34), _f64(s.F44)
    S11 = collect(ComplexF64, s.S11)
    S22 = collect(ComplexF64, s.S22)
    N = length(F11)
    all(length(v) == N for v in (F12, F22, F33, F34, F44, S11, S22)) ||
        throw(ArgumentError("the Spline tables must have one length"))
    p(v) = Ptr{Float64}(pointer(v))
    c = McrtScatterer(2, N, length(csrn) - 1, 0, s.ω, s.kl, p(csrn), p(F11), p(F12), p(F33),
                      p(F34), p(S11), p(S22), p(F22), p(F44))
    return c, Any[csrn, F11, F12, F22, F33, F34, F44, S11, S22]
end

mcrt_scatterer(s) = throw(ArgumentError("libmcrt_hip has no form for $(typeof(s))"))

"""`mcrt_geometry(g)` → (McrtGeometry, keep) for PlaneGeometry (plane.jl:28-38,
its Layer vector plane.jl:6-21, semi-infinite layers as d < 0) or
SphereGeometry (sphere.jl:6-21).  Layers that share a scatterer object share
its tables."""
function mcrt_geometry(g::PlaneGeometry)
    keep = Any[]
    seen = IdDict{Any, Ref{McrtScatterer}}()
    for l in g.layers
        haskey(seen, l.scatterer) && continue
        c, k = mcrt_scatterer(l.scatterer)
        r = Ref(c)
        seen[l.scatterer] = r
        push!(keep, r, k)
    end
    layers = [McrtLayer(l.d, l.n, l.τ, l.ω,
                        Base.unsafe_convert(Ptr{McrtScatterer}, seen[l.scatterer]))
              for l in g.layers]
    push!(keep, layers)
    geo = McrtGeometry(0, length(layers), g.θ, g.ϕ, pointer(layers), 0.0, 1.0, 0.0, 0.0, C_NULL)
    return geo, keep
end

function mcrt_geometry(g::SphereGeometry)
    c, k = mcrt_scatterer(g.scatterer)
    r = Ref(c)
    geo = McrtGeometry(1, 0, Float64(π), 0.0, C_NULL, g.R, g.n, g.τ, g.ω,
                       Base.unsafe_convert(Ptr{McrtScatterer}, r))
    return geo, Any[r, k]
end

"""`mcrt_config(cfg)` → (McrtConfig, keep): Config (config.jl:55-124) with its
angle vectors, as simulate reads them."""
function mcrt_config(cfg::Config)
    v = map(_f64, (cfg.ϕ, cfg.cosϕ, cfg.sinϕ, cfg.cosθ, cfg.sinθ, cfg.inorm, cfg.θb, cfg.cosθb,
                   cfg.sinθb, cfg.ϕb, cfg.cosϕb, cfg.sinϕb))
    p(x) = pointer(x)
    c = McrtConfig(cfg.number_of_rays, cfg.minimum_intensity, cfg.maximum_scattering_times,
                   cfg.maximum_threads, Int32(cfg.compute_cb), 0, cfg.τ₀,
                   cfg.Nϕ, p(v[1]), p(v[2]), p(v[3]),
                   cfg.Nθ, p(v[4]), p(v[5]), p(v[6]),
                   cfg.Nθb, p(v[7]), p(v[8]), p(v[9]),
                   cfg.Nϕb, p(v[10]), p(v[11]), p(v[12]))
    return c, Any[v...]
end

# ---------------------------------------------------------------- simulate
"""
    simulate_hip(cfg, g, I₀; seed = 42, precision = :fp64, stats = false)

The GPU branch of `simulate` (MonteCarloRadiativeTransfer.jl:36-56) on the
HIP library: returns `(Adt, Aabs, Aref, Atrans, Arem, Irt, Icb)` divided by N,
with `Irt` of size (4, Nθ, Nϕ) and `Icb` of size (4, Nθb, Nϕb) as the
reference lays them out.  `precision` is :fp64_plain (the structure of
gpu.jl), :fp64 or :mixed.  With `stats = true` the work counters of the solve
are returned as an eighth value.
"""
function simulate_hip(cfg::Config, g, I₀; seed::Integer = 42, precision::Symbol = :fp64,
                      stats::Bool = false)
    c, kc = mcrt_config(cfg)
    geo, kg = mcrt_geometry(g)
    I0 = collect(Float64, I₀)
    length(I0) == 4 || throw(ArgumentError("I₀ must have 4 Stokes components"))
    A = zeros(5)
    Irt = zeros(4, cfg.Nθ, cfg.Nϕ)
    Icb = zeros(4, max(cfg.Nθb, 1), max(cfg.Nϕb, 1))
    st = Ref(McrtStats())
    ctx = hip_context()
    rc = GC.@preserve kc kg I0 A Irt Icb begin
        ccall((:mcrt_simulate, libmcrt), Cint,
              (Ptr{Cvoid}, Ref{McrtConfig}, Ref{McrtGeometry}, Ptr{Float64}, UInt64, Int32,
               Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{McrtStats}),
              ctx, c, geo, I0, UInt64(seed), PRECISION[precision], A, Irt, Icb, st)
    end
    check(rc, ctx)
    out = (A[1], A[2], A[3], A[4], A[5], Irt, Icb)
    return stats ? (out..., st[]) : out
end

# ---------------------------------------------------------------- tables
"""
    pin_tables(f, scatterer)

Upload (or find) the scatterer's device tables and keep them resident while
`f()` runs — the counterpart of `run_rt`'s `adapt(CuArray, …)` before its
Stokes-column loop (:97-104, :194-199).  Every simulate of the context finds
pinned tables by their content.
"""
function pin_tables(f, s)
    c, k = mcrt_scatterer(s)
    h = Ref{UInt64}(0)
    ctx = hip_context()
    GC.@preserve k begin
        check(ccall((:mcrt_table_upload, libmcrt), Cint, (Ptr{Cvoid}, Ref{McrtScatterer}, Ref{UInt64}),
                    ctx, c, h), ctx)
    end
    try
        return f()
    finally
        ccall((:mcrt_table_release, libmcrt), Cint, (Ptr{Cvoid}, UInt64), ctx, h[])
    end
end

# ---------------------------------------------------------------- batches
"""
    simulate_batch_hip(jobs; precision = :mixed)

Many solves in one persistent launch: `jobs` is a vector of
`(cfg, g, I₀, seed)`; returns one 7-tuple per job, as simulate_hip.
"""
function simulate_batch_hip(jobs; precision::Symbol = :mixed)
    keep = Any[]
    cs = Ref{McrtConfig}[]
    gs = Ref{McrtGeometry}[]
    for (cfg, g, _, _) in jobs
        c, kc = mcrt_config(cfg)
        geo, kg = mcrt_geometry(g)
        push!(cs, Ref(c)); push!(gs, Ref(geo)); push!(keep, kc, kg)
    end
    ctx = hip_context()
    out = Vector{Any}(undef, length(jobs))
    GC.@preserve keep cs gs begin
        reqs = [McrtRequest(Base.unsafe_convert(Ptr{McrtConfig}, cs[i]),
                            Base.unsafe_convert(Ptr{McrtGeometry}, gs[i]),
                            Tuple(Float64.(jobs[i][3])), UInt64(jobs[i][4]), 0, 0,
                            PRECISION[precision], 0, 0, 0) for i in eachindex(jobs)]
        b = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:mcrt_batch_create, libmcrt), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{McrtRequest}, Ref{Ptr{Cvoid}}),
                    ctx, length(reqs), reqs, b), ctx)
        try
            check(ccall((:mcrt_batch_launch, libmcrt), Cint, (Ptr{Cvoid},), b[]), ctx)
            check(ccall((:mcrt_batch_wait, libmcrt), Cint, (Ptr{Cvoid},), b[]), ctx)
            for (i, (cfg, _, _, _)) in enumerate(jobs)
                A = zeros(5)
                Irt = zeros(4, cfg.Nθ, cfg.Nϕ)
                Icb = zeros(4, max(cfg.Nθb, 1), max(cfg.Nϕb, 1))
                check(ccall((:mcrt_batch_read, libmcrt), Cint,
                            (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                             Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                            b[], i - 1, A, Irt, Icb, C_NULL, C_NULL, C_NULL, C_NULL), ctx)
                out[i] = (A[1], A[2], A[3], A[4], A[5], Irt, Icb)
            end
        finally
            ccall((:mcrt_batch_destroy, libmcrt), Cvoid, (Ptr{Cvoid},), b[])
        end
    end
    return out
end
