# LangevInMediumHIP.jl -- the MI355X backend as an add-on module of LangevInMedium.jl.
#
# It adds two backend types next to CPUBackend/GPUBackend (src/backends.jl:14-44) and methods of
# LangevInMedium.Simulate.simulate_ensemble_bulk for them, the way src/simulate_gpu_wrapper.jl:15-108
# adds the CUDA ones.  The host preparation is the reference's own (Utils, Transport); every step runs
# in liblmhip.so through ccall (C ABI: include/lmhip.h).  Keyword arguments and return shapes follow
# simulate_cpu.jl:12-44,318-322 and simulate_cpu_general_coords.jl:18-29,128-132.
#
#   using LangevInMedium, LangevInMediumHIP
#   t, mom, pos = simulate_ensemble_bulk(HIPBackend(), r, p, dens, Tevo, Vevo, (xgrid, tgrid);
#                                        N_particles = N, ...)
#
# Keyword arguments beyond the reference's (extensions of the backend):
#   drag_model  :table (the reference's eta_D(T) m/E) or (:gamma_p, gamma0, Lambda):
#               eta(p, T) = gamma0 T^2/(1 + p^2/Lambda^2)
#   noise       :philox, or a fixed-noise record: an Array{Float64,3} of size (N, C, steps)
#               (Julia's column-major layout of the ABI's (steps, C, N) record; C from
#               lmh_noise_channels) read instead of the driver's randn draws
#   milne       Milne coordinates with dimensions = 2, momentum_dimensions = 3 (the reference's
#               Bjorken mode, bjorken_redshift implied): the positions gain the space-time rapidity
#               row eta_s (x_init may hold it as a third row, else eta_s = 0)
module LangevInMediumHIP

using LangevInMedium
import LangevInMedium.Simulate: simulate_ensemble_bulk
using LangevInMedium.Backends: AbstractBackend
using LangevInMedium.Transport: build_tau_drag_spline, build_taun_current_spline, build_juttner_invcdf
using LangevInMedium.Utils: check_momentum_dims, sample_initial_particles_milne!
import ChainRulesCore
using ChainRulesCore: NoTangent

export HIPBackend, HIP_GCBackend, GridEnsemble, final_observables

const liblmhip = get(ENV, "LMHIP_LIB", joinpath(@__DIR__, "..", "liblmhip.so"))

struct HIPBackend <: AbstractBackend end
struct HIP_GCBackend <: AbstractBackend end

# lmh_params (include/lmhip.h), field for field
struct LmhParams
    dt::Float64; t0::Float64; DsT::Float64; psi2::Float64; gamma0::Float64; Lambda::Float64
    seed::UInt64; stream::UInt64
    coords::Int32; collision::Int32; noise::Int32; integrator::Int32; drag_model::Int32
    relativistic::Int32; bjorken_redshift::Int32; position_diffusion::Int32
    reflecting_boundary::Int32; check_finite::Int32
end

const LMH_CARTESIAN, LMH_GENERAL_COORDS, LMH_MILNE = Int32(0), Int32(1), Int32(2)
const LMH_LANGEVIN, LMH_RTA, LMH_FLUID = Int32(0), Int32(1), Int32(2)
const LMH_NOISE_FIXED, LMH_NOISE_PHILOX = Int32(0), Int32(1)
const LMH_DRAG_TABLE, LMH_DRAG_GAMMA_P = Int32(0), Int32(1)
const LMH_TABLE_DRAG, LMH_TABLE_RTA = Int32(0), Int32(1)

mutable struct Ctx
    h::Ptr{Cvoid}
end

function Ctx(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:lmh_create, liblmhip), Cint, (Ref{Ptr{Cvoid}}, Int32), r, device)
    rc == 0 || error("lmh_create(device=$device) failed with status $rc")
    c = Ctx(r[])
    finalizer(x -> ccall((:lmh_destroy, liblmhip), Cvoid, (Ptr{Cvoid},), x.h), c)
    return c
end

check(c::Ctx, rc) = rc == 0 ||
    error(unsafe_string(ccall((:lmh_last_error, liblmhip), Cstring, (Ptr{Cvoid},), c.h)))

set_medium!(c, xg, tg, T, V, V2) = GC.@preserve xg tg T V V2 check(c, ccall((:lmh_set_medium_radial, liblmhip), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
    c.h, xg, length(xg), tg, length(tg), T, V, V2 === nothing ? C_NULL : V2))

set_table!(c, which, sp, Tmin, invdT, vals) = GC.@preserve vals check(c, ccall((:lmh_set_time_table, liblmhip), Cint,
    (Ptr{Cvoid}, Int32, Int32, Float64, Float64, Ptr{Float64}, Int32), c.h, which, sp, Tmin, invdT, vals, length(vals)))

set_juttner!(c, sp, inv, nU, nT, Tmin, invdT) = GC.@preserve inv check(c, ccall((:lmh_set_juttner_table, liblmhip), Cint,
    (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Int32, Float64, Float64), c.h, sp, inv, nU, nT, Tmin, invdT))

function set_particles!(c, pos::Matrix{Float64}, mom::Matrix{Float64}, m::Float64)
    N = size(pos, 2)
    counts = Int64[N]; masses = Float64[m]
    GC.@preserve pos mom counts masses check(c, ccall((:lmh_set_particles, liblmhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32, Int64, Int32, Ptr{Int64}, Ptr{Float64}),
        c.h, pos, size(pos, 1), mom, size(mom, 1), N, 1, counts, masses))
end

function get_particles(c, dims, pdim, N)
    pos = zeros(dims, N); mom = zeros(pdim, N)
    check(c, ccall((:lmh_get_particles, liblmhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), c.h, pos, mom))
    return pos, mom
end

run!(c, p::LmhParams, step0, n) = check(c, ccall((:lmh_run, liblmhip), Cint,
    (Ptr{Cvoid}, Ref{LmhParams}, Int64, Int64, Ptr{Float64}, Int32), c.h, p, step0, n, C_NULL, 0))

# steps step0 .. step0+n-1 reading the fixed-noise record's columns of those steps
function run!(c, p::LmhParams, step0, n, noise::Array{Float64,3})
    rec = noise[:, :, step0:step0 + n - 1]
    GC.@preserve rec check(c, ccall((:lmh_run, liblmhip), Cint,
        (Ptr{Cvoid}, Ref{LmhParams}, Int64, Int64, Ptr{Float64}, Int32), c.h, p, step0, n, rec, 0))
end

noise_channels(p::LmhParams, dims, pdim) =
    ccall((:lmh_noise_channels, liblmhip), Int32, (Ref{LmhParams}, Int32, Int32), p, dims, pdim)

function drag_params(drag_model)
    drag_model === :table && return LMH_DRAG_TABLE, 0.0, 1.0
    (drag_model isa Tuple && length(drag_model) == 3 && drag_model[1] === :gamma_p) ||
        error("drag_model must be :table or (:gamma_p, gamma0, Lambda)")
    return LMH_DRAG_GAMMA_P, Float64(drag_model[2]), Float64(drag_model[3])
end

boost_to_lab!(c, p::LmhParams, step) = check(c, ccall((:lmh_boost_to_lab, liblmhip), Cint,
    (Ptr{Cvoid}, Ref{LmhParams}, Int64), c.h, p, step))

# ---- initial state: utils.jl:33-172 and 499-558 with the draws on the device ---------------------
# lmh_fonll_tables (include/lmhip.h), field for field
struct LmhFonllTables
    r_grid::Ptr{Float64}; p_grid::Ptr{Float64}; cdf_r::Ptr{Float64}; n_r::Ptr{Float64}
    cdf_p::Ptr{Float64}; p_ok::Ptr{Float64}
    nr::Int32; np::Int32; cartesian::Int32; pad_::Int32
    rmax::Float64; fmax::Float64
end

# utils.jl:60-67,134-141: first node 0, strictly increasing by at least eps, last node exactly 1
function _monotone_normalised!(cdf)
    cdf[1] = 0.0
    for k in 2:length(cdf)
        cdf[k] = max(cdf[k], cdf[k - 1] + eps(Float64))
    end
    cdf ./= cdf[end]
    cdf[end] = 1.0
    return cdf
end

# The inverse-CDF tables of sample_particles_from_FONLL (utils.jl:33-146), built as the reference
# builds them; the device draws from them (lmh_sample_fonll).
function sample_fonll(c, r_grid, p_grid, f, N; cartesian::Bool, antithetic::Bool, seed)
    rg = collect(Float64, r_grid); pg = collect(Float64, p_grid)
    Nr, Np = length(rg), length(pg)
    F = size(f) == (Np, Nr) ? Matrix{Float64}(f) : size(f) == (Nr, Np) ? Matrix{Float64}(permutedims(f)) :
        error("sample_particles_from_FONLL: f_HQ_init_FONLL has size $(size(f)), expected ($Np, $Nr) (p×r) or ($Nr, $Np) (r×p).")
    dr = sum(diff(rg)) / (Nr - 1); dp = sum(diff(pg)) / (Np - 1)
    P = pg .* F
    cdf_p = zeros(Np, Nr); p_ok = zeros(Nr)
    for i in 1:Nr
        if sum(P[:, i]) > 0
            cdf_p[:, i] .= _monotone_normalised!(cumsum(P[:, i]) .* dp)
            p_ok[i] = 1.0
        end
    end
    cdf_r = zeros(Nr); n_r = zeros(Nr); rmax = 0.0; fmax = 0.0
    if cartesian
        n_r .= vec(sum(P; dims = 1)) .* dp
        rmax = rg[end]; fmax = maximum(n_r) + eps(Float64)
    else
        Prp = rg' .* pg .* F
        Z = sum(Prp) * dr * dp
        Z == 0 && error("Distribution normalization is zero.")
        cdf_r .= _monotone_normalised!(cumsum(vec(sum(Prp ./ Z; dims = 1)) .* dp) .* dr)
    end
    x = zeros(2, N); p = zeros(2, N)
    GC.@preserve rg pg cdf_r n_r cdf_p p_ok x p begin
        t = LmhFonllTables(pointer(rg), pointer(pg), pointer(cdf_r), pointer(n_r), pointer(cdf_p), pointer(p_ok),
                           Nr, Np, cartesian, 0, rmax, fmax)
        check(c, ccall((:lmh_sample_fonll, liblmhip), Cint,
            (Ptr{Cvoid}, Ref{LmhFonllTables}, Int64, Int32, UInt64, UInt64, Ptr{Float64}, Ptr{Float64}),
            c.h, t, N, antithetic, UInt64(seed), UInt64(0), x, p))
    end
    return x, p
end

# append_thermal_pz (utils.jl:533-558) on the device: T(|x|, t0) of the context's radial medium
function append_thermal_pz_device(c, mom::Matrix{Float64}, pos::Matrix{Float64}, m, t0; antithetic::Bool, seed)
    N = size(mom, 2)
    out = zeros(3, N)
    GC.@preserve mom pos out check(c, ccall((:lmh_append_thermal_pz, liblmhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int32, Int64, Float64, Float64, Int32, UInt64, UInt64, Ptr{Float64}),
        c.h, mom, pos, size(pos, 1), N, m, t0, antithetic, UInt64(seed), UInt64(0), out))
    return out
end

# ---- Cartesian: simulate_gpu_wrapper.jl:15-83 signature ----------------------------------------
function simulate_ensemble_bulk(::HIPBackend, r_grid_Langevin, p_grid_Langevin, heavy_quark_density,
        TemperatureEvolutionn, VelocityEvolutionn, SpaceTimeGrid;
        N_particles::Int = 10_000, Δt::Float64 = 0.001, initial_time::Float64 = 0.0,
        final_time::Float64 = 1.0, save_interval::Float64 = 0.1, m::Float64 = 1.0, DsT::Float64 = 0.2,
        DsT_linear::Bool = false, DsT_slope::Float64 = 1.765, DsT_offset::Float64 = -0.159,
        Tfo::Float64 = 0.156, DsT_quad::Bool = false, DsT_Tref::Float64 = 0.0, dimensions::Int = 3,
        cartesian_spatial_sampling = nothing, antithetic_momenta::Bool = false,
        position_diffusion::Bool = false, momentum_langevin::Bool = true,
        reflecting_boundary::Bool = false, collision_mode::Symbol = :langevin,
        x_init = nothing, p_init = nothing, V2Evolutionn = nothing, psi2::Float64 = 0.0,
        freezeout_capture::Bool = false, freezeout_interp::Bool = true, integrator_mode::Int = 0,
        relativistic::Bool = true, momentum_dimensions::Int = 0, bjorken_redshift::Bool = false,
        seed::Integer = 0, device::Integer = 0, drag_model = :table, noise = :philox, milne::Bool = false)
    (collision_mode == :langevin || collision_mode == :rta) ||
        error("simulate_ensemble_bulk(::HIPBackend): collision_mode=$(collision_mode) is not supported on GPU (only :langevin and :rta).")
    if milne
        (dimensions == 2 && momentum_dimensions == 3) ||
            error("milne needs dimensions = 2 and momentum_dimensions = 3 (x, y, eta_s; p_x, p_y, p_z)")
        (freezeout_capture || position_diffusion || reflecting_boundary) &&
            error("milne: no freeze-out capture, position diffusion or reflecting boundary")
        bjorken_redshift = true
    end
    dmodel, gamma0, Lambda = drag_params(drag_model)
    steps = floor(Int, (final_time - initial_time) / Δt)
    save_every = round(Int, save_interval / Δt)
    num_saves = div(steps, save_every)
    xgrid, tgrid = SpaceTimeGrid
    c = Ctx(device)
    # the medium first: the thermal p_z completion reads T(r, t0) from it
    set_medium!(c, collect(Float64, xgrid), collect(Float64, tgrid), Matrix{Float64}(TemperatureEvolutionn),
                Matrix{Float64}(VelocityEvolutionn), V2Evolutionn === nothing ? nothing : Matrix{Float64}(V2Evolutionn))
    eta0 = nothing
    if x_init !== nothing && p_init !== nothing
        x_matrix = Matrix{Float64}(x_init); p_matrix = Matrix{Float64}(p_init)
        if milne && size(x_matrix, 1) == 3
            eta0 = x_matrix[3:3, :]; x_matrix = x_matrix[1:2, :]
        end
    else
        # simulate_cpu.jl:57-94: Cartesian sampling by default, antithetic pairs on request
        cart = cartesian_spatial_sampling === nothing ? (dimensions == 1 || dimensions >= 2) : cartesian_spatial_sampling
        x_matrix, p_matrix = sample_fonll(c, r_grid_Langevin, p_grid_Langevin, heavy_quark_density, N_particles;
                                          cartesian = cart, antithetic = antithetic_momenta, seed = seed)
    end
    radial_mode = dimensions == 1
    if radial_mode
        # simulate_cpu.jl:102-136: r = |x|, p_r = p . e_r, a random e_r on the axis
        r = sqrt.(x_matrix[1, :] .^ 2 .+ x_matrix[2, :] .^ 2)
        r_axis_eps = max(1e-12, 0.5 * (length(r_grid_Langevin) >= 2 ? abs(float(r_grid_Langevin[2] - r_grid_Langevin[1])) : 0.0))
        reg = r .> r_axis_eps
        ex = similar(r); ey = similar(r)
        ex[reg] .= x_matrix[1, reg] ./ r[reg]; ey[reg] .= x_matrix[2, reg] ./ r[reg]
        ns = count(.!reg)
        if ns > 0
            rx = randn(ns); ry = randn(ns)
            invn = 1.0 ./ (sqrt.(rx .^ 2 .+ ry .^ 2) .+ eps())
            ex[.!reg] .= rx .* invn; ey[.!reg] .= ry .* invn
        end
        positions = reshape(r, 1, :)
        momenta = reshape(ex .* p_matrix[1, :] .+ ey .* p_matrix[2, :], 1, :)
    else
        positions = copy(x_matrix); momenta = copy(p_matrix)
    end
    pdim = momentum_dimensions <= 0 ? dimensions : momentum_dimensions
    check_momentum_dims(dimensions, pdim, radial_mode, bjorken_redshift, initial_time)
    if size(momenta, 1) < pdim
        # simulate_cpu.jl:147-151
        momenta = append_thermal_pz_device(c, Matrix{Float64}(momenta), Matrix{Float64}(positions), m, initial_time;
                                           antithetic = antithetic_momenta && x_init === nothing, seed = seed)
    end
    size(momenta, 1) == pdim || error("momentum rows $(size(momenta, 1)) ≠ momentum_dimensions $pdim")
    N = size(positions, 2)
    Tmin = max(float(minimum(TemperatureEvolutionn)), 0.0)
    Tmax = max(float(maximum(TemperatureEvolutionn)), Tmin + eps(Float64))
    if momentum_langevin && DsT > 0.0
        kw = (Tmin = Tmin, Tmax = Tmax, nT = 1024, DsT_linear = DsT_linear, DsT_slope = DsT_slope,
              DsT_offset = DsT_offset, Tfo = Tfo, DsT_quad = DsT_quad, DsT_Tref = DsT_Tref)
        if dmodel == LMH_DRAG_TABLE
            set_table!(c, LMH_TABLE_DRAG, 0, build_tau_drag_spline(m, DsT; kw...)...)
        else
            set_table!(c, LMH_TABLE_DRAG, 0, 0.0, 1.0, [0.0, 0.0])
        end
        if collision_mode == :rta
            set_table!(c, LMH_TABLE_RTA, 0, build_taun_current_spline(m, DsT; kw...)...)
            inv, nU, nT, Tt, idt = build_juttner_invcdf(m, pdim; Tmin = Tmin, Tmax = Tmax, nT = 256, nU = 2048, np = 8192)
            set_juttner!(c, 0, inv, nU, nT, Tt, idt)
        end
    else
        set_table!(c, LMH_TABLE_DRAG, 0, 0.0, 1.0, [0.0, 0.0])
    end
    if milne
        positions = vcat(positions, eta0 === nothing ? zeros(1, N) : eta0)
        dimensions = 3
    end
    set_particles!(c, positions, momenta, m)
    coll = (!momentum_langevin || DsT == 0.0) ? LMH_FLUID : (collision_mode == :rta ? LMH_RTA : LMH_LANGEVIN)
    fixed = noise isa AbstractArray
    p = LmhParams(Δt, initial_time, DsT, psi2, gamma0, Lambda, UInt64(seed), 0,
                  milne ? LMH_MILNE : LMH_CARTESIAN, coll, fixed ? LMH_NOISE_FIXED : LMH_NOISE_PHILOX,
                  Int32(integrator_mode), dmodel, relativistic, bjorken_redshift,
                  position_diffusion, reflecting_boundary, 1)
    if fixed
        C = noise_channels(p, dimensions, pdim)
        size(noise) == (N, C, steps) ||
            error("the fixed-noise record must be (N, C, steps) = ($N, $C, $steps), got $(size(noise))")
        noise = Array{Float64,3}(noise)
    else
        noise === :philox || error("noise must be :philox or a fixed-noise record")
    end
    step_run! = fixed ? ((s, n) -> run!(c, p, s, n, noise)) : ((s, n) -> run!(c, p, s, n))
    boost_to_lab!(c, p, 0)
    if freezeout_capture
        # simulate_gpu.jl:68-75,375-402: no history; the latched states are the result
        check(c, ccall((:lmh_freezeout_enable, liblmhip), Cint, (Ptr{Cvoid}, Float64, Int32), c.h, Tfo, freezeout_interp))
        step_run!(1, steps)
        fpos = zeros(dimensions, N); fmom = zeros(pdim, N); ftau = zeros(N); fflag = zeros(N)
        check(c, ccall((:lmh_freezeout_get, liblmhip), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), c.h, fpos, fmom, ftau, fflag))
        return (pos = fpos, mom = fmom, tau = ftau, flag = fflag)
    end
    pos0, mom0 = get_particles(c, dimensions, pdim, N)
    pos_hist = [pos0]; mom_hist = [mom0]
    step = 1
    while step <= steps
        n = min(save_every, steps - step + 1)
        step_run!(step, n)
        step += n
        if n == save_every && length(pos_hist) <= num_saves
            ps, ms = get_particles(c, dimensions, pdim, N)
            push!(pos_hist, ps); push!(mom_hist, ms)
        end
    end
    return range(initial_time, final_time, length = num_saves + 1), mom_hist, pos_hist
end

# ---- general coordinates (tau, r): simulate.jl:117-141 signature ---------------------------------
function simulate_ensemble_bulk(::HIP_GCBackend, T_profile_MIS, ur_profile_MIS, mu_profile_MIS,
        TemperatureEvolutionn, VelocityEvolutionn, SpaceTimeGrid;
        N_particles::Int = 10_000, Δt::Float64 = 0.001, initial_time::Float64 = 0.0,
        final_time::Float64 = 1.0, save_interval::Float64 = 0.1, m::Float64 = 1.0,
        DsT::Float64 = 0.2, dimensions::Int = 2, seed::Integer = 0, device::Integer = 0)
    steps = floor(Int, (final_time - initial_time) / Δt)
    save_every = Int(save_interval / Δt)
    num_saves = div(steps, save_every)
    xgrid, tgrid = SpaceTimeGrid
    positions, momenta = sample_initial_particles_milne!(m, dimensions, N_particles, initial_time,
        T_profile_MIS, ur_profile_MIS, mu_profile_MIS, (0., maximum(xgrid)), 200)
    c = Ctx(device)
    set_medium!(c, collect(Float64, xgrid), collect(Float64, tgrid), Matrix{Float64}(TemperatureEvolutionn),
                Matrix{Float64}(VelocityEvolutionn), nothing)
    if DsT > 0.0
        Tmin = max(float(minimum(TemperatureEvolutionn)), 0.0)
        Tmax = max(float(maximum(TemperatureEvolutionn)), Tmin + eps(Float64))
        set_table!(c, LMH_TABLE_DRAG, 0, build_tau_drag_spline(m, DsT; Tmin = Tmin, Tmax = Tmax, nT = 1024)...)
    else
        set_table!(c, LMH_TABLE_DRAG, 0, 0.0, 1.0, [0.0, 0.0])
    end
    set_particles!(c, Matrix{Float64}(positions), Matrix{Float64}(momenta), m)
    p = LmhParams(Δt, initial_time, DsT, 0.0, 0.0, 1.0, UInt64(seed), 0, LMH_GENERAL_COORDS,
                  LMH_LANGEVIN, LMH_NOISE_PHILOX, 0, 0, 1, 0, 0, 0, 1)
    mom_hist = [abs.(momenta[2, :])]
    pos_hist = [reshape(positions[2, :], 1, :)]
    step = 1
    while step <= steps
        n = min(save_every, steps - step + 1)
        run!(c, p, step, n)
        step += n
        if n == save_every && length(mom_hist) <= num_saves
            ps, ms = get_particles(c, 2, 2, N_particles)
            push!(mom_hist, sqrt.(sum(ms .^ 2, dims = 1))[:])
            push!(pos_hist, reshape(ps[2, :], 1, :))
        end
    end
    return range(initial_time, final_time, length = num_saves + 1), mom_hist, pos_hist
end

# ---- differentiable final-state observables: the pathwise adjoint behind a ChainRulesCore.rrule --
# theta = (DsT, gamma0, Lambda).  The ensemble fixes the 2+1D medium, the initial state, the step
# count and the options; final_observables returns lmh_observables of the final lab momenta.
struct LmhObsSpec
    nbins::Int32; species::Int32; pt_min::Float64; pt_max::Float64; sigma::Float64
end
struct LmhAdjointSpec
    checkpoint_every::Int32; keep_tape::Int32; dtau_dDsT::Float64
end

struct GridEnsemble
    c::Ctx
    x0::Float64; dx::Float64; nx::Int; y0::Float64; dy::Float64; ny::Int
    tgrid::Vector{Float64}; T::Array{Float64,3}; vx::Array{Float64,3}; vy::Array{Float64,3}
    pos::Matrix{Float64}; mom::Matrix{Float64}; m::Float64
    dt::Float64; t0::Float64; nsteps::Int; drag_model::Int32; seed::UInt64
    checkpoint_every::Int; keep_tape::Bool; obs::LmhObsSpec
    current::Base.RefValue{Int}     # the forward pass whose adjoint record the context holds (0: none)
end

const _RECORDS = Ref(0)

function GridEnsemble(x0, dx, nx, y0, dy, ny, tgrid, T, vx, vy, pos, mom, m; dt, t0, nsteps,
                      drag_model = 0, seed = 0, checkpoint_every = 50, keep_tape = false, nbins = 40,
                      pt_min = 0.0, pt_max = 10.0, sigma = 0.0, device = 0)
    c = Ctx(device)
    GC.@preserve tgrid T vx vy check(c, ccall((:lmh_set_medium_grid2p1, liblmhip), Cint,
        (Ptr{Cvoid}, Float64, Float64, Int32, Float64, Float64, Int32, Ptr{Float64}, Int32,
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        c.h, x0, dx, nx, y0, dy, ny, tgrid, length(tgrid), T, vx, vy))
    return GridEnsemble(c, x0, dx, nx, y0, dy, ny, tgrid, T, vx, vy, Matrix{Float64}(pos),
                        Matrix{Float64}(mom), m, dt, t0, nsteps, Int32(drag_model), UInt64(seed),
                        checkpoint_every, keep_tape, LmhObsSpec(nbins, -1, pt_min, pt_max, sigma), Ref(0))
end

function _reset!(e::GridEnsemble, theta)
    e.current[] = 0                 # the context's adjoint record no longer describes its state
    DsT, g0, Lam = theta
    Tmin = max(minimum(e.T), 0.0); Tmax = max(maximum(e.T), Tmin + eps(Float64))
    e.drag_model == 0 && set_table!(e.c, LMH_TABLE_DRAG, 0, build_tau_drag_spline(e.m, DsT; Tmin = Tmin, Tmax = Tmax, nT = 1024)...)
    set_particles!(e.c, e.pos, e.mom, e.m)
    p = LmhParams(e.dt, e.t0, DsT, 0.0, g0, Lam, e.seed, 0, LMH_CARTESIAN, LMH_LANGEVIN,
                  LMH_NOISE_PHILOX, 0, e.drag_model, 1, 0, 0, 0, 0)
    boost_to_lab!(e.c, p, 0)
    return p
end

function _observables(e::GridEnsemble)
    out = zeros(5 + e.obs.nbins)
    check(e.c, ccall((:lmh_observables, liblmhip), Cint, (Ptr{Cvoid}, Ref{LmhObsSpec}, Ptr{Float64}),
                     e.c.h, e.obs, out))
    return out
end

function final_observables(theta::AbstractVector{<:Real}, e::GridEnsemble)
    p = _reset!(e, theta)
    run!(e.c, p, 1, e.nsteps)
    return _observables(e)
end

# the checkpointed forward pass at theta; the context then holds its adjoint record
function _adjoint_forward!(e::GridEnsemble, theta, record)
    p = _reset!(e, theta)
    spec = LmhAdjointSpec(e.checkpoint_every, Int32(e.keep_tape), theta[1] > 0 ? 1.0 / theta[1] : 0.0)
    check(e.c, ccall((:lmh_adjoint_forward, liblmhip), Cint,
        (Ptr{Cvoid}, Ref{LmhParams}, Int64, Int64, Ref{LmhAdjointSpec}), e.c.h, p, 1, e.nsteps, spec))
    e.current[] = record
end

function ChainRulesCore.rrule(::typeof(final_observables), theta::AbstractVector{<:Real}, e::GridEnsemble)
    th = collect(Float64, theta)
    record = (_RECORDS[] += 1)
    _adjoint_forward!(e, th, record)
    obs = _observables(e)
    function final_observables_pullback(dobs)
        w = collect(Float64, ChainRulesCore.unthunk(dobs))
        # another forward pass (or a plain run) replaced this one's record: run it again first
        e.current[] == record || _adjoint_forward!(e, th, record)
        g = zeros(3)
        check(e.c, ccall((:lmh_adjoint_backward, liblmhip), Cint,
            (Ptr{Cvoid}, Ref{LmhObsSpec}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            e.c.h, e.obs, w, g, C_NULL, C_NULL))
        return NoTangent(), g, NoTangent()
    end
    return obs, final_observables_pullback
end

end # module
