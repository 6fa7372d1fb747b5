# The C ABI of libclarabel_hip.so (include/clarabel_hip.h) for Julia: struct
# mirrors with the header's field order and C layout, and ccall wrappers that
# turn a non-zero return code into an exception carrying clh_last_error().

const CLH_ZERO, CLH_NONNEG, CLH_SOC, CLH_EXP, CLH_POW, CLH_PSD = 0, 1, 2, 3, 4, 5
const CLH_METHOD_HIP, CLH_METHOD_HIPMIXED = 0, 1

# clh_schedule
struct ClhSchedule
    relax_max_cols::Int32
    relax_zero_pct::Int32
    group_lanes::Int32
    blocks_per_cu::Int32
    grid_blocks::Int32
    tile_min_cols::Int32
    dense_min_cols::Int32
    use_tuned::Int32
end
ClhSchedule() = ClhSchedule(0, 0, 0, 0, 0, 0, 0, 0)

# clh_settings: the fields of src/settings.jl:70-148 the GPU path uses, then
# the library's own (mixed-precision switch, engine, schedule)
struct ClhSettings
    max_iter::Int32
    time_limit::Float64
    verbose::Int32
    max_step_fraction::Float64
    tol_gap_abs::Float64
    tol_gap_rel::Float64
    tol_feas::Float64
    tol_infeas_abs::Float64
    tol_infeas_rel::Float64
    tol_ktratio::Float64
    reduced_tol_gap_abs::Float64
    reduced_tol_gap_rel::Float64
    reduced_tol_feas::Float64
    reduced_tol_infeas_abs::Float64
    reduced_tol_infeas_rel::Float64
    reduced_tol_ktratio::Float64
    equilibrate_enable::Int32
    equilibrate_max_iter::Int32
    equilibrate_min_scaling::Float64
    equilibrate_max_scaling::Float64
    linesearch_backtrack_step::Float64
    min_switch_step_length::Float64
    min_terminate_step_length::Float64
    static_regularization_enable::Int32
    static_regularization_constant::Float64
    static_regularization_proportional::Float64
    dynamic_regularization_enable::Int32
    dynamic_regularization_eps::Float64
    dynamic_regularization_delta::Float64
    iterative_refinement_enable::Int32
    iterative_refinement_reltol::Float64
    iterative_refinement_abstol::Float64
    iterative_refinement_max_iter::Int32
    iterative_refinement_stop_ratio::Float64
    direct_solve_method::Int32
    soc_threshold::Int32
    mixed_switch_reltol::Float64
    mixed_fp64_static_reg::Float64
    serve_engine::Int32
    dynamic_regularization_keep_magnitude::Int32
    schedule::ClhSchedule
end

# clh_info
struct ClhInfo
    status::Int32
    iterations::Int32
    obj_val::Float64
    obj_val_dual::Float64
    r_prim::Float64
    r_dual::Float64
    solve_time::Float64
    setup_time::Float64
    gap_abs::Float64
    gap_rel::Float64
    ktratio::Float64
    mu::Float64
    ir_steps::Int32
    kkt_solves::Int32
    nnz_L::Int64
    factor_flops::Float64
    fp32_factors::Int32
    switched_fp64::Int32
    nnz_K::Int64
    kkt_dim::Int32
    n_supernodes::Int32
    n_levels::Int32
    pad::Int32
    phase_seconds::NTuple{8,Float64}
end

# clh_tune_report
struct ClhTuneReport
    default_rate::Float64
    tuned_rate::Float64
    trials::Int32
    classes::Int32
    seconds::Float64
end

# clh_ldl_info
struct ClhLdlInfo
    nnz_L::Int64
    factor_flops::Float64
    n_supernodes::Int32
    n_levels::Int32
    n_positive::Int32
    n_negative::Int32
    last_dynreg::Int32
    last_ir_steps::Int32
    last_refine_error::Float64
    grid_blocks::Int32
    pad_ldl::Int32
end

last_error() = unsafe_string(ccall((:clh_last_error, libclh[]), Cstring, ()))

function check(rc::Integer, what::AbstractString)
    rc == 0 && return nothing
    msg = "$what: " * last_error()
    rc < 0 ? throw(ArgumentError(msg)) : error(msg)
end

function default_settings()
    r = Ref{ClhSettings}()
    check(ccall((:clh_default_settings, libclh[]), Cint, (Ref{ClhSettings},), r), "clh_default_settings")
    return r[]
end

# Clarabel.Settings -> clh_settings: every field with the same name is
# copied; :hip / :cudss give the FP64 method, :hipmixed / :cudssmixed the
# FP32-factor method (whose static regularizer the library sets to
# sqrt(eps(Float32)), as gpu_preprocess does, src/solver.jl:604-615)
function to_clh(s, schedule::ClhSchedule = ClhSchedule())
    c = default_settings()
    names = fieldnames(ClhSettings)
    vals = Any[getfield(c, f) for f in names]
    for (i, f) in enumerate(names)
        if f === :direct_solve_method
            vals[i] = Int32(s.direct_solve_method in (:hipmixed, :cudssmixed) ? CLH_METHOD_HIPMIXED : CLH_METHOD_HIP)
        elseif f === :schedule
            vals[i] = schedule
        elseif hasfield(typeof(s), f)
            vals[i] = convert(fieldtype(ClhSettings, f), getfield(s, f))
        end
    end
    return ClhSettings(vals...)
end
