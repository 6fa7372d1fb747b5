# direct_solve_method = :hip / :hipmixed for Clarabel.Solver: the whole
# interior-point loop on the MI355X through libclarabel_hip.so.
#
# The reference builds its GPU objects in setup! (src/solver.jl:150-166:
# DefaultProblemDataGPU, CompositeConeGPU, DefaultKKTSystemGPU) and runs each
# IPM step as its own launch from solve! (src/solver.jl:231-422).  Here setup!
# hands the problem to clh_solver_create (structure analysis, device arena)
# and solve! runs clh_solve, the whole loop on the device; the solver's data
# field holds the handle.  Routing, in Clarabel itself:
#   setup!:  s.settings.direct_solve_method in (:hip, :hipmixed) &&
#                return hip_setup!(s, P, q, A, b, cones)
#   solve!:  is_hip(s.data) && return hip_solve!(s)
#   update_q! (and P, A, b):  is_hip(s.data) && return hip_update_q!(s, q)
# with those functions declared in Clarabel (is_hip(x) = false, the others
# without methods) and given their methods here.

mutable struct HIPProblemData{T} <: Clarabel.AbstractProblemData{T}
    handle::Ptr{Cvoid}
    n::Int
    m::Int
    nnzP::Int
    nnzA::Int
    function HIPProblemData{T}(h, n, m, nnzP, nnzA) where {T}
        d = new{T}(h, n, m, nnzP, nnzA)
        finalizer(d) do x
            x.handle == C_NULL || ccall((:clh_solver_destroy, libclh[]), Cvoid, (Ptr{Cvoid},), x.handle)
            x.handle = C_NULL
        end
        return d
    end
end

# cone codes of include/clarabel_hip.h, in the order of src/cones/cone_api.jl:24-56
cone_code(c::Clarabel.ZeroConeT) = (CLH_ZERO, Int32(c.dim), 0.0)
cone_code(c::Clarabel.NonnegativeConeT) = (CLH_NONNEG, Int32(c.dim), 0.0)
cone_code(c::Clarabel.SecondOrderConeT) = (CLH_SOC, Int32(c.dim), 0.0)
cone_code(c::Clarabel.ExponentialConeT) = (CLH_EXP, Int32(3), 0.0)
cone_code(c::Clarabel.PowerConeT) = (CLH_POW, Int32(3), Float64(c.α))
cone_code(c::Clarabel.PSDTriangleConeT) = (CLH_PSD, Int32(c.dim), 0.0)   # the matrix side
cone_code(c) = throw(ArgumentError("cone $(typeof(c)) is not supported by direct_solve_method = :hip"))

# 0-based Int32 CSC as the ABI takes it; P keeps its upper triangle (the
# library ignores entries below the diagonal, as triu(P) in src/problemdata.jl:19-23)
csc32(M::SparseMatrixCSC) = (Int32.(M.colptr .- 1), Int32.(M.rowval .- 1), Vector{Float64}(M.nzval))

Clarabel.is_hip(::HIPProblemData) = true

function Clarabel.hip_setup!(s::Clarabel.Solver{T}, P::SparseMatrixCSC{T}, q::Vector{T}, A::SparseMatrixCSC{T},
                    b::Vector{T}, cones; schedule::ClhSchedule = ClhSchedule()) where {T}
    T === Float64 || throw(ArgumentError("direct_solve_method = :hip needs Float64 data"))
    (m, n) = size(A)
    size(P) == (n, n) && length(q) == n && length(b) == m || throw(DimensionMismatch())
    s.info = Clarabel.DefaultInfo{T}()
    s.solution = Clarabel.DefaultSolution{T}(n, m, false)
    Pp, Pi, Px = csc32(triu(P))
    Ap, Ai, Ax = csc32(A)
    codes = map(cone_code, cones)
    ct = Int32[c[1] for c in codes]
    cd = Int32[c[2] for c in codes]
    cp = Float64[c[3] for c in codes]
    cs = Ref(to_clh(s.settings, schedule))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    t0 = time()
    check(ccall((:clh_solver_create, libclh[]), Cint,
                (Ref{Ptr{Cvoid}}, Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Int32, Ptr{Int32}, Ptr{Int32},
                 Ptr{Float64}, Ref{ClhSettings}),
                h, n, m, Pp, Pi, Px, q, Ap, Ai, Ax, b, length(ct), ct, cd, cp, cs), "clh_solver_create")
    s.data = HIPProblemData{T}(h[], n, m, length(Px), length(Ax))
    s.info.setup_time = time() - t0
    return s
end

function Clarabel.hip_solve!(s::Clarabel.Solver{T}) where {T}
    d = s.data::HIPProblemData{T}
    check(ccall((:clh_solve, libclh[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), d.handle, C_NULL), "clh_solve")
    x = zeros(T, d.n); z = zeros(T, d.m); sv = zeros(T, d.m)
    info = Ref{ClhInfo}()
    check(ccall((:clh_get_solution, libclh[]), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{ClhInfo}),
                d.handle, x, sv, z, info), "clh_get_solution")
    i = info[]
    sol = s.solution
    sol.x = x; sol.z = z; sol.s = sv
    sol.status = Clarabel.SolverStatus(i.status)   # the same numbering (src/statuscodes.jl:25-37)
    sol.obj_val = i.obj_val; sol.obj_val_dual = i.obj_val_dual
    sol.setup_time = s.info.setup_time; sol.solve_time = i.solve_time
    sol.total_time = sol.setup_time + sol.solve_time
    sol.iterations = UInt32(i.iterations)
    sol.r_prim = i.r_prim; sol.r_dual = i.r_dual
    s.info.status = sol.status
    s.info.iterations = UInt32(i.iterations)
    s.info.cost_primal = i.obj_val; s.info.cost_dual = i.obj_val_dual
    s.info.res_primal = i.r_prim; s.info.res_dual = i.r_dual
    s.info.gap_abs = i.gap_abs; s.info.gap_rel = i.gap_rel; s.info.ktratio = i.ktratio; s.info.μ = i.mu
    s.info.solve_time = i.solve_time
    return s.solution
end

# update_P! / update_q! / update_A! / update_b! (src/data_updating.jl:26-202):
# new values for the pattern given at setup, kept in the first solve's
# equilibration
for (f, nm) in ((:hip_update_P!, :clh_update_P), (:hip_update_q!, :clh_update_q),
                (:hip_update_A!, :clh_update_A), (:hip_update_b!, :clh_update_b))
    @eval function Clarabel.$f(s::Clarabel.Solver{T}, v::AbstractVector{T}) where {T}
        d = s.data::HIPProblemData{T}
        vv = Vector{Float64}(v)
        check(ccall(($(QuoteNode(nm)), libclh[]), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.handle, vv), $(string(nm)))
        return s
    end
end

# the schedule chosen by timing candidates on the device (before the first
# solve; kept in the library's cache for later solvers with use_tuned)
function Clarabel.hip_autotune!(s::Clarabel.Solver; max_seconds::Real = 120)
    d = s.data::HIPProblemData
    r = Ref{ClhTuneReport}()
    check(ccall((:clh_solver_autotune, libclh[]), Cint, (Ptr{Cvoid}, Float64, Ref{ClhTuneReport}),
                d.handle, Float64(max_seconds), r), "clh_solver_autotune")
    return r[]
end
