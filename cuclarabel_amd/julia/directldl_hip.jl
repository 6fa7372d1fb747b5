# The LDL solver plugin on the MI355X: Clarabel's GPU LDL protocol
# (src/kktsolvers/gpu/directldl_cudss.jl:5-57: constructor on the full
# symmetric CSR KKT with the iterate vectors, refactor!(s) re-reading the
# values the caller edited in place, solve!(s, x, b)) on AMDGPU.jl arrays,
# registered as :hipldl (FP64 factors) and :hipldlmixed (FP32 factors and
# solves, the caller's refinement in FP64, as directldl_mixed_cudss.jl:5-78),
# for a KKT solver that drives its own loop on AMDGPU arrays.  (:hip and
# :hipmixed name the whole-solver path of solver_hip.jl.)
#
# The constructor takes the pivot signs as a fourth argument: cuDSS infers
# them, this factorization applies QDLDL's dynamic regularization, which
# needs them.  At the call site (kktsolver_directldl_gpu.jl:90) that is
#   GPUsolver = applicable(GPUsolverT{T}, KKT, x, b, Dsigns) ?
#               GPUsolverT{T}(KKT, x, b, Dsigns) : GPUsolverT{T}(KKT, x, b)
# and the KKT there is built as a ROCSparseMatrixCSR when the method is :hip.

struct HIPDirectLDLSolver{T} <: Clarabel.AbstractDirectLDLSolver{T}
    KKT::ROCSparseMatrixCSR{T}
    x::ROCVector{T}
    b::ROCVector{T}
    handle::Base.RefValue{Ptr{Cvoid}}
end
struct HIPMixedDirectLDLSolver{T} <: Clarabel.AbstractDirectLDLSolver{T}
    inner::HIPDirectLDLSolver{T}
end

function _ldl_create(KKT::ROCSparseMatrixCSR{Float64}, Dsigns, precision::Integer)
    LinearAlgebra.checksquare(KKT)
    n = size(KKT, 1)
    rowptr = Array(KKT.rowPtr)   # 1-based Int32: the structure analysis runs on the host, once
    colval = Array(KKT.colVal)
    signs = Int32.(Array(Dsigns))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:clh_ldl_create, libclh[]), Cint,
                (Ref{Ptr{Cvoid}}, Int32, Ptr{Int32}, Ptr{Int32}, Int32, Ptr{Int32}, Int32),
                h, n, rowptr, colval, 1, signs, precision), "clh_ldl_create")
    return h
end

function HIPDirectLDLSolver{T}(KKT::ROCSparseMatrixCSR{T}, x::ROCVector{T}, b::ROCVector{T}, Dsigns) where {T}
    T === Float64 || throw(ArgumentError(":hipldl factors Float64 KKT matrices"))
    s = HIPDirectLDLSolver{T}(KKT, x, b, _ldl_create(KKT, Dsigns, 0))
    finalizer(_ -> ccall((:clh_ldl_destroy, libclh[]), Cvoid, (Ptr{Cvoid},), s.handle[]), s.handle)
    refactor!(s)
    return s
end
function HIPMixedDirectLDLSolver{T}(KKT::ROCSparseMatrixCSR{T}, x::ROCVector{T}, b::ROCVector{T}, Dsigns) where {T}
    T === Float64 || throw(ArgumentError(":hipldlmixed takes the Float64 KKT matrix and factors it in Float32"))
    inner = HIPDirectLDLSolver{T}(KKT, x, b, _ldl_create(KKT, Dsigns, 1))
    finalizer(_ -> ccall((:clh_ldl_destroy, libclh[]), Cvoid, (Ptr{Cvoid},), inner.handle[]), inner.handle)
    s = HIPMixedDirectLDLSolver{T}(inner)
    refactor!(s)
    return s
end

Clarabel.ldlsolver_constructor(::Val{:hipldl}) = HIPDirectLDLSolver
Clarabel.ldlsolver_matrix_shape(::Val{:hipldl}) = :full
Clarabel.ldlsolver_constructor(::Val{:hipldlmixed}) = HIPMixedDirectLDLSolver
Clarabel.ldlsolver_matrix_shape(::Val{:hipldlmixed}) = :full

# refactor! (directldl_cudss.jl:33-45): the caller has regularized KKT.nzVal in
# place (_kktsolver_regularize_and_refactor!), so no static term is added
# here; dynamic regularization with QDLDL's constants (src/settings.jl)
function Clarabel.refactor!(s::HIPDirectLDLSolver{T}) where {T}
    rc = ccall((:clh_ldl_refactor, libclh[]), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Float64, Int32, Float64, Float64, Ptr{Cvoid}),
               s.handle[], pointer(s.KKT.nzVal), 0.0, 1, 1e-13, 2e-7, C_NULL)
    rc < 0 && check(rc, "clh_ldl_refactor")
    return rc == 0   # 1: a zero or non-finite pivot (the protocol's Bool)
end
Clarabel.refactor!(s::HIPMixedDirectLDLSolver) = Clarabel.refactor!(s.inner)

# solve! (directldl_cudss.jl:50-57)
function Clarabel.solve!(s::HIPDirectLDLSolver{T}, x::ROCVector{T}, b::ROCVector{T}) where {T}
    check(ccall((:clh_ldl_solve, libclh[]), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                s.handle[], pointer(x), pointer(b), C_NULL), "clh_ldl_solve")
    return nothing
end
Clarabel.solve!(s::HIPMixedDirectLDLSolver{T}, x::ROCVector{T}, b::ROCVector{T}) where {T} =
    Clarabel.solve!(s.inner, x, b)
