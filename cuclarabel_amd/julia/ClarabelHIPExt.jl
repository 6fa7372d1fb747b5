# ClarabelHIPExt: the MI355X engine behind Clarabel.jl, as a package extension
# in the pattern of ext/PardisoExt.jl (weak dependency AMDGPU; Clarabel's
# Project.toml gains
#     [weakdeps]    AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#     [extensions]  ClarabelHIPExt = "AMDGPU"
# and this directory's files go to ext/).  Two seams, as in the reference:
#  - direct_solve_method = :hip / :hipmixed for the whole solver
#    (solver_hip.jl: the interior-point loop on the device), and
#  - the LDL plugin protocol on AMDGPU arrays (directldl_hip.jl), for
#    Clarabel's own GPU KKT solver driving the factorization.
# The library is libclarabel_hip.so, found through ENV["CLARABEL_HIP_LIB"]
# (built by `python -m cuclarabel_amd.build`).
module ClarabelHIPExt

using Clarabel, SparseArrays, LinearAlgebra
using AMDGPU
using AMDGPU.rocSPARSE: ROCSparseMatrixCSR

const libclh = Ref{String}("libclarabel_hip.so")

function __init__()
    libclh[] = get(ENV, "CLARABEL_HIP_LIB", libclh[])
end

include("clh_abi.jl")
include("solver_hip.jl")
include("directldl_hip.jl")

end
