# smearFEM_hip.jl -- ccall shim that routes smearFEM's assemble-and-solve path to
# the gfx950 library (include/smearfem.h).  Include it after `using smearFEM`
# (or in place of src/fem.jl's assemble_system) to get the same signatures:
#
#   assemble_system(ne, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=nothing, Young=1, ν=0.3)
#       -> SparseMatrixCSC{Float64,Int64}                       (src/fem.jl:135)
#   solve_system(K, f, fixed, q_d; precision=:fp64, rtol=1e-10)
#       -> q = q_d + C * (K_free \ C'(f - K q_d))               (examples/vector3D.jl:314-322)
#   assemble_and_solve(ne, NodeList, IEN, ndim, FunctionClass, nDof, ID, Young, ν, f, fixed, q_d;
#                      precision=:split, preconditioner=:block_jacobi)
#       -> q, with K kept on the device (no export), so the node-block formats apply
#   apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim, FunctionClass, ID, nDof=3)
#       -> b, the surface penalty                                (examples/vector3D.jl:175-264)
#   displacement_sweep(ne, deltas; Young=40, ν=0.4, β=100, rtol=1e-10)
#       -> [q(d) for d in deltas], the example's loop with all d solved together
#          (meshgrid + inflate_sphere + K + β*b + setboundaryCond, :270-322)
#
# gaussian_quadrature / basis_function stay pure Julia (src/fem.jl:21, :48).
# Status codes become the exceptions the reference raises on the same inputs.
module SmearFEMHip

using SparseArrays
using LinearAlgebra

export assemble_system, solve_system, assemble_and_solve, apply_boundary_conditions, displacement_sweep

const libsfem = get(ENV, "SMEARFEM_HIP_LIB",
                    joinpath(@__DIR__, "..", "smearfem", "libsmearfem_hip.so"))

const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function context()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:sfem_ctx_create, libsfem), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r)
        st == 0 || error("sfem_ctx_create failed with status $st")
        CTX[] = r[]
        atexit(() -> ccall((:sfem_ctx_destroy, libsfem), Cvoid, (Ptr{Cvoid},), CTX[]))
    end
    return CTX[]
end

function check(ctx, st)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:sfem_last_error, libsfem), Cstring, (Ptr{Cvoid},), ctx))
    st == -1 && throw(ArgumentError(msg))
    st == -2 && throw(BoundsError(msg))
    st == -3 && throw(DimensionMismatch(msg))
    st == -4 && throw(LinearAlgebra.SingularException(0))
    st == -5 && throw(UndefVarError(Symbol(msg)))
    error("smearfem_hip status $st: $msg")
end

function export_csc(ctx, Kp)
    m, n, nnz, bs = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0), Ref{Cint}(0)
    ccall((:sfem_csr_info, libsfem), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Cint}),
          Kp, m, n, nnz, bs)
    colptr = Vector{Int64}(undef, n[] + 1)
    rowval = Vector{Int64}(undef, nnz[])
    nzval = Vector{Float64}(undef, nnz[])
    st = ccall((:sfem_export_csc1, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
               ctx, Kp, colptr, rowval, nzval)
    check(ctx, st)
    return SparseMatrixCSC(m[], n[], colptr, rowval, nzval)
end

function assemble_system(ne, NodeList, IEN, ndim, FunctionClass="Q1", nDof=1, ID=nothing, Young=1, ν=0.3)
    ctx = context()
    nodes = Matrix{Float64}(NodeList)
    ien = Matrix{Int64}(IEN)
    idm = ID === nothing ? Matrix{Int64}(undef, 0, 0) : Matrix{Int64}(ID)
    young = Young isa AbstractVector ? Vector{Float64}(Young) : [Float64(Young)]
    cls = FunctionClass == "Q1" ? 1 : FunctionClass == "Q2" ? 2 : 0
    out = Ref{Ptr{Cvoid}}(C_NULL)
    st = GC.@preserve nodes ien idm young ccall((:sfem_assemble, libsfem), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Cint, Int64,
         Ptr{Int64}, Int64, Ptr{Float64}, Int64, Float64, Ref{Ptr{Cvoid}}),
        ctx, ndim, cls, nDof, nodes, size(nodes, 2), ien, size(ien, 1), size(ien, 2), ne^ndim,
        ID === nothing ? C_NULL : pointer(idm), size(idm, 1), young, length(young), ν, out)
    check(ctx, st)
    try
        return export_csc(ctx, out[])
    finally
        ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), out[])
    end
end

# Field order and types follow sfem_solve_opts / sfem_solve_stats.
mutable struct SolveOpts
    precision::Cint
    format::Cint
    rtol::Float64
    max_iter::Int64
    reliable_delta::Float64
    check_every::Cint
    time_kernels::Cint
    x0_zero::Cint
    preconditioner::Cint
end

mutable struct SolveStats
    iterations::Int64
    fp64_updates::Int64
    spmv_launches::Int64
    spmv64_launches::Int64
    rel_residual::Float64
    solve_ms::Float64
    spmv_ms::Float64
    spmv_bytes::Float64
    spmv_timed::Int64
    converged::Cint
    format::Cint
    stop_reason::Cint
    floor_residual::Float64
    smooth_ms::Float64
    smooth_timed::Int64
    smooth_bytes::Float64
end

"""
    solve_system(K, f, fixed, q_d; precision=:fp64, rtol=1e-10, max_iter=200000)

`fixed[i] == true` marks a prescribed dof (the columns C = I[:, free] drops);
`q_d` holds the prescribed values.  Returns q = q_d + C*q_f with
K_free q_f = C'(f - K q_d), solved by Jacobi-PCG on the GPU.
"""
function solve_system(K::SparseMatrixCSC{Float64,Int64}, f, fixed, q_d; precision=:fp64, rtol=1e-10,
                      max_iter=200000)
    ctx = context()
    n = size(K, 1)
    Kp = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:sfem_csr_from_csc1, libsfem), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
               ctx, size(K, 1), size(K, 2), K.colptr, K.rowval, K.nzval, Kp)
    check(ctx, st)
    fv = f === nothing ? zeros(n) : Vector{Float64}(vec(f))
    mask = Vector{UInt8}(fixed .!= 0)
    qd = q_d === nothing ? zeros(n) : Vector{Float64}(vec(q_d))
    opts = SolveOpts(precision_code(precision), 0, rtol, max_iter, 0.1, 64, 0, 1, 0)
    stats = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0.0, 0.0, 0, 0.0)
    q = zeros(n)
    try
        st = ccall((:sfem_solve, libsfem), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{UInt8}, Ptr{Float64}, Ref{SolveOpts},
                    Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
                   ctx, Kp[], fv, mask, qd, opts, q, C_NULL, stats)
        check(ctx, st)
    finally
        ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), Kp[])
    end
    return q
end

precision_code(p::Symbol) = p === :fp64 ? 0 : p === :mixed ? 1 : p === :fp32_matrix ? 2 : p === :split ? 3 :
                            throw(ArgumentError("unknown precision $p"))
preconditioner_code(p::Symbol) = p === :jacobi ? 0 : p === :block_jacobi ? 1 : p === :multigrid ? 2 :
                                 throw(ArgumentError("unknown preconditioner $p"))

# the one-shot solve of sfem_solve, split so that the multigrid hierarchy of a
# structured grid (dims elements per direction; nothing: the grid sfem_assemble
# found K's mesh to be) can be built in between
function solve_with_grid(ctx, Kp, fv, mask, qd, opts, dims)
    n = length(fv)
    bc, sys = Ref{Ptr{Cvoid}}(C_NULL), Ref{Ptr{Cvoid}}(C_NULL)
    stats = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0.0, 0.0, 0, 0.0)
    q = zeros(n)
    d = dims === nothing ? Int64[] : Int64[dims...]
    try
        check(ctx, ccall((:sfem_bc_create, libsfem), Cint,
                         (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{UInt8}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                         ctx, n, fv, mask, qd, bc))
        check(ctx, ccall((:sfem_eliminate, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                         ctx, Kp, bc[], sys))
        GC.@preserve d check(ctx, ccall((:sfem_system_set_grid, libsfem), Cint,
                                        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Cvoid}),
                                        ctx, sys[], length(d), isempty(d) ? Ptr{Int64}(C_NULL) : pointer(d),
                                        C_NULL))
        check(ctx, ccall((:sfem_system_solve, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{SolveOpts}, Ref{SolveStats}),
                         ctx, sys[], opts, stats))
        check(ctx, ccall((:sfem_system_solution, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                         ctx, sys[], q, C_NULL))
    finally
        sys[] != C_NULL && ccall((:sfem_system_destroy, libsfem), Cvoid, (Ptr{Cvoid},), sys[])
        bc[] != C_NULL && ccall((:sfem_bc_destroy, libsfem), Cvoid, (Ptr{Cvoid},), bc[])
    end
    return q
end

"""
    assemble_and_solve(ne, NodeList, IEN, ndim, FunctionClass, nDof, ID, Young, ν, f, fixed, q_d;
                       precision=:split, preconditioner=:block_jacobi, rtol=1e-10, max_iter=200000,
                       grid=nothing)

assemble_system (src/fem.jl:135) and the example's elimination and solve
(examples/vector3D.jl:314-322) in one call; K stays on the device, so the
node-block operator formats (split precision, block-Jacobi) are available when
ID is node-interleaved.  `preconditioner=:multigrid` uses the geometric
multigrid preconditioner (precision :fp64 or :split) on the grid of the
meshgrid numbering the nodes follow: the one sfem_assemble recognised
(grid=nothing), or `grid=(nx, ny, nz)` elements per direction.
"""
function assemble_and_solve(ne, NodeList, IEN, ndim, FunctionClass, nDof, ID, Young, ν, f, fixed, q_d;
                            precision=:split, preconditioner=:block_jacobi, rtol=1e-10, max_iter=200000,
                            grid=nothing)
    ctx = context()
    nodes = Matrix{Float64}(NodeList)
    ien = Matrix{Int64}(IEN)
    idm = ID === nothing ? Matrix{Int64}(undef, 0, 0) : Matrix{Int64}(ID)
    young = Young isa AbstractVector ? Vector{Float64}(Young) : [Float64(Young)]
    cls = FunctionClass == "Q1" ? 1 : FunctionClass == "Q2" ? 2 : 0
    Kp = Ref{Ptr{Cvoid}}(C_NULL)
    st = GC.@preserve nodes ien idm young ccall((:sfem_assemble, libsfem), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Cint, Int64,
         Ptr{Int64}, Int64, Ptr{Float64}, Int64, Float64, Ref{Ptr{Cvoid}}),
        ctx, ndim, cls, nDof, nodes, size(nodes, 2), ien, size(ien, 1), size(ien, 2), ne^ndim,
        ID === nothing ? C_NULL : pointer(idm), size(idm, 1), young, length(young), ν, Kp)
    check(ctx, st)
    m = Ref{Int64}(0)
    ccall((:sfem_csr_info, libsfem), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Cint}),
          Kp[], m, C_NULL, C_NULL, C_NULL)
    n = m[]
    fv = f === nothing ? zeros(n) : Vector{Float64}(vec(f))
    mask = Vector{UInt8}(fixed .!= 0)
    qd = q_d === nothing ? zeros(n) : Vector{Float64}(vec(q_d))
    opts = SolveOpts(precision_code(precision), 0, rtol, max_iter, 0.1, 64, 0, 1,
                     preconditioner_code(preconditioner))
    if preconditioner === :multigrid
        try
            return solve_with_grid(ctx, Kp[], fv, mask, qd, opts, grid)
        finally
            ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), Kp[])
        end
    end
    stats = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0.0, 0.0, 0, 0.0)
    q = zeros(n)
    try
        st = ccall((:sfem_solve, libsfem), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{UInt8}, Ptr{Float64}, Ref{SolveOpts},
                    Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
                   ctx, Kp[], fv, mask, qd, opts, q, C_NULL, stats)
        check(ctx, st)
    finally
        ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), Kp[])
    end
    return q
end

"""
    apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim, FunctionClass, ID=nothing, nDof=3)

The surface penalty b of examples/vector3D.jl:175-264 assembled on the GPU
(IEN and FunctionClass are unused there too).  Same errors: BoundsError at
the first bad face, DimensionMismatch unless ndim = 3, 4-node faces, nDof = 3.
"""
function apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim, FunctionClass="Q1", ID=nothing,
                                   nDof=3)
    ctx = context()
    nodes = Matrix{Float64}(NodeList)
    top = Matrix{Int64}(IEN_top)
    btm = Matrix{Int64}(IEN_btm)
    idm = Matrix{Int64}(ID)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    st = GC.@preserve nodes top btm idm ccall((:sfem_assemble_surface, libsfem), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Cint, Ptr{Int64}, Int64,
         Ref{Ptr{Cvoid}}),
        ctx, ndim, nDof, nodes, size(nodes, 2), btm, top, size(btm, 1), size(btm, 2), idm, size(idm, 1), out)
    check(ctx, st)
    try
        return export_csc(ctx, out[])
    finally
        ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), out[])
    end
end

# sfem_batch_problem_stats (per right-hand side of sfem_system_solve_many)
struct ColumnStats
    iterations::Int64
    restarts::Int64
    rel_residual::Float64
    n_free::Int64
    stop_reason::Cint
    floor_residual::Float64
end

"""
    displacement_sweep(ne, deltas; Young=40, ν=0.4, β=100, rtol=1e-10, multigrid=iseven(ne))

The example's solve loop (examples/vector3D.jl:270-322) on the GPU: the unit
cube meshed with ne^3 Hex8 and inflated to the cylinder (K and b both on it,
as inflate_sphere overwrites NodeList there), K_bar = K + β b, u_z = 0 on
z = 0 and u_z = -d on z = 1, q(d) = q_d + C K_free^-1 C'(-K_bar q_d) for every d.
With `multigrid` the displacements are solved one after the other with the
geometric multigrid preconditioner of the ne^3 grid (built once; each solve
starts from the previous solution); otherwise together, up to 64 per pass
over K_free.  Returns a vector of q (3 (ne+1)^3 values each, node-interleaved
as ID).
"""
function displacement_sweep(ne, deltas; Young=40.0, ν=0.4, β=100.0, rtol=1e-10, max_iter=200000,
                            multigrid=iseven(ne))
    ctx = context()
    mesh, faces, K, b, Kbar, bc, sys = (Ref{Ptr{Cvoid}}(C_NULL) for _ in 1:7)
    dims = Int64[ne, ne, ne]
    box = Float64[0, 1, 0, 1, 0, 1]
    young = [Float64(Young)]
    try
        check(ctx, ccall((:sfem_meshgrid, libsfem), Cint,
                         (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Float64}, Cint, Ptr{Float64}, Int64, Float64,
                          Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                         ctx, 3, dims, box, 1, young, 1, ν, mesh, faces))
        check(ctx, ccall((:sfem_assemble_mesh, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}),
                         ctx, mesh[], 0, K))
        check(ctx, ccall((:sfem_assemble_mesh, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}),
                         ctx, faces[], 0, b))
        check(ctx, ccall((:sfem_csr_add, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                         ctx, K[], β, b[], Kbar))
        n = 3 * (ne + 1)^3
        # setboundaryCond's fixed set (:155-166): u_z on z == 0 and on z == 1
        axis, coord, dof = Cint[3, 3], Float64[0, 1], Cint[3, 3]
        value = Float64[0, -first(deltas)]
        check(ctx, ccall((:sfem_bc_planes, libsfem), Cint,
                         (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Ptr{Cint}, Ptr{Float64}, Ptr{Cint}, Ptr{Float64},
                          Ptr{Float64}, Ref{Ptr{Cvoid}}),
                         ctx, mesh[], n, 2, axis, coord, dof, value, C_NULL, bc))
        check(ctx, ccall((:sfem_eliminate, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                         ctx, Kbar[], bc[], sys))
        if multigrid
            check(ctx, ccall((:sfem_system_set_grid, libsfem), Cint,
                             (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Cvoid}), ctx, sys[], 3, dims, C_NULL))
            qs = Vector{Vector{Float64}}()
            stats = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0.0, 0.0, 0, 0.0)
            for (i, d) in enumerate(deltas)
                if i > 1
                    bcd = Ref{Ptr{Cvoid}}(C_NULL)
                    value[2] = -d
                    check(ctx, ccall((:sfem_bc_planes, libsfem), Cint,
                                     (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Ptr{Cint}, Ptr{Float64}, Ptr{Cint},
                                      Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                                     ctx, mesh[], n, 2, axis, coord, dof, value, C_NULL, bcd))
                    try
                        check(ctx, ccall((:sfem_system_set_bc, libsfem), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                                         ctx, sys[], bcd[]))
                    finally
                        ccall((:sfem_bc_destroy, libsfem), Cvoid, (Ptr{Cvoid},), bcd[])
                    end
                end
                mgopts = SolveOpts(0, 0, rtol, max_iter, 0.1, 8, 0, i == 1 ? 1 : 0, 2)
                check(ctx, ccall((:sfem_system_solve, libsfem), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ref{SolveOpts}, Ref{SolveStats}), ctx, sys[], mgopts, stats))
                q = zeros(n)
                check(ctx, ccall((:sfem_system_solution, libsfem), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx, sys[], q, C_NULL))
                push!(qs, q)
            end
            return qs
        end
        # node z coordinates of the generated mesh (x fastest, then y, then z)
        z = [Float64(k) / ne for k in 0:ne for j in 0:ne for i in 0:ne]
        z[end - (ne + 1)^2 + 1:end] .= 1.0
        qs = Vector{Vector{Float64}}()
        opts = SolveOpts(0, 0, rtol, max_iter, 0.1, 32, 0, 1, 0)
        for group in Iterators.partition(collect(deltas), 64)
            k = length(group)
            qd = zeros(n, k)
            for (j, d) in enumerate(group), (m, zm) in enumerate(z)
                zm == 1.0 && (qd[3m, j] = -d)
            end
            q = zeros(n, k)
            per = Vector{ColumnStats}(undef, k)
            check(ctx, ccall((:sfem_system_solve_many, libsfem), Cint,
                             (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ref{SolveOpts}, Ptr{Float64},
                              Ptr{ColumnStats}, Ptr{Float64}),
                             ctx, sys[], k, qd, C_NULL, opts, q, per, C_NULL))
            append!(qs, [q[:, j] for j in 1:k])
        end
        return qs
    finally
        sys[] != C_NULL && ccall((:sfem_system_destroy, libsfem), Cvoid, (Ptr{Cvoid},), sys[])
        bc[] != C_NULL && ccall((:sfem_bc_destroy, libsfem), Cvoid, (Ptr{Cvoid},), bc[])
        for h in (Kbar, b, K)
            h[] != C_NULL && ccall((:sfem_csr_destroy, libsfem), Cvoid, (Ptr{Cvoid},), h[])
        end
        for h in (faces, mesh)
            h[] != C_NULL && ccall((:sfem_mesh_destroy, libsfem), Cvoid, (Ptr{Cvoid},), h[])
        end
    end
end

end # module
