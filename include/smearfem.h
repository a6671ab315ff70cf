/*
 * smearfem.h -- C ABI of the MI355X (gfx950) assemble-and-solve library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status (SFEM_OK == 0, negative == error class).  No entry
 * point prints or exits.  The message of the last failure on a context is
 * available from sfem_last_error().
 *
 * Array conventions follow the Julia caller (column-major, 1-based ids):
 *   NodeList  ndim x n_nodes  Float64   x[d + ndim*node]
 *   IEN       ien_rows x nen  Int64     ien[a*ien_rows + e]   (1-based node ids)
 *   ID        id_rows x n_dof Int64     id[d*id_rows + node]  (1-based dof ids)
 * so a Julia caller passes its arrays as they are (zero copy on the host).
 *
 * Reference interfaces replaced (file:line in soshalaw/smearFEM.jl):
 *   sfem_gaussian_quadrature  <- gaussian_quadrature        src/fem.jl:21
 *   sfem_basis_function       <- basis_function             src/fem.jl:48
 *   sfem_assemble             <- assemble_system            src/fem.jl:135 (returns CSC via sfem_export_csc1)
 *   sfem_bc_create + sfem_eliminate
 *                             <- setboundaryCond + C'*K*C   examples/vector3D.jl:133-173, 314-316
 *   sfem_system_solve         <- inv(Matrix(K_free))*C'*(-K*q_d)   examples/vector3D.jl:318-320
 *   sfem_system_solution      <- q = q_d + C*q_f            examples/vector3D.jl:322
 *   sfem_meshgrid             <- meshgrid (+ inflate_sphere) examples/vector3D.jl:10-130, src/PostProcess.jl:30
 *   sfem_assemble_surface     <- apply_boundary_conditions  examples/vector3D.jl:175-264
 *   sfem_csr_add              <- K_bar = K + beta*b         examples/vector3D.jl:308
 *   sfem_bc_planes            <- setboundaryCond            examples/vector3D.jl:133-173
 *   sfem_system_set_bc / sfem_system_solve_many
 *                             <- the displacement sweep     examples/vector3D.jl:310-322
 *   sfem_system_set_grid      <- no counterpart: the multigrid preconditioner of the solve that
 *                                replaces inv(Matrix(K_free)) (examples/vector3D.jl:318)
 *
 * Threading: a context owns one HIP stream.  Calls on one context are not
 * re-entrant; different contexts may be used from different threads.
 */
#ifndef SMEARFEM_H
#define SMEARFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped by the Julia shim to the exceptions the
 *      reference raises on the same inputs) ---------------------------- */
#define SFEM_OK                 0
#define SFEM_ERR_ARGUMENT      -1  /* ArgumentError: dof id < 1 reaching sparse() (src/fem.jl:253), bad option  */
#define SFEM_ERR_BOUNDS        -2  /* BoundsError: IEN/ID index out of range (src/fem.jl:180,243), Q2 with nDof>=2 */
#define SFEM_ERR_DIMENSION     -3  /* DimensionMismatch: ndim==1 (src/fem.jl:75,192), nDof != ndim, nen mismatch */
#define SFEM_ERR_SINGULAR      -4  /* SingularException: det(Jac)==0 (src/fem.jl:195)                            */
#define SFEM_ERR_UNDEFVAR      -5  /* UndefVarError: nGaussPoints not in {2,3} (src/fem.jl:30), unknown class   */
#define SFEM_ERR_NOT_CONVERGED -6  /* solver reached max_iter before rtol                                       */
#define SFEM_ERR_MEMORY        -7  /* device allocation failed or HBM budget exceeded                           */
#define SFEM_ERR_DEVICE        -8  /* HIP runtime failure                                                       */
#define SFEM_ERR_UNSUPPORTED   -9  /* beyond the library's index limits (e.g. >= 2^31 dofs, >= 2^32 COO entries) */
#define SFEM_ERR_STATE        -10  /* handle used with the wrong context or after destroy                       */

/* FunctionClass ("Q1" / "Q2", src/fem.jl:50,77) */
#define SFEM_Q1 1
#define SFEM_Q2 2

typedef struct sfem_ctx    sfem_ctx;     /* device, stream, allocator, last error       */
typedef struct sfem_mesh   sfem_mesh;    /* device-resident mesh + material inputs     */
typedef struct sfem_csr    sfem_csr;     /* device-resident assembled K (block CSC)    */
typedef struct sfem_bc     sfem_bc;      /* device-resident loads + Dirichlet data     */
typedef struct sfem_system sfem_system;  /* device-resident K_ff, rhs, PCG state       */

/* ---- lifecycle ------------------------------------------------------- */
int         sfem_ctx_create(int device, sfem_ctx** out);
void        sfem_ctx_destroy(sfem_ctx* ctx);
const char* sfem_last_error(const sfem_ctx* ctx);
/* device name, CU count, HBM bytes; any pointer may be NULL */
int         sfem_device_info(sfem_ctx* ctx, char* name, size_t name_len, int* n_cu,
                             int64_t* hbm_bytes);
/* bytes the context's allocator currently holds on the device */
int64_t     sfem_ctx_device_bytes(const sfem_ctx* ctx);

/* ---- scalar utilities (host) ----------------------------------------- */
/* gaussian_quadrature(a,b,n), src/fem.jl:21-31.  n in {2,3}, else
 * SFEM_ERR_UNDEFVAR (the reference leaves xi undefined).  xi, w: n doubles. */
int sfem_gaussian_quadrature(double a, double b, int n, double* xi, double* w);

/* basis_function(xi,eta,zeta,FunctionClass), src/fem.jl:48-114.
 * nargs = number of coordinates given (1: xi, 2: xi,eta, 3: xi,eta,zeta).
 * N gets nfun values, dN gets nfun x ncol values column-major (Julia layout),
 * where ncol = the reference's ΔN column count (Q1: 1D -> the 1x2 row, so
 * nfun=2 and dN = [-0.5, 0.5] with *ncol = 2 and *nrow = 1).
 * Q2 needs nargs >= 2 and ignores zeta, as the reference does. */
int sfem_basis_function(int nargs, double xi, double eta, double zeta, int function_class,
                        double* N, double* dN, int* nfun, int* nrow, int* ncol);

/* ---- assembly: assemble_system, src/fem.jl:135-256 -------------------- */
/* Upload mesh inputs once (H2D); assembly then runs from HBM.
 *   n_elem      = ne^ndim in the reference (element count actually used)
 *   young       = n_young values: 1 (the reference's scalar Young) or n_elem
 *                 (per-element modulus; Ke is linear in E)
 *   id          = may be NULL when n_dof == 1 (reference ignores ID then)  */
int sfem_mesh_create(sfem_ctx* ctx, int ndim, int function_class, int n_dof,
                     const double* node_list, int64_t n_nodes,
                     const int64_t* ien, int64_t ien_rows, int nen, int64_t n_elem,
                     const int64_t* id, int64_t id_rows,
                     const double* young, int64_t n_young, double nu,
                     sfem_mesh** out);
void sfem_mesh_destroy(sfem_mesh* mesh);
/* Declare a Q4 / Hex8 vector mesh (n_dof = ndim) the structured grid of the
 * reference's meshgrid (examples/vector3D.jl:10-130): ne[d] elements along
 * direction d, node i + (nx+1)(j + (ny+1)k) (0-based) at grid point (i,j,k),
 * elements numbered x fastest with one corner order for all, ID the
 * node-interleaved numbering ndof*(node-1)+l.  Checked on the device
 * (SFEM_ERR_ARGUMENT if the mesh is not that grid).  sfem_assemble_mesh then
 * takes every node pair's elements from the grid instead of sorting keys;
 * the result is bitwise the same.  sfem_meshgrid meshes are declared already. */
int  sfem_mesh_set_grid(sfem_ctx* ctx, sfem_mesh* mesh, int ndim, const int64_t* ne);
/* The same declaration without the element counts: they are read from element
 * 1's corners (nodes 0, 1, nx+1, nx+2, ... of the meshgrid numbering), then
 * the mesh is checked as by sfem_mesh_set_grid.  *ndim = 0 (and SFEM_OK) when
 * the mesh is not such a grid; the mesh is then left as it was.  ndim / ne
 * (3 values, 0 beyond ndim) may be NULL.  No counterpart in the reference:
 * sfem_assemble calls it, so assemble_system's own signature (src/fem.jl:135)
 * reaches the grid-pattern assembly and the matrix-free multigrid levels. */
int  sfem_mesh_detect_grid(sfem_ctx* ctx, sfem_mesh* mesh, int* ndim, int64_t* ne);

/* flags for sfem_assemble_mesh */
#define SFEM_ASM_DEFAULT       0
#define SFEM_ASM_FORCE_SCALAR  1   /* sort dof-pair keys even when ID is node-blocked */
#define SFEM_ASM_EXACT_SUMS    2   /* on a declared grid, sum the element matrices themselves */

/* Assemble K = sparse(E,J,V) on the device.  Pattern is the union of the
 * element couplings (explicit zeros kept), duplicates are summed in
 * ascending COO index order, dims = max row/col id -- the semantics of
 * Julia's sparse(I,J,V) called at src/fem.jl:253.  On a declared grid whose
 * element matrices are all s_e Ke(0) to within 1e-13 of Ke(0)'s largest
 * entry (s_e = Ke(e)[1,1] / Ke(0)[1,1], every element matrix computed and
 * compared), K's values are summed from s_e and Ke(0) in the same order
 * instead of from the stored element matrices: the same pattern, values
 * within that deviation, K bitwise symmetric (sfem_csr_grid_values says
 * which form was taken; SFEM_ASM_EXACT_SUMS keeps the element sums). */
int sfem_assemble_mesh(sfem_ctx* ctx, const sfem_mesh* mesh, int flags, sfem_csr** out);

/* One-shot drop-in for assemble_system: sfem_mesh_create +
 * sfem_mesh_detect_grid + sfem_assemble_mesh with SFEM_ASM_DEFAULT (the
 * same pattern either way; the values as sfem_assemble_mesh says). */
int sfem_assemble(sfem_ctx* ctx, int ndim, int function_class, int n_dof,
                  const double* node_list, int64_t n_nodes,
                  const int64_t* ien, int64_t ien_rows, int nen, int64_t n_elem,
                  const int64_t* id, int64_t id_rows,
                  const double* young, int64_t n_young, double nu,
                  sfem_csr** out);

/* m, n, stored entries, and the block size the device copy is kept in
 * (n_dof when ID is node-blocked, else 1).  Any pointer may be NULL. */
int  sfem_csr_info(const sfem_csr* K, int64_t* m, int64_t* n, int64_t* nnz, int* block);
/* The grid K was assembled on (sfem_mesh_set_grid / sfem_mesh_detect_grid):
 * *ndim = 0 when none; ne = 3 element counts (0 beyond ndim).  No counterpart
 * in the reference (its K carries no mesh). */
int  sfem_csr_grid(const sfem_csr* K, int* ndim, int64_t* ne);
/* How K's values were formed on its grid: *from_scales = 1 when summed from
 * s_e Ke(0), 0 from the element matrices; *ke_dev = max over elements of
 * |Ke(e) / s_e - Ke(0)| / max |Ke(0)| (+inf when not measured).  No
 * counterpart in the reference.  Either pointer may be NULL. */
int  sfem_csr_grid_values(const sfem_csr* K, int* from_scales, double* ke_dev);

/* Fill caller-allocated host buffers with the 1-based CSC arrays of
 * SparseMatrixCSC{Float64,Int64}: colptr (n+1), rowval (nnz), nzval (nnz). */
int  sfem_export_csc1(sfem_ctx* ctx, const sfem_csr* K, int64_t* colptr, int64_t* rowval,
                      double* nzval);
void sfem_csr_destroy(sfem_csr* K);
/* Upload a host SparseMatrixCSC (1-based, e.g. K_bar = K + beta*b built on
 * the host as examples/vector3D.jl:308 does) for sfem_bc_create/eliminate/
 * solve.  Stored as scalar rows (block 1); export is refused for it. */
int  sfem_csr_from_csc1(sfem_ctx* ctx, int64_t m, int64_t n, const int64_t* colptr,
                        const int64_t* rowval, const double* nzval, sfem_csr** out);
/* Device time of the last assembly's stages (ms): validate, element Ke,
 * keys, sort, segmented reduction, total.  ms: 6 doubles. */
int  sfem_csr_stage_ms(const sfem_csr* K, double* ms);

/* ---- Dirichlet elimination, examples/vector3D.jl:133-173, 314-322 ----- */
/* n = size(K,1).  f: load vector (NULL = 0).  fixed: n bytes, 1 = prescribed
 * dof (the complement of the columns C = I[:,free] keeps).  q_d: prescribed
 * values on fixed dofs (NULL = 0; entries on free dofs are ignored). */
int  sfem_bc_create(sfem_ctx* ctx, int64_t n, const double* f, const uint8_t* fixed,
                    const double* q_d, sfem_bc** out);
void sfem_bc_destroy(sfem_bc* bc);

/* Build K_ff = C' K C and rhs = C'(f - K q_d) on the device.  The solver
 * formats are built here: the FP64 CSR row pointers (its columns and values
 * on the first solve that uses the CSR format), and, when K is node-blocked,
 * the node-block formats (a node with some dofs fixed keeps them as a zero
 * row / column with a unit diagonal, so they stay exactly 0 in the solve).
 * The system holds a reference on K (its solves and reactions read K), so the
 * caller may call sfem_csr_destroy(K) at any time; K's memory is released
 * when the last system eliminated from it is destroyed. */
int  sfem_eliminate(sfem_ctx* ctx, const sfem_csr* K, const sfem_bc* bc, sfem_system** out);
int  sfem_system_info(const sfem_system* sys, int64_t* n_free, int64_t* nnz_ff, int* block);
/* Device time of the elimination that built the system (ms). */
int  sfem_system_eliminate_ms(const sfem_system* sys, double* ms);
void sfem_system_destroy(sfem_system* sys);

/* ---- solve ------------------------------------------------------------ */
#define SFEM_PREC_FP64   0   /* plain FP64 Jacobi-PCG on FP64 CSR                       */
#define SFEM_PREC_MIXED  1   /* FP32 operator and CG vectors, FP64 residual replacement  */
#define SFEM_PREC_FP32_MATRIX 2 /* FP32 operator values, FP64 CG vectors and residuals    */
#define SFEM_PREC_SPLIT  3   /* operator as FP32 + 16-bit correction (6 B/value, 2^-39
                                relative), FP64 CG vectors; node-block format only       */

#define SFEM_FMT_AUTO    0
#define SFEM_FMT_CSR     1   /* scalar CSR, int32 columns                               */
#define SFEM_FMT_BSR     2   /* node-block (n_dof x n_dof) sliced format, int32 columns */

typedef struct {
    int     precision;       /* SFEM_PREC_*                                              */
    int     format;          /* SFEM_FMT_*                                               */
    double  rtol;            /* stop when ||rhs - K_ff x||_2 <= rtol ||rhs||_2 (FP64)   */
    int64_t max_iter;        /* cap on CG iterations (all precisions)                    */
    double  reliable_delta;  /* mixed: FP64 update when the FP32 residual drops by this  */
    int     check_every;     /* iterations enqueued between host convergence polls      */
    int     time_kernels;    /* 1: time every SpMV launch with HIP events                */
    int     x0_zero;         /* 1: start from x = 0 (else from the system's current x)   */
    int     preconditioner;  /* SFEM_PC_*                                                */
} sfem_solve_opts;

#define SFEM_PC_JACOBI        0   /* z = r / diag(K_ff)                                   */
#define SFEM_PC_BLOCK_JACOBI  1   /* z = (node diagonal block)^-1 r; node-block format only */
#define SFEM_PC_MULTIGRID     2   /* z = one geometric multigrid V-cycle (sfem_system_set_grid
                                     first); precision FP64, SPLIT or FP32_MATRIX          */

typedef struct {
    int64_t iterations;      /* CG iterations performed                                  */
    int64_t fp64_updates;    /* mixed: FP64 residual replacements performed              */
    int64_t spmv_launches;   /* operator applications in the iteration (any precision)   */
    int64_t spmv64_launches; /* FP64 residual SpMVs                                      */
    double  rel_residual;    /* final ||rhs - K_ff x|| / ||rhs|| evaluated in FP64       */
    double  solve_ms;        /* device time of the solve (HIP events)                    */
    double  spmv_ms;         /* summed device time of timed SpMV launches                */
    double  spmv_bytes;      /* algorithmic bytes of one iteration SpMV                  */
    int64_t spmv_timed;      /* number of SpMV launches that were timed                  */
    int     converged;
    int     format;          /* format actually used                                     */
    int     stop_reason;     /* 1 converged, 2 breakdown, 3 stagnation, 4 iteration cap  */
    double  floor_residual;  /* attainable relative residual estimate at the last FP64
                                residual: eps || |rhs| + |K_ff||x| || / ||rhs||          */
    double  smooth_ms;       /* multigrid, time_kernels: summed device time of the timed
                                finest-level smoothing steps (FP32 operator)             */
    int64_t smooth_timed;    /* number of those launches                                 */
    double  smooth_bytes;    /* algorithmic bytes of one finest-level smoothing step     */
} sfem_solve_stats;

void sfem_solve_opts_default(sfem_solve_opts* opts);

int  sfem_system_solve(sfem_ctx* ctx, sfem_system* sys, const sfem_solve_opts* opts,
                       sfem_solve_stats* stats);
/* The same solve, allocating nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The CG vectors,
 * partial sums and state live in one file-scope device array per GPU of
 * 40.5 MiB (held until the process ends), so the system's solver vectors may
 * have at most 2^20 = 1,048,576 entries (free dofs; with partially fixed
 * nodes, node rows x n_dof), else SFEM_ERR_UNSUPPORTED.  Point or block
 * Jacobi, any precision; multigrid is refused (its V-cycle allocates).  The
 * operator format the options read must exist: the first sfem_system_solve
 * with the same precision and format builds it (else SFEM_ERR_STATE).  No
 * kernel timing (stats->spmv_ms is 0; solve_ms is host wall time).  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (from different contexts or threads) race and must not be made. */
int  sfem_system_solve_noalloc(sfem_ctx* ctx, sfem_system* sys, const sfem_solve_opts* opts,
                               sfem_solve_stats* stats);
/* q = q_d + C q_f (n values, FP64), and optionally reactions K q - f on the
 * fixed dofs (n values; zero on free dofs). */
int  sfem_system_solution(sfem_ctx* ctx, const sfem_system* sys, double* q,
                          double* reactions);

/* ---- geometric multigrid preconditioner --------------------------------
 * No counterpart in the reference (its solve is a dense inverse,
 * examples/vector3D.jl:316-320); the solution is the same one, to rtol on the
 * FP64 true residual.  For node-block systems (n_dof = ndim) whose nodes form
 * the structured grid of the reference's meshgrid numbering
 * (examples/vector3D.jl:67-101: node = i + (nx+1)(j + (ny+1) k), x fastest),
 * with dims[d] elements along direction d.  Any geometry, material field and
 * fixed set; K may hold only couplings between grid neighbours (element
 * stiffness, surface penalty), else SFEM_ERR_ARGUMENT.
 *   Levels: the grid coarsened by 2 along each direction whose element count
 *   is even, until a level has at most coarse_nodes nodes (at least once);
 *   coarse operators are Galerkin products P'AP with multilinear P on the
 *   free dofs; the coarsest (at most 400 nodes, else SFEM_ERR_UNSUPPORTED)
 *   is solved with a dense inverse.  Smoothing: `degree` Chebyshev steps in
 *   (node diagonal block)^-1 K on [lo, hi] x lambda_max before and after the
 *   coarse correction, lambda_max by power_iters power iterations.
 * The hierarchy depends on K and the fixed set only: sfem_system_set_bc keeps
 * it.  Solve with opts.preconditioner = SFEM_PC_MULTIGRID. */
#define SFEM_MG_MAX_LEVELS 24
typedef struct {
    int     degree;        /* Chebyshev steps before and after the coarse correction (2)  */
    int     coarse_degree; /* the same on levels below the finest; 0: degree (0)          */
    double  lo, hi;        /* smoothing interval, x lambda_max (0.1, 1.1)                 */
    int     power_iters;   /* power iterations per level for lambda_max (8)               */
    int64_t coarse_nodes;  /* stop coarsening at this many nodes (64)                     */
    int     fine_fp16;     /* 1: the finest-level smoother reads FP16 values of K_ff / 2^e
                              (max |a| / 2^e in (1/2, 1]) instead of FP32 (1)             */
    int     coarse_fp16;   /* 1: the same for the levels between finest and coarsest (1)  */
    int     matrix_free;   /* finest level applied from the element scales when K was
                              assembled on this grid (sfem_mesh_set_grid) and every element
                              matrix is s_e Ke(0): 1 the smoother (deviation <= 1e-6),
                              2 also the FP64 outer operator (<= 1e-12); 0 off (2)        */
} sfem_mg_opts;

typedef struct {
    int     levels;
    int     degree;
    int     fine_fp16;
    int     coarse_fp16;                     /* FP16 smoothers below the finest level   */
    int64_t nodes[SFEM_MG_MAX_LEVELS];
    int64_t dims[SFEM_MG_MAX_LEVELS][3];     /* elements per direction (0 beyond ndim)  */
    double  lambda_max[SFEM_MG_MAX_LEVELS];  /* estimate; 0 on the coarsest level         */
    int64_t coarse_dofs;
    double  setup_ms;                        /* host wall time of the build             */
    int     matrix_free;                     /* 0, 1, 2 as sfem_mg_opts.matrix_free     */
    int     pad2;
    double  ke_deviation;                    /* max |Ke(e)/s_e - Ke(0)| / max |Ke(0)|;
                                                +inf when K was not assembled on the grid */
} sfem_mg_info;

void sfem_mg_opts_default(sfem_mg_opts* opts);
/* Builds the hierarchy (opts NULL: defaults).  ndim = 0 with dims = NULL: the
 * grid K was assembled on (SFEM_ERR_ARGUMENT if it was not). */
int  sfem_system_set_grid(sfem_ctx* ctx, sfem_system* sys, int ndim, const int64_t* dims,
                          const sfem_mg_opts* opts);
int  sfem_system_mg_info(const sfem_system* sys, sfem_mg_info* info);
/* Level >= 1 operator as stored for the next Galerkin product (FP64, fixed
 * rows and columns zero): stencil[(s*B2 + a*n_dof + b)*nodes + I] =
 * A[(I,a),(I+off_s,b)], B2 = n_dof^2, off_s = (s%3-1, s/3%3-1, s/9-1) (3D) or
 * (s%3-1, s/3-1) (2D); fixed: nodes*n_dof flags.  Either may be NULL. */
int  sfem_system_mg_stencil(sfem_ctx* ctx, const sfem_system* sys, int level, double* stencil,
                            uint8_t* fixed);

/* One-shot: bc_create + eliminate + solve + solution. */
int  sfem_solve(sfem_ctx* ctx, const sfem_csr* K, const double* f, const uint8_t* fixed,
                const double* q_d, const sfem_solve_opts* opts, double* q, double* reactions,
                sfem_solve_stats* stats);

/* ---- the example's boundary machinery and displacement sweep -----------
 * examples/vector3D.jl:10-338 and src/PostProcess.jl:30-46.  Meshes, face
 * lists, penalty matrices and constraints stay on the device. */

/* apply_boundary_conditions(ne, NodeList, IEN, IEN_top, IEN_btm, ndim,
 * FunctionClass, ID, nDof) (examples/vector3D.jl:175-264): the surface
 * penalty b = sparse(E,J,V) with, per face and 2x2 Gauss point,
 * V += |t1 x t2| * N_a N_b on equal dof components (bottom faces first in
 * COO order, then top).  IEN_btm / IEN_top: n_faces x nfn, column-major,
 * 1-based; ID: id_rows x n_dof, column-major.  Only ndim = 3, nfn = 4,
 * n_dof = 3 runs in the reference; anything else is SFEM_ERR_DIMENSION.
 * Errors follow the reference's face loop (BoundsError at the first face
 * whose nodes or ids are out of range). */
int  sfem_surface_mesh_create(sfem_ctx* ctx, int ndim, int n_dof, const double* node_list, int64_t n_nodes,
                              const int64_t* ien_btm, const int64_t* ien_top, int64_t n_faces, int nfn,
                              const int64_t* id, int64_t id_rows, sfem_mesh** out);
int  sfem_assemble_surface(sfem_ctx* ctx, int ndim, int n_dof, const double* node_list, int64_t n_nodes,
                           const int64_t* ien_btm, const int64_t* ien_top, int64_t n_faces, int nfn,
                           const int64_t* id, int64_t id_rows, sfem_csr** out);

/* K_bar = K + beta*b (examples/vector3D.jl:308): union pattern, values
 * K_ij + fl(beta*b_ij) (K_ij + 0.0 where b has no entry, 0.0 + fl(beta*b_ij)
 * where K has none).  SparseArrays drops entries whose sum is exactly zero
 * from the stored pattern; this block format keeps them as stored zeros.
 * Different sizes: SFEM_ERR_DIMENSION (as Julia's +). */
int  sfem_csr_add(sfem_ctx* ctx, const sfem_csr* K, double beta, const sfem_csr* b, sfem_csr** out);

/* y = K x (n values in and out, FP64) for an assembled matrix (K, b,
 * K + beta*b: all bitwise symmetric) or an uploaded CSC.  With b the surface
 * penalty and x = the traction t repeated per node, y is the consistent load
 * vector of a uniform traction on those faces, int N_a t dA (no counterpart
 * in the reference, which passes loads as arrays). */
int  sfem_csr_apply(sfem_ctx* ctx, const sfem_csr* K, const double* x, double* y);

/* Structured mesh on the device: meshgrid(x0,x1,y0,y1,z0,z1,ne,ndim)
 * (examples/vector3D.jl:10-130) generalised to ne[0] x ne[1] (x ne[2])
 * elements; node numbering x fastest, Q4/Hex8 IEN in the reference's
 * order, ID[m,l] = ndim(m-1)+l, coordinates x0 + i*(x1-x0)/ne with the end
 * point exact.  box = {x0,x1,y0,y1[,z0,z1]}.  SFEM_MESH_INFLATE maps (x, y)
 * as inflate_sphere(NodeList, x0, x1, y0, y1) (src/PostProcess.jl:30-46).
 * faces (3D, may be NULL): the bottom (z = z0) and top (z = z1) face lists
 * IEN_btm / IEN_top (:96-115) as a surface mesh for sfem_assemble_mesh. */
#define SFEM_MESH_INFLATE 1
int  sfem_meshgrid(sfem_ctx* ctx, int ndim, const int64_t* ne, const double* box, int flags,
                   const double* young, int64_t n_young, double nu, sfem_mesh** mesh, sfem_mesh** faces);
/* Sizes of a mesh, and its arrays in Julia's layouts (NodeList ndim x nNodes,
 * IEN nElem x nen, ID nNodes x nDof; column-major, 1-based).  Any output
 * pointer may be NULL. */
int  sfem_mesh_info(const sfem_mesh* mesh, int* ndim, int* nen, int* n_dof, int64_t* n_nodes, int64_t* n_elem);
int  sfem_mesh_export(sfem_ctx* ctx, const sfem_mesh* mesh, double* node_list, int64_t* ien, int64_t* id);

/* Dirichlet constraints from coordinate planes, the rule of setboundaryCond
 * (examples/vector3D.jl:133-173) generalised: rule r fixes dof dof[r]
 * (1-based) of every node whose coordinate axis[r] (1-based) equals
 * coord[r] exactly, with value value[r]; for each (node, dof) the first
 * matching rule wins (the reference's if/elseif).  f: n_total loads or NULL. */
int  sfem_bc_planes(sfem_ctx* ctx, const sfem_mesh* mesh, int64_t n_total, int nrules, const int* axis,
                    const double* coord, const int* dof, const double* value, const double* f, sfem_bc** out);

/* New loads and prescribed values for an eliminated system whose fixed set
 * is unchanged (the displacement sweep, examples/vector3D.jl:310-322):
 * rhs = C'(f - K q_d) summed as sfem_eliminate does (bitwise the same rhs);
 * the solution is kept as the next solve's start (opts.x0_zero = 0).  A
 * different fixed set is SFEM_ERR_ARGUMENT. */
int  sfem_system_set_bc(sfem_ctx* ctx, sfem_system* sys, const sfem_bc* bc);

/* ---- problems generated on the device from a compact spec -------------
 * The load and constraint front end of the example (meshgrid,
 * examples/vector3D.jl:10-130; setboundaryCond, :133-173) moved to the
 * device, plus consistent loads, which the reference passes as arrays.  A
 * structured Q4 (ndim 2) / Hex8 (ndim 3) elasticity problem:
 *   mesh     meshgrid of box with dims elements per direction (the numbering,
 *            corner order and coordinates of sfem_meshgrid)
 *   material E per element log-uniform in [young_lo, young_hi] from seed
 *            (E = young_lo when equal), nu for every element
 *   clamp    every dof of the nodes on the face x_clamp_axis = box min
 *            (clamp_side 0) or max (1)
 *   loads    consistent nodal loads of a uniform traction (force per area)
 *            on the face load_axis / load_side and of a body force (per
 *            volume), 2-point Gauss per cell / face, each node summing its
 *            cells in ascending order; plus N(0, noise^2) on every free dof
 *            of the loaded face.
 * Random draws are counter-based on (seed, stream, index):
 *   z = splitmix64 finaliser (x ^ x>>30, * 0xBF58476D1CE4E5B9, ^ >>27,
 *       * 0x94D049BB133111EB, ^ >>31) of (seed ^ stream*0xD1B54A32D192ED03
 *       ^ index*0x9E3779B97F4A7C15) + 0x9E3779B97F4A7C15 (mod 2^64),
 *   u = ((z >> 12) + 1/2) 2^-52;
 *   E_e = exp(log lo + u(1, e) (log hi - log lo));
 *   noise on local dof k = noise sqrt(-2 log u(2, 2k)) cos(2 pi u(2, 2k+1)). */
typedef struct {
    int      ndim;
    int      pad0;
    int64_t  dims[3];        /* elements per direction (dims[2] unused in 2D)            */
    double   box[6];         /* x0, x1, y0, y1[, z0, z1]                                 */
    double   young_lo, young_hi;
    double   nu;
    uint64_t seed;
    int      clamp_axis, clamp_side;
    int      load_axis, load_side;
    double   traction[3];    /* on the loaded face, force per unit area                  */
    double   body[3];        /* force per unit volume                                    */
    double   noise;          /* sigma of the random nodal loads on the loaded face       */
} sfem_problem_spec;

/* The problem's mesh (declared its grid) and loads / constraints on the
 * device; then sfem_assemble_mesh + sfem_eliminate as for any mesh. */
int  sfem_spec_build(sfem_ctx* ctx, const sfem_problem_spec* spec, sfem_mesh** mesh, sfem_bc** bc);
/* The same inputs copied to the host (for checking or output): NodeList
 * (ndim x n_nodes), Young per element, f and the fixed mask (ndim*n_nodes).
 * Any pointer may be NULL. */
int  sfem_spec_export(sfem_ctx* ctx, const sfem_problem_spec* spec, double* node_list, double* young, double* f,
                      uint8_t* fixed);


/* ---- batches: many meshes, one persistent solver launch ---------------
 * No counterpart in the reference (each assemble_system / backslash call
 * there handles one mesh, src/fem.jl:135-256); every problem is the same
 * maths as a single-mesh call.  Problems are Q4 (ndim 2) or Hex8 (ndim 3)
 * vector elasticity with node-interleaved dofs (ID[m,l] = ndim*(m-1)+l, as
 * examples/vector3D.jl:33 builds it).  A node may have some dofs fixed and
 * others free (a roller plane): its fixed dofs are held at q_d exactly.
 *   node_list    ndim x N_total, node-major, problems concatenated
 *   node_offsets nprob+1 (node_offsets[p] = first node of problem p, 0-based)
 *   ien          E_total x nen column-major, 1-based node ids local to the
 *                element's problem
 *   elem_offsets nprob+1
 *   young, nu    one value per element
 *   f, fixed, q_d  ndim*N_total values in the interleaved numbering */
typedef struct sfem_batch sfem_batch;

typedef struct {
    int64_t iterations;
    int64_t fp64_updates;
    double  rel_residual;     /* final FP64 ||rhs - K_ff x|| / ||rhs||                   */
    int64_t n_free;
    int     stop_reason;      /* 1 converged, 2 breakdown, 3 stagnation, 4 iteration cap */
    double  floor_residual;   /* attainable relative residual estimate (see sfem_solve_stats) */
} sfem_batch_problem_stats;

/* k right-hand sides of one eliminated system at once (the sweep's
 * K_free q_f = C'(f - K_bar q_d) for k displacements, examples/vector3D.jl:
 * 310-322): column j uses q_d[:, j] and f[:, j] (n x k column-major; NULL =
 * zeros, prescribed values read on the fixed dofs only) and returns
 * q[:, j] = q_d + C q_f.  One FP64 point-Jacobi PCG per column, all columns
 * sharing each pass over K_ff; each column stops when its FP64 true residual
 * meets opts->rtol (restarting from it otherwise).  1 <= k <= 64; per: k
 * entries (fp64_updates = restarts), may be NULL; solve_ms may be NULL.
 * SFEM_ERR_NOT_CONVERGED if a column misses rtol. */
int  sfem_system_solve_many(sfem_ctx* ctx, sfem_system* sys, int k, const double* q_d, const double* f,
                            const sfem_solve_opts* opts, double* q, sfem_batch_problem_stats* per,
                            double* solve_ms);

typedef struct {
    double  solve_ms;         /* device time of all launches (HIP events)                 */
    int64_t launches;
    int64_t max_iterations;
    int64_t sum_iterations;
    int64_t not_converged;
    int     grid;             /* blocks of the persistent launch                         */
    int64_t device_bytes;     /* operator + vectors held by the batch                     */
    int     barrier_grouped;  /* 1: the launches used the per-XCD grouped grid barrier    */
} sfem_batch_stats;

int  sfem_batch_create(sfem_ctx* ctx, int64_t nprob, int ndim, int n_dof, const double* node_list,
                       const int64_t* node_offsets, const int64_t* ien, int nen,
                       const int64_t* elem_offsets, const double* young, const double* nu,
                       const double* f, const uint8_t* fixed, const double* q_d, sfem_batch** out);
/* A batch of problems generated on the device from specs (one dimension for
 * all): no host-side mesh arrays are built or copied. */
int  sfem_batch_create_specs(sfem_ctx* ctx, int64_t nprob, const sfem_problem_spec* specs, sfem_batch** out);
/* n_total = ndim*N_total; n_padded = dofs of the padded operator; slices of 64 node rows. */
int  sfem_batch_info(const sfem_batch* b, int64_t* n_total, int64_t* n_padded, int64_t* slices,
                     double* assemble_ms, double* eliminate_ms);
/* Solve every problem.  item_slices: slices per work item (changes the
 * reduction order, so results are bitwise reproducible per item_slices);
 * blocks_per_cu: persistent grid = blocks_per_cu * CUs (0 = default; capped
 * at residency).  opts->check_every > 0 splits the solve into launches of
 * that many iterations; <= 0 runs it in one launch.  per: nprob entries or
 * NULL.  Returns SFEM_ERR_NOT_CONVERGED if any problem did not converge. */
int  sfem_batch_solve(sfem_ctx* ctx, sfem_batch* b, const sfem_solve_opts* opts, int item_slices,
                      int blocks_per_cu, sfem_batch_problem_stats* per, sfem_batch_stats* stats);
/* q = q_d + C x for all problems (n_total values, FP64). */
int  sfem_batch_solution(sfem_ctx* ctx, const sfem_batch* b, double* q);
/* Largest resident blocks per CU of the persistent kernel for a precision. */
int  sfem_batch_max_blocks_per_cu(sfem_ctx* ctx, int precision, int* blocks);
void sfem_batch_destroy(sfem_batch* b);

/* ---- serving: continuous batching over the persistent solver -----------
 * No counterpart in the reference.  Requests (one Q4 or Hex8 problem each,
 * inputs as a one-problem batch) queue on the host; each step admits queued
 * requests in arrival order while their device footprint fits the HBM
 * budget and a slot is free (assembly + elimination on the device), runs
 * iters_per_step iterations of every active request in one persistent
 * launch, and retires the finished ones (solution copied to the host, device
 * memory released).  A request's result does not depend on what else is
 * active: it is bitwise the result it gets alone.
 * With preconditioner = SFEM_PC_MULTIGRID, requests submitted with their
 * grid (sfem_server_submit_grid, sfem_server_submit_spec) are solved
 * run-to-completion instead, up to mg_workers at once: each step hands the
 * oldest queued such requests that fit the HBM budget to idle worker
 * threads, each of which assembles, eliminates, builds the multigrid
 * hierarchy and solves (FP64 CG, sfem_system_set_grid) on its own context
 * and stream, so their kernels overlap on the device beside the persistent
 * launch for the other active requests.  Steps with nothing else to do wait
 * up to 1 ms for a worker to finish.  A grid that cannot be coarsened, and
 * requests without a grid, go to the persistent launch with node-block
 * Jacobi. */
typedef struct sfem_server sfem_server;

typedef struct {
    int     precision;        /* SFEM_PREC_*                                              */
    double  rtol;
    int64_t max_iter;
    double  reliable_delta;
    int     item_slices;
    int     blocks_per_cu;
    int64_t hbm_budget;       /* bytes admitted requests may hold on the device           */
    int     max_active;       /* cap on concurrently active requests                      */
    int     iters_per_step;   /* CG iterations per persistent launch                      */
    int     preconditioner;   /* SFEM_PC_*                                                */
    int     mg_workers;       /* multigrid servers: requests solved at once, each by a host
                                 thread with its own stream (4)                            */
} sfem_server_opts;

#define SFEM_REQ_QUEUED  0
#define SFEM_REQ_ACTIVE  1
#define SFEM_REQ_DONE    2
#define SFEM_REQ_FAILED  3

void sfem_server_opts_default(sfem_server_opts* opts);
int  sfem_server_create(sfem_ctx* ctx, const sfem_server_opts* opts, sfem_server** out);
/* Queue one problem; copies the inputs.  young/nu: one value per element.
 * f / q_d may be NULL (zero).  *id receives the request id (0, 1, ...). */
int  sfem_server_submit(sfem_server* s, int ndim, const double* node_list, int64_t n_nodes,
                        const int64_t* ien, int64_t n_elem, const double* young, const double* nu,
                        const double* f, const uint8_t* fixed, const double* q_d, int64_t* id);
/* sfem_server_submit for a mesh whose nodes are the structured grid of the
 * meshgrid numbering (examples/vector3D.jl:67-101) with dims[d] elements
 * along direction d. */
int  sfem_server_submit_grid(sfem_server* s, int ndim, const int64_t* dims, const double* node_list,
                             int64_t n_nodes, const int64_t* ien, int64_t n_elem, const double* young,
                             const double* nu, const double* f, const uint8_t* fixed, const double* q_d,
                             int64_t* id);
/* A request generated on the device from a spec (sfem_problem_spec): no
 * host arrays are built or copied; the spec is its whole input.  Solved
 * like a grid request. */
int  sfem_server_submit_spec(sfem_server* s, const sfem_problem_spec* spec, int64_t* id);
/* One admit / iterate / retire step.  Outputs (may be NULL): requests that
 * finished in this step, and device ms of the step's solver work. */
int  sfem_server_step(sfem_server* s, int* n_finished, double* launch_ms);
/* State of a request (SFEM_REQ_*), its stats once done, and host times in
 * ms since the server was created: [submitted, admitted, finished]. */
int  sfem_server_poll(const sfem_server* s, int64_t id, int* state, sfem_batch_problem_stats* st,
                      double* times3);
/* Copy a finished request's solution (ndim*n_nodes values) and forget it. */
int  sfem_server_take(sfem_server* s, int64_t id, double* q);
/* queued / active request counts and device bytes held by active requests. */
int  sfem_server_info(const sfem_server* s, int64_t* queued, int64_t* active, int64_t* bytes);
void sfem_server_destroy(sfem_server* s);

/* ---- measurement helpers --------------------------------------------- */
/* y = K_ff x in the given precision/format, `reps` times, on the system's
 * own vectors; returns the average device time per launch (HIP events). */
int  sfem_system_spmv_bench(sfem_ctx* ctx, sfem_system* sys, int precision, int format,
                            int reps, double* avg_ms, double* bytes);
/* Stream-copy `bytes` bytes device-to-device with the library's copy kernel
 * `reps` times; returns average ms per launch (streaming HBM ceiling). */
int  sfem_stream_copy_bench(sfem_ctx* ctx, int64_t bytes, int reps, double* avg_ms);
/* Wait for all work on the context's stream. */
int  sfem_ctx_synchronize(sfem_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SMEARFEM_H */
