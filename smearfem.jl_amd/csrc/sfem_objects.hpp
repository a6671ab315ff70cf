// sfem_objects.hpp -- device-resident objects behind the opaque C-ABI handles,
// and the host launchers each translation unit provides.
#pragma once

#include <memory>

#include "sfem_internal.hpp"

// Mesh and material inputs, uploaded once.
struct sfem_mesh {
    sfem_ctx* ctx = nullptr;
    int ndim = 0, cls = 0, ndof = 0, nen = 0;
    int64_t n_nodes = 0, n_elem = 0, ien_rows = 0, id_rows = 0, n_young = 0;
    double nu = 0.3;
    sfem::DevBuf<double> coords;   // ndim x n_nodes, node-major (x[d + ndim*node])
    sfem::DevBuf<int32_t> ien;     // n_elem x nen element-major, 0-based; -1 = outside NodeList
    sfem::DevBuf<int32_t> dof;     // id_rows x ndof node-major, 0-based; -1 = id < 1, -2 = id > 2^31
    sfem::DevBuf<double> young;    // n_young
    sfem::DevBuf<double> nu_e;     // per-element Poisson ratio (batches); empty = scalar nu
    // Error the reference raises before any element work can be checked on the
    // device (structural mismatches found at element 1, or IEN too short):
    // (code, key) with key = element*8 + step, or code 0.
    int pre_code = 0;
    int64_t pre_key = INT64_MAX;
    // Surface mesh of the penalty b: elements are the bottom faces then the
    // top faces (nen = 4 in 3D); errors are reported per face index, i.e.
    // element % key_period.
    bool face = false;
    int64_t key_period = 0;
    // Structured grid (sfem_mesh_set_grid): grid_ne elements per direction,
    // nodes and elements in the meshgrid numbering (x fastest), the local
    // index of each cell corner (code ox + 2 oy + 4 oz), identity node ids.
    // Assembly then takes every node pair's elements from the grid instead
    // of sorting keys.
    int grid_nd = 0;
    int64_t grid_ne[3] = {0, 0, 0};
    int grid_corner[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Assembled K, kept as a block CSC: column-unit c holds blocks (row-unit r,
// bs x bs values with val[i*bs + j] = K[r*bs+i, c*bs+j]).  bs = ndof when ID is
// node-blocked (ID[m,l] = ndof*g(m) + l), else 1.
struct sfem_csr {
    sfem_ctx* ctx = nullptr;
    // Handle references: the caller's (sfem_csr_destroy drops it) plus one per
    // system eliminated from this matrix (sfem_system_destroy drops it), so a
    // caller may destroy K right after sfem_eliminate.
    mutable std::atomic<int> refs{1};
    int64_t m = 0, n = 0, nnz = 0;
    int bs = 1;
    bool transposed = false;  // stores K' (uploaded from a host CSC)
    int64_t nunits = 0, nnzb = 0;
    sfem::DevBuf<int64_t> bcolptr;  // nunits + 1
    sfem::DevBuf<int32_t> browidx;  // nnzb
    sfem::DevBuf<int32_t> bcolidx;  // nnzb (column unit of each block)
    sfem::DevBuf<double> bval;      // nnzb * bs * bs
    // Assembled on a structured grid (sfem_mesh_set_grid): its element counts
    // and corner order, and how far the element matrices are from multiples
    // of the first one, Ke(e) = s_e Ke(0): ke_dev = max over elements and
    // entries of |Ke(e) / s_e - Ke(0)| / max |Ke(0)| (+inf when not measured).
    // On a uniform grid with one Poisson ratio it is rounding (the elements'
    // coordinates differ in their last bits); multigrid then applies the
    // finest-level operator from s_e and Ke(0) without reading K.
    int grid_nd = 0;
    int64_t grid_ne[3] = {0, 0, 0};
    int grid_corner[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double ke_dev = HUGE_VAL;
    bool from_scales = false;       // values summed from s_e Ke(0), not from the element matrices
    sfem::DevBuf<double> escale;    // n_elem: s_e = Ke(e)[0][0] / Ke(0)[0][0]
    std::vector<double> ke0;        // Ke(0), blocks (a*nen + b)*bs*bs + i*bs + j (local corners a, b)
    // device time per assembly stage (ms)
    double t_validate = 0, t_element = 0, t_keys = 0, t_sort = 0, t_reduce = 0, t_total = 0;
};

struct sfem_bc {
    sfem_ctx* ctx = nullptr;
    int64_t n = 0;
    sfem::DevBuf<double> f;        // n
    sfem::DevBuf<uint8_t> fixed;   // n
    sfem::DevBuf<double> qd;       // n (0 on free dofs)
    bool qd_zero = false;          // known: no prescribed value is nonzero (rhs is f on the free dofs)
};

namespace sfem {

// ---- element.hip ----
struct ValidateOut {
    unsigned long long err_key;   // min over elements of element*8 + step
    int max_unit;                 // max 0-based unit (node / dof / node-block) referenced
    int min_dof;                  // min stored dof code over used nodes (-1: id<1, -2: too large)
    int blocked;                  // 1 while every used node's ID is ndof*g + l
    int pad;
};
// d_out: kValidateCopies records; block b of the validation reduces into copy
// b % kValidateCopies (one record for every block serialises its atomics),
// the host combines them (combine_validate).  The element kernels report
// singular Jacobians into copy 0.
constexpr int kValidateCopies = 8;
void launch_validate(sfem_ctx* ctx, const sfem_mesh* m, ValidateOut* d_out);
inline ValidateOut combine_validate(const ValidateOut* v) {
    ValidateOut r = v[0];
    for (int i = 1; i < kValidateCopies; ++i) {
        r.err_key = v[i].err_key < r.err_key ? v[i].err_key : r.err_key;
        r.max_unit = v[i].max_unit > r.max_unit ? v[i].max_unit : r.max_unit;
        r.min_dof = v[i].min_dof < r.min_dof ? v[i].min_dof : r.min_dof;
        r.blocked = r.blocked && v[i].blocked;
    }
    return r;
}
void launch_element_stiffness(sfem_ctx* ctx, const sfem_mesh* m, double* V, ValidateOut* d_out);
// The element-matrix check of a structured-grid assembly, done while the
// element kernel stores Ke: with ke0 = element 0's Ke (computed first, by
// launch_element_zero), escale[e] = s_e = Ke(e)[0][0] / Ke(0)[0][0] and, per
// launch block b of the element kernel, bmax[b] = max |Ke(e) (1 / s_e) - Ke(0)|
// over its elements' entries (NaN if any is).  Hex8 / Q4 vector meshes only.
// max_nan: the larger of a and b, NaN when either is.
__device__ __forceinline__ double max_nan(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }
struct ElemCheck {
    const double* ke0;
    double* escale;
    double* bmax;
};
unsigned element_launch_blocks(const sfem_mesh* m);
void launch_element_zero(sfem_ctx* ctx, const sfem_mesh* m, double* ke0, ValidateOut* d_out);
void launch_element_checked(sfem_ctx* ctx, const sfem_mesh* m, double* V, ValidateOut* d_out,
                            const ElemCheck& chk);
// Emit (key, payload) for every unit pair; returns pair count.
int64_t launch_keys(sfem_ctx* ctx, const sfem_mesh* m, int bs, int row_bits, uint64_t* keys,
                    uint32_t* payload);

// ---- radix_sort.hip ----
// Stable LSD sort of (key, value) pairs on the low `key_bits` bits.  Buffers
// are ping-ponged; on return *keys / *vals point at the sorted data (either
// the input or the alternate buffer).
void radix_sort_pairs(sfem_ctx* ctx, uint64_t** keys, uint64_t* keys_alt, uint32_t** vals,
                      uint32_t* vals_alt, int64_t n, int key_bits);
// Exclusive scan of n uint32 values (in may equal out); returns the total.
uint64_t exclusive_scan_u32(sfem_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n);
// Same without a host sync; total written to d_total if non-null.
void exclusive_scan_u32_dev(sfem_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n,
                            uint32_t* d_total);
// Exclusive scan into 64-bit offsets; returns the total.
int64_t exclusive_scan_i64(sfem_ctx* ctx, const int64_t* in, int64_t* out, int64_t n);

// ---- surface.hip ----
// COO values of the surface penalty, same layout as launch_element_stiffness.
void launch_face_mass(sfem_ctx* ctx, const sfem_mesh* m, double* V);

// ---- assemble.hip ----
void assemble(sfem_ctx* ctx, const sfem_mesh* m, int flags, sfem_csr* K);
void mesh_set_grid(sfem_ctx* ctx, sfem_mesh* m, int ndim, const int64_t* ne);
// Declares the grid when the mesh is one (element counts from element 1's
// corners, then mesh_set_grid's check); false otherwise, mesh unchanged.
bool mesh_detect_grid(sfem_ctx* ctx, sfem_mesh* m, int64_t* ne);
void export_csc1(sfem_ctx* ctx, const sfem_csr* K, int64_t* colptr, int64_t* rowval, double* nzval);
void upload_csc1(sfem_ctx* ctx, int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const double* nzval, sfem_csr* K);

}  // namespace sfem

namespace sfem {
// Arrival counters of a last-block reduction (last_block_arrive): 8 group
// counters and a top counter, each on its own 256-byte line.
struct Arrivals {
    unsigned c[9 * 64];
};
}  // namespace sfem

// Device scalars of one CG solve; updated only by the last block of each
// reduction, read by every kernel of the next step.
struct CgState {
    double rz, rz_prev, beta, alpha, rr, bb, maxrr, tol2bb, delta2, rr_true;
    double zcoef;        // p_new = zcoef*z + beta*p; 1 except right after a residual replacement
    double best_rr;      // smallest recurrence |r|^2 seen (stagnation check)
    double best_true;    // smallest true |r|^2 at a replacement
    double floor_rr;     // 16 (eps || |rhs| + |K||x| ||)^2 at the last replacement
    int done;            // 1 converged, 2 breakdown, 3 stagnated
    int need_update;     // the true FP64 residual is due
    int fold_pending;    // FP32 solution increment to be folded into x64
    int stall;           // replacements in a row without progress
    long long iters, updates, best_it;
    long long last_upd;  // iteration of the latest replacement
    long long best_true_it;  // iteration of best_true
    sfem::Arrivals arrive;
};

// K_ff and its solver formats, plus the data to recombine the solution.
namespace sfem {
struct MgHierarchy;  // mg.hip
struct MgDelete {
    void operator()(MgHierarchy* h) const;
};
}  // namespace sfem

struct sfem_system {
    sfem_ctx* ctx = nullptr;
    const sfem_csr* K = nullptr;   // referenced (sfem_csr::refs) until the system is destroyed
    int64_t n = 0, nf = 0;         // all dofs, free dofs
    int bs = 1;                    // block size of K
    sfem::DevBuf<int32_t> free_of; // n: free index or -1
    sfem::DevBuf<double> qd, f;    // n
    sfem::DevBuf<double> rhs;      // nf: C'(f - K q_d)
    sfem::DevBuf<double> x;        // nf: FP64 solution
    // scalar CSR of K_ff, built on first use: rowptr and nnz (-1 until then,
    // ensure_rowptr), col / val64 (ensure_csr)
    int64_t nnz = -1;
    sfem::DevBuf<int32_t> rowptr, col;
    sfem::DevBuf<double> val64;
    sfem::DevBuf<float> val32;     // built on first mixed CSR use
    sfem::DevBuf<double> dinv64;
    sfem::DevBuf<float> dinv32;
    // node-block sliced format (when bs > 1 and every node is all-fixed or all-free)
    bool has_block = false;
    int64_t nfb = 0, nslices = 0, nnzb = 0, sell_blocks = 0;
    sfem::DevBuf<int64_t> sl_ptr;  // nslices + 1, in blocks (64 per slot)
    sfem::DevBuf<int32_t> sl_w;    // nslices: slots per slice
    sfem::DevBuf<int32_t> scol;    // sell_blocks
    sfem::DevBuf<double> sval64;   // sell_blocks * bs*bs, planes of 64
    sfem::DevBuf<float> sval32;
    sfem::DevBuf<int16_t> sval16;  // correction of sval32 toward sval64 (split precision)
    sfem::DevBuf<double> binv64;   // inverse node diagonal blocks, planar: [(j*bs+i)*nfb + node]
    sfem::DevBuf<float> binv32;
    // Partially fixed nodes: the block rows carry their fixed dofs with a zero
    // row / column and a unit diagonal; the block solver then works on padded
    // vectors of nfb*bs values, mapped to the free dofs by pad_of (-1 = fixed).
    bool partial = false;
    sfem::DevBuf<int32_t> pad_of;
    // node -> block row (-1: every dof fixed); the sliced values (scol,
    // sval*) are filled on first use (ensure_sell), the point-Jacobi
    // 1/diag too (ensure_dinv)
    sfem::DevBuf<int32_t> ufree_of;
    bool sell_ready = false;
    sfem::DevBuf<double> prhs, px, pdinv64;
    sfem::DevBuf<float> pdinv32;
    double t_eliminate_ms = 0;
    // structured grid of the nodes (sfem_system_set_grid) and the multigrid
    // hierarchy built on it; it depends on K and the fixed mask only, so new
    // loads or prescribed values (system_set_bc) keep it
    int grid_nd = 0;
    int64_t grid_dims[3] = {0, 0, 0};
    sfem_mg_opts mg_opts{};
    std::unique_ptr<sfem::MgHierarchy, sfem::MgDelete> mg;
};

namespace sfem {
// ---- eliminate.hip ----
void eliminate(sfem_ctx* ctx, const sfem_csr* K, const sfem_bc* bc, sfem_system* S);
void system_solution(sfem_ctx* ctx, const sfem_system* S, double* q, double* reactions);
void system_set_bc(sfem_ctx* ctx, sfem_system* S, const sfem_bc* bc);
// the scalar CSR of K_ff: row pointers and nnz (ensure_rowptr), columns and
// FP64 values (ensure_csr), built on first use
void ensure_sell(sfem_ctx* ctx, sfem_system* S);
void ensure_dinv(sfem_ctx* ctx, sfem_system* S);
void ensure_rowptr(sfem_ctx* ctx, sfem_system* S);
void ensure_csr(sfem_ctx* ctx, sfem_system* S);
// padded block vector <- compact free-dof vector (0 on padded fixed dofs), and back
void pad_gather(sfem_ctx* ctx, const sfem_system* S, const double* compact, double* padded);
void pad_scatter(sfem_ctx* ctx, const sfem_system* S, const double* padded, double* compact);
// k right-hand sides C'(f_j - K q_d_j) into rhs (n_free x kp, row-major; columns >= k zero).
void launch_rhs_many(sfem_ctx* ctx, const sfem_system* S, int k, int kp, const double* qd, const double* f,
                     double* rhs);
// ---- pcg.hip ----
void system_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st);
// the solve on the per-device workspace array (no device allocation)
constexpr int64_t kNoAllocMaxDofs = int64_t(1) << 20;
void system_solve_noalloc(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st);
void* noalloc_scratch_address();  // for the current device
size_t noalloc_scratch_bytes();
void system_spmv_bench(sfem_ctx* ctx, sfem_system* S, int precision, int format, int reps, double* avg_ms,
                       double* bytes);
void stream_copy_bench(sfem_ctx* ctx, int64_t bytes, int reps, double* avg_ms);
// ---- mg.hip ----
void mg_opts_default(sfem_mg_opts* o);
void system_set_grid(sfem_ctx* ctx, sfem_system* S, int ndim, const int64_t* dims, const sfem_mg_opts* opts);
void mg_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st);
void mg_info(const sfem_system* S, sfem_mg_info* info);
void mg_stencil(sfem_ctx* ctx, const sfem_system* S, int level, double* stencil, uint8_t* fixed);
}  // namespace sfem
