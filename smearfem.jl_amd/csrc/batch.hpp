// batch.hpp -- the table-driven persistent PCG engine shared by batches and
// the serving loop: per-problem descriptors, work items, CG state.
#pragma once

#include <utility>
#include <vector>

#include "sfem_objects.hpp"

namespace sfem {

// CG state of one problem (device memory; written by the persistent kernel).
struct ProbState {
    double rtol, delta;  // set by the host
    double bb, tol2bb, delta2, maxrr;
    double rz, rz_prev, beta, zcoef;
    double rr_true, best_rr, best_true, rr_rec, floor_rr;
    int64_t max_iter, iters, updates, best_it, last_upd, best_true_it;
    int32_t init_pending, need_update, fold_pending, done;  // done: 0 running, 1 converged, 2 breakdown,
    int32_t stall, capped, pad0, pad1;                      // 3 stagnation, 4 iteration cap
};

// One problem of a table: its sliced operator and vectors (padded layout,
// 64-row slices, vectors of nslices*64*bs entries).  Pointers are typed by the
// launch's precision: sval is the FP32 plane in mixed and split modes (plus
// sval16 in split mode), the FP64 plane otherwise; dinv/binv/r/z/p/q/xs are
// float in mixed mode.  binv: inverse node diagonal blocks, planar with
// stride binv_stride (node rows of the operator); used when block_pc is 1.
struct ProbDesc {
    const int64_t* sl_ptr;
    const int32_t* sl_w;
    const int32_t* scol;
    const void* sval;
    const int16_t* sval16;
    const double* sval64;
    const double* rhs;
    double* x64;
    void* xs;
    void* r;
    void* z;
    void* p;
    void* q;
    const void* dinv;
    const void* binv;
    ProbState* st;
    int64_t binv_stride;
    int32_t bs, item0, item1, block_pc;
};

// A run of slices [slice0, slice1) of one problem, processed by one block.
struct WorkItem {
    int32_t prob, pad;
    int64_t slice0, slice1;
};

struct TableLaunch {
    const ProbDesc* probs;
    int nprob;
    const WorkItem* items;
    int nitems;
    double* part;   // 4 per item
    double* partr;  // 4 per item
    int64_t iters;  // loop passes in this launch
    unsigned* bar_ws;                 // grid barrier workspace (kBarWords, zeroed per launch)
    unsigned long long bar_timeout;   // barrier wait limit, 100 MHz ticks
    int bar_grouped;                  // allow the per-XCD grouped barrier
};

struct BatchProb {
    int64_t slice0, slice1, n_free;
};

size_t prec_vec_bytes(int precision);
int table_max_blocks_per_cu(int precision);
int table_grid(sfem_ctx* ctx, int precision, int blocks_per_cu);
void table_launch(sfem_ctx* ctx, int precision, const TableLaunch& a, int grid);
void table_fold(sfem_ctx* ctx, double* x64, void* xs, int64_t n);
ProbState initial_state(const sfem_solve_opts& o);

}  // namespace sfem

struct sfem_batch {
    sfem_ctx* ctx = nullptr;
    int64_t nprob = 0;
    int ndim = 0, ndof = 0, bs = 0;
    int64_t n = 0, npad = 0, nslices = 0, sell_blocks = 0;
    std::vector<int64_t> node_off;
    std::vector<sfem::BatchProb> hprobs;
    // inputs kept only while the padded operator is built
    sfem_mesh* mesh = nullptr;
    sfem_csr* K = nullptr;
    sfem_bc* bc = nullptr;
    sfem::DevBuf<int32_t> pad_of;  // node unit -> padded row (-1 fixed)
    sfem::DevBuf<uint8_t> fixed;   // dof mask (a node may have fixed and free dofs)
    sfem::DevBuf<int64_t> sl_ptr;
    sfem::DevBuf<int32_t> sl_w, scol;
    sfem::DevBuf<double> sval64, rhs, x, dinv64, qd;
    sfem::DevBuf<float> sval32, dinv32, binv32;
    sfem::DevBuf<int16_t> sval16;
    sfem::DevBuf<double> binv64;  // planar, stride nslices * 64
    // CG work vectors, typed by the last precision (work_vb bytes per entry)
    size_t work_vb = 0;
    sfem::DevBuf<uint8_t> wr, wz, wp, wq, wxs;
    sfem::DevBuf<sfem::ProbState> st;
    double t_assemble_ms = 0, t_eliminate_ms = 0;
    ~sfem_batch();
};

namespace sfem {

sfem_batch* batch_create(sfem_ctx* ctx, int64_t nprob, int ndim, int n_dof, const double* node_list,
                         const int64_t* node_offsets, const int64_t* ien, int nen, const int64_t* elem_offsets,
                         const double* young, const double* nu, const double* f, const uint8_t* fixed,
                         const double* q_d);
// problems already on the device (one concatenated mesh, node ids offset per
// problem, and its bc); takes ownership of mesh and bc
sfem_batch* batch_from_device(sfem_ctx* ctx, int64_t nprob, sfem_mesh* mesh, sfem_bc* bc,
                              const std::vector<int64_t>& node_offsets, const std::vector<int64_t>& nfree);
// structured problems generated from specs (spec.hip): one concatenated
// device mesh and bc; one problem: the mesh is declared its grid
void spec_build(sfem_ctx* ctx, int64_t nprob, const sfem_problem_spec* specs, sfem_mesh** mesh, sfem_bc** bc,
                std::vector<int64_t>* node_off);
int64_t spec_free_dofs(const sfem_problem_spec& s);
void batch_items(const sfem_batch* B, int item_slices, std::vector<WorkItem>* items,
                 std::vector<std::pair<int, int>>* ranges);
void batch_ensure_work(sfem_ctx* ctx, sfem_batch* B, int precision);
ProbDesc batch_desc(sfem_batch* B, int precision, int preconditioner, ProbState* st, int item0, int item1);
void batch_reset(sfem_ctx* ctx, sfem_batch* B, int precision, const sfem_solve_opts& o, ProbState* d_st);
void batch_solution(sfem_ctx* ctx, const sfem_batch* B, double* q);
int64_t batch_device_bytes(const sfem_batch* B, int precision);

}  // namespace sfem
