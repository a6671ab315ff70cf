// serve.hip -- continuous batching of solve requests over the persistent
// table solver (batch.hip).
//
// Requests queue on the host.  Each step: (1) admit queued requests in
// arrival order while a slot is free and their device footprint fits the HBM
// budget -- each admitted request is assembled and eliminated on the device
// into its own padded sliced operator (a one-problem batch) and gets a CG
// state slot; (2) run iters_per_step iterations of every active request in
// one cooperative launch; (3) retire finished requests: solution to the host,
// device memory back to the allocator.  Requests that arrive while others
// run join at the next step and share the launch with them.
//
// A request's arithmetic depends only on its own operator, item_slices and
// precision (per-problem reductions over its own items, state in memory), so
// its result is bitwise the same alone or in any mix.
//
// Multigrid servers solve requests that come with their grid (or as a spec,
// generated on the device) to completion with the multigrid preconditioner
// (mg.hip), several at once: mg_workers host threads, each with its own
// context (stream and allocator), take queued grid requests in arrival
// order while the HBM budget allows, so the requests' kernels overlap on the
// device beside the persistent launch.  A worker's arithmetic depends only on
// its request, so results are bitwise the same alone or in any mix.  A grid
// that cannot be coarsened joins the persistent launch.
//
// No counterpart in the reference (one blocking solve per call there,
// examples/vector3D.jl:314-322).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace {

struct Request {
    int64_t id = 0;
    int state = SFEM_REQ_QUEUED;
    int ndim = 0;
    bool grid = false;      // nodes form the structured grid of dims
    bool spec = false;      // generated on the device from sp (no host arrays)
    sfem_problem_spec sp{};
    int64_t dims[3] = {0, 0, 0};
    int64_t n_nodes = 0, n_elem = 0;
    std::vector<double> nodes, young, nu, f, qd;
    std::vector<int64_t> ien;
    std::vector<uint8_t> fixed;
    int64_t est_resident = 0, est_transient = 0, bytes = 0;
    sfem_batch* B = nullptr;
    int slot = -1;
    std::vector<double> result;
    sfem_batch_problem_stats stats{};
    double t_sub = 0, t_adm = 0, t_fin = 0;
    std::string error;
};

// A host thread solving multigrid requests on its own context.
struct Worker {
    std::thread th;
    sfem_ctx* ctx = nullptr;
    int64_t rid = -1;      // assigned request, -1 when idle
    int64_t reserved = 0;  // HBM budget bytes held for it
    bool finished = false; // solved, not yet harvested by a step
    bool coarsens = true;  // false: the grid has no multigrid hierarchy
    int result = SFEM_REQ_ACTIVE;  // the request's state once harvested (DONE / FAILED)
    double ms = 0;         // wall time of the solve
};

}  // namespace

struct sfem_server {
    sfem_ctx* ctx = nullptr;
    sfem_server_opts o{};
    std::vector<std::unique_ptr<Worker>> workers;
    mutable std::mutex mu;             // request map and states, worker slots
    std::condition_variable cv_work;   // a worker has a request, or the server stops
    std::condition_variable cv_done;   // a worker finished
    bool stopping = false;
    std::map<int64_t, Request> reqs;
    std::deque<int64_t> queue;   // requests for the persistent launch
    std::deque<int64_t> gqueue;  // multigrid servers: requests solved one at a time
    int table_pc = SFEM_PC_JACOBI;
    std::vector<int64_t> active;
    std::vector<int> free_slots;
    sfem::DevBuf<sfem::ProbState> pool;
    std::vector<sfem::ProbState> hpool;
    sfem::DevBuf<sfem::ProbDesc> ddesc;
    sfem::DevBuf<sfem::WorkItem> ditems;
    sfem::DevBuf<double> part, partr;
    int64_t bytes_in_use = 0, next_id = 0;
    int grid = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::chrono::steady_clock::time_point t0;
    ~sfem_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv_work.notify_all();
        for (auto& w : workers) {
            if (w->th.joinable()) w->th.join();
            if (w->ctx) sfem_ctx_destroy(w->ctx);
        }
        for (auto& kv : reqs)
            if (kv.second.B) sfem_batch_destroy(kv.second.B);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace sfem {
namespace {

// Conservative device footprint of a request before it is built: every row
// padded to the largest possible node degree (elements at the node x nodes
// per element), both operator precisions, the vectors, and the transient
// assembly buffers (element matrices, sort keys, block CSC).
void estimate_bytes(Request& r, int precision) {
    const int nen = 1 << r.ndim, bs = r.ndim;
    int64_t maxdeg = 1;
    if (r.spec) {  // a structured grid: at most 3^ndim neighbours
        maxdeg = std::min<int64_t>(r.n_nodes, r.ndim == 3 ? 27 : 9);
    } else {
        std::vector<int32_t> cnt((size_t)r.n_nodes, 0);
        for (int64_t i = 0; i < r.n_elem * nen; ++i) {
            const int64_t v = r.ien[(size_t)i];
            if (v >= 1 && v <= r.n_nodes) ++cnt[(size_t)(v - 1)];
        }
        for (int32_t c : cnt)
            maxdeg = std::max<int64_t>(maxdeg, std::min<int64_t>(r.n_nodes, (int64_t)c * (nen - 1) + 1));
    }
    const int64_t rows = (r.n_nodes + 63) / 64 * 64;
    const int64_t slots = rows * maxdeg;
    const int64_t b2 = (int64_t)bs * bs, vb = (int64_t)prec_vec_bytes(precision);
    r.est_resident = slots * (4 + 14 * b2) + rows / 64 * 12 + rows * bs * (8 * 3 + 4 + 4 * vb + 4) +
                     rows * b2 * 12 + r.n_nodes * 4 + r.n_nodes * bs * 8;
    const int64_t ke = (int64_t)nen * bs * nen * bs;
    r.est_transient = r.est_resident + r.n_elem * (ke * 8 + (int64_t)nen * nen * 24) + slots * (8 + b2 * 8);
}

double launch_ms(sfem_server* S) {
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, S->e0, S->e1));
    return ms;
}

void admit(sfem_server* S) {
    sfem_ctx* ctx = S->ctx;
    while (!S->queue.empty() && !S->free_slots.empty()) {
        Request& r = S->reqs.at(S->queue.front());
        const bool fits = S->bytes_in_use + r.est_resident <= S->o.hbm_budget &&
                          S->bytes_in_use + r.est_transient <= S->o.hbm_budget;
        if (!fits) {
            if (!S->active.empty()) break;  // wait for memory to free up
            r.state = SFEM_REQ_FAILED;      // cannot fit even on an idle device
            r.error = "request does not fit the HBM budget";
            r.t_fin = S->now_ms();
            S->queue.pop_front();
            continue;
        }
        S->queue.pop_front();
        const int64_t nen = (int64_t)1 << r.ndim;
        const int64_t noff[2] = {0, r.n_nodes}, eoff[2] = {0, r.n_elem};
        sfem_batch* B = nullptr;
        const int st = guard(ctx, [&] {
            if (r.spec) {  // generated on the device
                sfem_mesh* mesh = nullptr;
                sfem_bc* bc = nullptr;
                std::vector<int64_t> off;
                spec_build(ctx, 1, &r.sp, &mesh, &bc, &off);
                ctx_retain(ctx);  // released by the destroy calls batch_from_device makes
                ctx_retain(ctx);
                B = batch_from_device(ctx, 1, mesh, bc, off, {spec_free_dofs(r.sp)});
                return;
            }
            B = batch_create(ctx, 1, r.ndim, r.ndim, r.nodes.data(), noff, r.ien.data(), (int)nen, eoff,
                             r.young.data(), r.nu.data(), r.f.empty() ? nullptr : r.f.data(), r.fixed.data(),
                             r.qd.empty() ? nullptr : r.qd.data());
        });
        if (st) {
            r.state = SFEM_REQ_FAILED;
            r.error = ctx->last_error;
            r.t_fin = S->now_ms();
            continue;
        }
        ctx_retain(ctx);  // held by B, released by sfem_batch_destroy
        r.B = B;
        sfem_solve_opts o;
        sfem_solve_opts_default(&o);
        o.precision = S->o.precision;
        o.rtol = S->o.rtol;
        o.max_iter = S->o.max_iter;
        o.reliable_delta = S->o.reliable_delta;
        o.x0_zero = 1;
        r.slot = S->free_slots.back();
        S->free_slots.pop_back();
        batch_reset(ctx, B, o.precision, o, nullptr);
        const ProbState init = initial_state(o);
        h2d(ctx, S->pool.get() + r.slot, &init, sizeof(ProbState));
        r.bytes = batch_device_bytes(B, o.precision);
        S->bytes_in_use += r.bytes;
        r.state = SFEM_REQ_ACTIVE;
        r.t_adm = S->now_ms();
        S->active.push_back(r.id);
        // the host copies of the inputs are no longer needed
        std::vector<double>().swap(r.nodes);
        std::vector<int64_t>().swap(r.ien);
    }
}

void retire(sfem_server* S, int* n_finished) {
    sfem_ctx* ctx = S->ctx;
    std::vector<int64_t> still;
    for (int64_t id : S->active) {
        Request& r = S->reqs.at(id);
        const ProbState& st = S->hpool[(size_t)r.slot];
        if (!st.done) {
            still.push_back(id);
            continue;
        }
        if (S->o.precision == SFEM_PREC_MIXED) table_fold(ctx, r.B->x.get(), r.B->wxs.get(), r.B->npad);
        r.result.resize((size_t)r.B->n);
        batch_solution(ctx, r.B, r.result.data());
        r.stats.iterations = st.iters;
        r.stats.fp64_updates = st.updates;
        r.stats.rel_residual = st.bb > 0 ? std::sqrt(st.rr_true / st.bb) : 0.0;
        r.stats.n_free = r.B->hprobs[0].n_free;
        r.stats.stop_reason = st.done;
        r.stats.floor_residual = st.bb > 0 ? std::sqrt(st.floor_rr / 16.0 / st.bb) : 0.0;
        r.state = SFEM_REQ_DONE;
        r.t_fin = S->now_ms();
        S->bytes_in_use -= r.bytes;
        sfem_batch_destroy(r.B);
        r.B = nullptr;
        S->free_slots.push_back(r.slot);
        r.slot = -1;
        if (n_finished) ++*n_finished;
    }
    S->active.swap(still);
}

// One request on its own on ctx (a worker's context): build, assemble,
// eliminate, multigrid hierarchy, FP64 CG.  Writes the result, stats and
// error into r and returns the request state (SFEM_REQ_DONE / FAILED), or
// SFEM_REQ_QUEUED when the grid has no multigrid hierarchy (the request then
// joins the persistent launch, its inputs kept).  Touches no other request.
int solve_grid_request(sfem_server* S, sfem_ctx* ctx, Request& r, double* ms) {
    const int nd = r.ndim, nen = 1 << nd;
    const int64_t N = r.n_nodes, E = r.n_elem, n = N * nd;
    sfem_mesh* mesh = nullptr;
    sfem_csr* K = nullptr;
    sfem_bc* bc = nullptr;
    sfem_system* sys = nullptr;
    auto release = [&] {
        if (sys) sfem_system_destroy(sys);
        if (bc) sfem_bc_destroy(bc);
        if (K) sfem_csr_destroy(K);
        if (mesh) sfem_mesh_destroy(mesh);
    };
    auto ok = [&](int st) {
        if (st) fail(st, ctx->last_error);
    };
    bool coarsens = true;
    int state = SFEM_REQ_FAILED;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        if (r.spec) {
            ok(sfem_spec_build(ctx, &r.sp, &mesh, &bc));
        } else {
            std::vector<int64_t> gid((size_t)n);
            for (int64_t m = 0; m < N; ++m)
                for (int l = 0; l < nd; ++l) gid[(size_t)((int64_t)l * N + m)] = (int64_t)nd * m + l + 1;
            ok(sfem_mesh_create(ctx, nd, SFEM_Q1, nd, r.nodes.data(), N, r.ien.data(), E, nen, E, gid.data(), N,
                                r.young.data(), E, 0.3, &mesh));
            mesh->nu_e.reset(ctx, (size_t)E);
            h2d(ctx, mesh->nu_e.get(), r.nu.data(), sizeof(double) * E);
            try {
                mesh_set_grid(ctx, mesh, nd, r.dims);  // grid-pattern assembly when the mesh is the grid
            } catch (const Error& e) {
                if (e.code != SFEM_ERR_ARGUMENT) throw;
            }
            ok(sfem_bc_create(ctx, n, r.f.empty() ? nullptr : r.f.data(), r.fixed.data(),
                              r.qd.empty() ? nullptr : r.qd.data(), &bc));
        }
        ok(sfem_assemble_mesh(ctx, mesh, SFEM_ASM_DEFAULT, &K));
        if (K->n != n || K->bs != nd) fail(SFEM_ERR_ARGUMENT, "server: every node must belong to an element");
        ok(sfem_eliminate(ctx, K, bc, &sys));
        try {
            system_set_grid(ctx, sys, nd, r.dims, nullptr);
        } catch (const Error& e) {
            if (e.code != SFEM_ERR_UNSUPPORTED) throw;
            coarsens = false;
        }
        if (coarsens) {
            sfem_solve_opts o;
            sfem_solve_opts_default(&o);
            o.precision = SFEM_PREC_FP64;
            o.format = SFEM_FMT_BSR;
            o.rtol = S->o.rtol;
            o.max_iter = S->o.max_iter;
            o.x0_zero = 1;
            o.preconditioner = SFEM_PC_MULTIGRID;
            sfem_solve_stats st{};
            try {
                system_solve(ctx, sys, &o, &st);
            } catch (const Error& e) {
                if (e.code != SFEM_ERR_NOT_CONVERGED) throw;
            }
            r.result.resize((size_t)n);
            system_solution(ctx, sys, r.result.data(), nullptr);
            r.stats.iterations = st.iterations;
            r.stats.fp64_updates = st.fp64_updates;
            r.stats.rel_residual = st.rel_residual;
            r.stats.n_free = sys->nf;
            r.stats.stop_reason = st.stop_reason;
            r.stats.floor_residual = st.floor_residual;
            // stagnation at the FP64 floor keeps the best attainable answer: done,
            // as in the persistent launch (stop_reason 3 tells the caller)
            state = st.converged || st.stop_reason == 3 ? SFEM_REQ_DONE : SFEM_REQ_FAILED;
            if (state == SFEM_REQ_FAILED)
                r.error = "request did not converge (multigrid CG, stop reason " + std::to_string(st.stop_reason) +
                          ", relative residual " + std::to_string(st.rel_residual) + ")";
        }
    } catch (const Error& e) {
        release();
        r.error = e.msg;
        return SFEM_REQ_FAILED;
    }
    release();
    if (!coarsens) return SFEM_REQ_QUEUED;
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return state;
}

// A worker: waits for a request, solves it on its own context, reports back.
void worker_loop(sfem_server* S, Worker* W) {
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_work.wait(lk, [&] { return S->stopping || (W->rid >= 0 && !W->finished); });
        if (S->stopping) return;
        Request& r = S->reqs.at(W->rid);  // map nodes stay put; no one else touches r meanwhile
        lk.unlock();
        double ms = 0;
        const int state = solve_grid_request(S, W->ctx, r, &ms);
        lk.lock();
        // the request stays ACTIVE until a step harvests it (a caller polling
        // it then finds its slot released as well); its finish time is now
        W->coarsens = state != SFEM_REQ_QUEUED;
        W->result = state;
        if (W->coarsens) r.t_fin = S->now_ms();
        W->ms = ms;
        W->finished = true;
        S->cv_done.notify_all();
    }
}

// Harvest finished workers, then hand queued grid requests to idle workers
// in arrival order while the HBM budget allows.  Caller holds S->mu.
void step_grid(sfem_server* S, int* n_finished, double* ms) {
    for (auto& w : S->workers) {
        if (w->rid < 0 || !w->finished) continue;
        Request& r = S->reqs.at(w->rid);
        S->bytes_in_use -= w->reserved;
        if (w->coarsens) {
            *ms += w->ms;
            r.state = w->result;
            if (r.state == SFEM_REQ_DONE && n_finished) ++*n_finished;
            if (r.state == SFEM_REQ_DONE) {
                std::vector<double>().swap(r.nodes);
                std::vector<int64_t>().swap(r.ien);
            }
        } else {  // no multigrid hierarchy: the persistent launch
            r.state = SFEM_REQ_QUEUED;
            r.t_adm = 0;
            S->queue.push_back(r.id);
        }
        w->rid = -1;
        w->reserved = 0;
        w->finished = false;
    }
    bool assigned = false;
    while (!S->gqueue.empty()) {
        Worker* idle = nullptr;
        bool busy = false;
        for (auto& w : S->workers) {
            if (w->rid < 0) idle = idle ? idle : w.get();
            else busy = true;
        }
        if (!idle) break;
        Request& r = S->reqs.at(S->gqueue.front());
        // the request's build and hierarchy are held only while it is solved
        const int64_t need = r.est_transient + r.est_transient / 2;
        if (S->bytes_in_use + need > S->o.hbm_budget) {
            if (!S->active.empty() || busy) break;  // wait for memory to free up
            r.state = SFEM_REQ_FAILED;
            r.error = "request does not fit the HBM budget";
            r.t_fin = S->now_ms();
            S->gqueue.pop_front();
            continue;
        }
        S->gqueue.pop_front();
        r.state = SFEM_REQ_ACTIVE;
        r.t_adm = S->now_ms();
        S->bytes_in_use += need;
        idle->rid = r.id;
        idle->reserved = need;
        idle->finished = false;
        assigned = true;
    }
    if (assigned) S->cv_work.notify_all();
}

int64_t busy_workers(const sfem_server* S) {
    int64_t k = 0;
    for (const auto& w : S->workers) k += w->rid >= 0 ? 1 : 0;
    return k;
}

void step(sfem_server* S, int* n_finished, double* ms) {
    sfem_ctx* ctx = S->ctx;
    if (n_finished) *n_finished = 0;
    if (ms) *ms = 0;
    double grid_ms = 0;
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->workers.empty()) {
        step_grid(S, n_finished, &grid_ms);
        // nothing for the persistent launch while workers solve: wait (briefly)
        // for one of them, so that a polling caller does not spin
        if (S->queue.empty() && S->active.empty() && busy_workers(S) > 0 && !(n_finished && *n_finished)) {
            S->cv_done.wait_for(lk, std::chrono::milliseconds(1));
            step_grid(S, n_finished, &grid_ms);
        }
    }
    if (ms) *ms = grid_ms;
    admit(S);
    lk.unlock();
    if (S->active.empty()) return;
    std::vector<ProbDesc> descs;
    std::vector<WorkItem> items;
    for (int64_t id : S->active) {
        Request& r = S->reqs.at(id);
        std::vector<WorkItem> its;
        std::vector<std::pair<int, int>> ranges;
        batch_items(r.B, S->o.item_slices, &its, &ranges);
        const int base = (int)items.size();
        for (WorkItem w : its) {
            w.prob = (int32_t)descs.size();
            items.push_back(w);
        }
        descs.push_back(batch_desc(r.B, S->o.precision, S->table_pc, S->pool.get() + r.slot, base,
                                   (int)items.size()));
    }
    if (S->ddesc.size() < descs.size()) S->ddesc.reset(ctx, descs.size() * 2);
    if (S->ditems.size() < items.size()) {
        S->ditems.reset(ctx, items.size() * 2);
        S->part.reset(ctx, items.size() * 8);
        S->partr.reset(ctx, items.size() * 8);
    }
    h2d(ctx, S->ddesc.get(), descs.data(), sizeof(ProbDesc) * descs.size());
    h2d(ctx, S->ditems.get(), items.data(), sizeof(WorkItem) * items.size());
    TableLaunch a{};
    a.probs = S->ddesc.get();
    a.nprob = (int)descs.size();
    a.items = S->ditems.get();
    a.nitems = (int)items.size();
    a.part = S->part.get();
    a.partr = S->partr.get();
    a.iters = S->o.iters_per_step;
    SFEM_HIP(hipEventRecord(S->e0, ctx->stream));
    table_launch(ctx, S->o.precision, a, S->grid);
    SFEM_HIP(hipEventRecord(S->e1, ctx->stream));
    d2h(ctx, S->hpool.data(), S->pool.get(), sizeof(ProbState) * S->hpool.size());
    if (ms) *ms = grid_ms + launch_ms(S);
    retire(S, n_finished);
}

}  // namespace
}  // namespace sfem

using namespace sfem;

extern "C" {

void sfem_server_opts_default(sfem_server_opts* o) {
    if (!o) return;
    o->precision = SFEM_PREC_FP64;
    o->rtol = 1e-10;
    o->max_iter = 100000;
    o->reliable_delta = 0.1;
    o->item_slices = 8;
    o->blocks_per_cu = 0;
    o->hbm_budget = (int64_t)64 << 30;
    o->max_active = 256;
    o->iters_per_step = 64;
    o->preconditioner = SFEM_PC_JACOBI;
    o->mg_workers = 4;
}

int sfem_server_create(sfem_ctx* ctx, const sfem_server_opts* opts, sfem_server** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out) { ctx->last_error = "server_create: NULL out"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    std::unique_ptr<sfem_server> S(new (std::nothrow) sfem_server());
    if (!S) return SFEM_ERR_MEMORY;
    const int st = guard(ctx, [&] {
        sfem_server_opts_default(&S->o);
        if (opts) S->o = *opts;
        if (S->o.max_active < 1 || S->o.iters_per_step < 1 || S->o.item_slices < 1 || S->o.hbm_budget <= 0)
            fail(SFEM_ERR_ARGUMENT, "server_create: max_active, iters_per_step, item_slices and hbm_budget must be > 0");
        if (S->o.precision < SFEM_PREC_FP64 || S->o.precision > SFEM_PREC_SPLIT)
            fail(SFEM_ERR_ARGUMENT, "server_create: unknown precision");
        if (S->o.preconditioner != SFEM_PC_JACOBI && S->o.preconditioner != SFEM_PC_BLOCK_JACOBI &&
            S->o.preconditioner != SFEM_PC_MULTIGRID)
            fail(SFEM_ERR_ARGUMENT, "server_create: unknown preconditioner");
        S->table_pc = S->o.preconditioner == SFEM_PC_MULTIGRID ? SFEM_PC_BLOCK_JACOBI : S->o.preconditioner;
        S->ctx = ctx;
        S->grid = table_grid(ctx, S->o.precision, S->o.blocks_per_cu);
        S->pool.reset(ctx, (size_t)S->o.max_active);
        S->hpool.assign((size_t)S->o.max_active, ProbState{});
        for (int i = S->o.max_active - 1; i >= 0; --i) S->free_slots.push_back(i);
        SFEM_HIP(hipEventCreate(&S->e0));
        SFEM_HIP(hipEventCreate(&S->e1));
        S->t0 = std::chrono::steady_clock::now();
        if (S->o.preconditioner == SFEM_PC_MULTIGRID) {
            if (S->o.mg_workers < 1 || S->o.mg_workers > 64)
                fail(SFEM_ERR_ARGUMENT, "server_create: mg_workers must lie in [1, 64]");
            for (int i = 0; i < S->o.mg_workers; ++i) {
                std::unique_ptr<Worker> w(new Worker());
                const int cs = sfem_ctx_create(ctx->device, &w->ctx);
                if (cs) fail(cs, "server_create: worker context");
                S->workers.push_back(std::move(w));
            }
            for (auto& w : S->workers) w->th = std::thread(worker_loop, S.get(), w.get());
        }
    });
    if (st) return st;
    ctx_retain(ctx);
    *out = S.release();
    return SFEM_OK;
}

int sfem_server_submit_grid(sfem_server* S, int ndim, const int64_t* dims, const double* node_list, int64_t n_nodes,
                            const int64_t* ien, int64_t n_elem, const double* young, const double* nu,
                            const double* f, const uint8_t* fixed, const double* q_d, int64_t* id) {
    if (!S) return SFEM_ERR_ARGUMENT;
    return guard(S->ctx, [&] {
        if (ndim != 2 && ndim != 3) fail(SFEM_ERR_UNSUPPORTED, "server_submit: Q4 (2D) or Hex8 (3D) only");
        int64_t grid_nodes = 1;
        if (dims)
            for (int d = 0; d < ndim; ++d) {
                if (dims[d] < 1) fail(SFEM_ERR_ARGUMENT, "server_submit: element counts must be >= 1");
                grid_nodes *= dims[d] + 1;
            }
        if (dims && grid_nodes != n_nodes) fail(SFEM_ERR_ARGUMENT, "server_submit: dims do not match n_nodes");
        if (!node_list || !ien || !young || !nu || !fixed || n_nodes <= 0 || n_elem <= 0)
            fail(SFEM_ERR_ARGUMENT, "server_submit: NULL array or empty mesh");
        const int64_t nen = (int64_t)1 << ndim, n = n_nodes * ndim;
        Request r;
        r.id = S->next_id++;
        r.ndim = ndim;
        if (dims) {
            r.grid = true;
            for (int d = 0; d < ndim; ++d) r.dims[d] = dims[d];
        }
        r.n_nodes = n_nodes;
        r.n_elem = n_elem;
        r.nodes.assign(node_list, node_list + ndim * n_nodes);
        r.ien.assign(ien, ien + nen * n_elem);
        r.young.assign(young, young + n_elem);
        r.nu.assign(nu, nu + n_elem);
        if (f) r.f.assign(f, f + n);
        if (q_d) r.qd.assign(q_d, q_d + n);
        r.fixed.assign(fixed, fixed + n);
        estimate_bytes(r, S->o.precision);
        r.t_sub = S->now_ms();
        const int64_t rid = r.id;
        const bool to_grid = r.grid && S->o.preconditioner == SFEM_PC_MULTIGRID;
        std::lock_guard<std::mutex> lk(S->mu);
        S->reqs.emplace(rid, std::move(r));
        (to_grid ? S->gqueue : S->queue).push_back(rid);
        if (id) *id = rid;
    });
}

int sfem_server_submit_spec(sfem_server* S, const sfem_problem_spec* spec, int64_t* id) {
    if (!S) return SFEM_ERR_ARGUMENT;
    return guard(S->ctx, [&] {
        if (!spec) fail(SFEM_ERR_ARGUMENT, "server_submit_spec: NULL spec");
        const sfem_problem_spec& sp = *spec;
        if (sp.ndim != 2 && sp.ndim != 3) fail(SFEM_ERR_UNSUPPORTED, "server_submit: Q4 (2D) or Hex8 (3D) only");
        Request r;
        r.id = S->next_id++;
        r.ndim = sp.ndim;
        r.grid = true;
        r.spec = true;
        r.sp = sp;
        r.n_nodes = 1;
        r.n_elem = 1;
        for (int d = 0; d < sp.ndim; ++d) {
            if (sp.dims[d] < 1) fail(SFEM_ERR_ARGUMENT, "server_submit: element counts must be >= 1");
            r.dims[d] = sp.dims[d];
            r.n_nodes *= sp.dims[d] + 1;
            r.n_elem *= sp.dims[d];
        }
        estimate_bytes(r, S->o.precision);
        r.t_sub = S->now_ms();
        const int64_t rid = r.id;
        const bool to_grid = S->o.preconditioner == SFEM_PC_MULTIGRID;
        std::lock_guard<std::mutex> lk(S->mu);
        S->reqs.emplace(rid, std::move(r));
        (to_grid ? S->gqueue : S->queue).push_back(rid);
        if (id) *id = rid;
    });
}

int sfem_server_submit(sfem_server* S, int ndim, const double* node_list, int64_t n_nodes, const int64_t* ien,
                       int64_t n_elem, const double* young, const double* nu, const double* f, const uint8_t* fixed,
                       const double* q_d, int64_t* id) {
    return sfem_server_submit_grid(S, ndim, nullptr, node_list, n_nodes, ien, n_elem, young, nu, f, fixed, q_d, id);
}

int sfem_server_step(sfem_server* S, int* n_finished, double* launch_ms) {
    if (!S) return SFEM_ERR_ARGUMENT;
    return guard(S->ctx, [&] { step(S, n_finished, launch_ms); });
}

int sfem_server_poll(const sfem_server* S, int64_t id, int* state, sfem_batch_problem_stats* st, double* times3) {
    if (!S) return SFEM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) {
        S->ctx->last_error = "server_poll: unknown request id";
        return SFEM_ERR_ARGUMENT;
    }
    const Request& r = it->second;
    if (state) *state = r.state;
    if (st) *st = r.stats;
    if (times3) {
        times3[0] = r.t_sub;
        times3[1] = r.t_adm;
        times3[2] = r.t_fin;
    }
    if (r.state == SFEM_REQ_FAILED) S->ctx->last_error = r.error;
    return SFEM_OK;
}

int sfem_server_take(sfem_server* S, int64_t id, double* q) {
    if (!S) return SFEM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) {
        S->ctx->last_error = "server_take: unknown request id";
        return SFEM_ERR_ARGUMENT;
    }
    Request& r = it->second;
    if (r.state == SFEM_REQ_FAILED) {
        S->ctx->last_error = r.error;
        S->reqs.erase(it);
        return SFEM_ERR_STATE;
    }
    if (r.state != SFEM_REQ_DONE) {
        S->ctx->last_error = "server_take: request not finished";
        return SFEM_ERR_STATE;
    }
    if (q) std::copy(r.result.begin(), r.result.end(), q);
    S->reqs.erase(it);
    return SFEM_OK;
}

int sfem_server_info(const sfem_server* S, int64_t* queued, int64_t* active, int64_t* bytes) {
    if (!S) return SFEM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    if (queued) *queued = (int64_t)(S->queue.size() + S->gqueue.size());
    if (active) *active = (int64_t)S->active.size() + busy_workers(S);
    if (bytes) *bytes = S->bytes_in_use;
    return SFEM_OK;
}

void sfem_server_destroy(sfem_server* S) {
    if (!S) return;
    sfem_ctx* ctx = S->ctx;
    delete S;
    ctx_release(ctx);
}

}  // extern "C"
