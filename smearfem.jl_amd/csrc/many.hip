// many.hip -- k right-hand sides of one eliminated system solved together:
// an independent FP64 Jacobi-PCG per column, every iteration sharing one pass
// over K_ff (multi-vector SpMV, the k columns of a row in one wavefront).
//
// Reference: the example solves K_free q_f = C'(-K_bar q_d) once per
// prescribed displacement d of its sweep (examples/vector3D.jl:310-322) with
// the same K_free every time; solving all of them at once reads K_free once
// per iteration instead of once per displacement.
//
// Layout: X, R, Z, P, Q, RHS are n_free x KP row-major (KP = the power of two
// >= k, columns >= k stay zero).  Per-column sums are formed per block in a
// fixed order over a fixed grid and then over the blocks in order, so results
// do not depend on timing.  Column states live on the device; the host looks
// at them every check_every iterations.  A column whose recurrence residual
// reaches the tolerance gets an FP64 true residual rhs - K_ff x; it is done if
// that meets the tolerance, otherwise it restarts from the true residual.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sfem_objects.hpp"
#include "solver_common.hpp"

namespace sfem {

namespace {

constexpr int kThreadsM = 256;

enum ColState : int {
    kActive = 0,
    kPending = 1,    // recurrence below tolerance or iteration cap: true residual due
    kRestart = 2,    // true residual above tolerance: restart from it
    kDone = 3,       // converged (true residual)
    kBreakdown = 4,  // p'q <= 0 or not finite
    kCapped = 5,     // iteration cap reached
    kStalled = 6,    // too many restarts without meeting the tolerance
};

struct ColumnState {
    double rz, alpha, beta, tol2, bb, rr_rec, rr_true;
    long long iters, restarts;
    int state, capped;
};

template <int KP>
__device__ __forceinline__ void block_partial(double v, double* sh, double* out) {
    constexpr int RPB = kThreadsM / KP;
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < KP) {
        double s = 0.0;
        for (int r = 0; r < RPB; ++r) s += sh[r * KP + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Q = K_ff P; part[b*KP + j] = sum over block b's rows of P.Q in column j.
template <int KP>
__global__ __launch_bounds__(kThreadsM) void spmm_kernel(const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ col,
                                                         const double* __restrict__ val, int64_t nf, int64_t chunk,
                                                         const double* __restrict__ P, double* __restrict__ Q,
                                                         double* __restrict__ part) {
    const int64_t lb = logical_block();
    constexpr int RPB = kThreadsM / KP;
    __shared__ double sh[kThreadsM];
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    double dot = 0.0;
    for (int64_t row = lb * chunk + rl; row < min(nf, (lb + 1) * chunk); row += RPB) {
        double acc = 0.0;
        const int32_t e1 = rowptr[row + 1];
#pragma unroll 4
        for (int32_t e = rowptr[row]; e < e1; ++e) acc = fma(val[e], P[(int64_t)col[e] * KP + j], acc);
        Q[row * KP + j] = acc;
        dot = fma(P[row * KP + j], acc, dot);
    }
    if (part) block_partial<KP>(dot, sh, part + lb * KP);
}

// Q = K_ff P from K's node blocks (bs = 3).  K is bitwise symmetric, so
// column unit u of its block CSC is node row u: rows 3u+jj of K_ff are
// sum over blocks (r, v) of v[i*3+jj] x_i with x_i = P[free_of[3r+i]] (0 for a
// fixed dof).  Each gathered node row of P serves three output rows, a third
// of the gather traffic of the scalar rows; partially fixed nodes need no
// separate format.  Four blocks' row ids, values and free-dof ids are loaded
// before their gathers (scheduling barriers keep them batched); the
// arithmetic order is that of one block at a time.
template <int KP>
__global__ __launch_bounds__(kThreadsM) void spmm_blk3_kernel(const int64_t* __restrict__ bcolptr,
                                                              const int32_t* __restrict__ browidx,
                                                              const double* __restrict__ bval,
                                                              const int32_t* __restrict__ free_of, int64_t nunits,
                                                              int64_t chunk, const double* __restrict__ P,
                                                              double* __restrict__ Q, double* __restrict__ part) {
    const int64_t lb = logical_block();
    constexpr int UPB = kThreadsM / KP;
    __shared__ double sh[kThreadsM];
    const int j = threadIdx.x % KP, ul = threadIdx.x / KP;
    double dot = 0.0;
    const int64_t u1 = min(nunits, (lb + 1) * chunk);
    for (int64_t u = lb * chunk + ul; u < u1; u += UPB) {
        const int32_t f0 = free_of[3 * u], f1 = free_of[3 * u + 1], f2 = free_of[3 * u + 2];
        if (f0 < 0 && f1 < 0 && f2 < 0) continue;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        const int64_t e1 = bcolptr[u + 1];
        int64_t e = bcolptr[u];
        for (; e + 4 <= e1; e += 4) {
            int64_t r[4];
            double v[4][9];
            int32_t g[4][3];
            double x[4][3];
#pragma unroll
            for (int t = 0; t < 4; ++t) r[t] = browidx[e + t];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int m = 0; m < 9; ++m) v[t][m] = bval[(e + t) * 9 + m];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < 3; ++i) g[t][i] = free_of[3 * r[t] + i];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < 3; ++i) x[t][i] = g[t][i] >= 0 ? P[(int64_t)g[t][i] * KP + j] : 0.0;
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                a0 = fma(v[t][0], x[t][0], fma(v[t][3], x[t][1], fma(v[t][6], x[t][2], a0)));
                a1 = fma(v[t][1], x[t][0], fma(v[t][4], x[t][1], fma(v[t][7], x[t][2], a1)));
                a2 = fma(v[t][2], x[t][0], fma(v[t][5], x[t][1], fma(v[t][8], x[t][2], a2)));
            }
        }
        for (; e < e1; ++e) {
            const int64_t r = browidx[e];
            const double* v = bval + e * 9;
            const int32_t g0 = free_of[3 * r], g1 = free_of[3 * r + 1], g2 = free_of[3 * r + 2];
            const double x0 = g0 >= 0 ? P[(int64_t)g0 * KP + j] : 0.0;
            const double x1 = g1 >= 0 ? P[(int64_t)g1 * KP + j] : 0.0;
            const double x2 = g2 >= 0 ? P[(int64_t)g2 * KP + j] : 0.0;
            a0 = fma(v[0], x0, fma(v[3], x1, fma(v[6], x2, a0)));
            a1 = fma(v[1], x0, fma(v[4], x1, fma(v[7], x2, a1)));
            a2 = fma(v[2], x0, fma(v[5], x1, fma(v[8], x2, a2)));
        }
        if (f0 >= 0) { Q[(int64_t)f0 * KP + j] = a0; dot = fma(P[(int64_t)f0 * KP + j], a0, dot); }
        if (f1 >= 0) { Q[(int64_t)f1 * KP + j] = a1; dot = fma(P[(int64_t)f1 * KP + j], a1, dot); }
        if (f2 >= 0) { Q[(int64_t)f2 * KP + j] = a2; dot = fma(P[(int64_t)f2 * KP + j], a2, dot); }
    }
    if (part) block_partial<KP>(dot, sh, part + lb * KP);
}

// x += a p; r -= a q; z = r / diag; partials of r.z and r.r.
template <int KP>
__global__ __launch_bounds__(kThreadsM) void update_kernel(int64_t nf, int64_t chunk, const ColumnState* __restrict__ cs,
                                                           double* __restrict__ X, double* __restrict__ R,
                                                           double* __restrict__ Z, const double* __restrict__ P,
                                                           const double* __restrict__ Q,
                                                           const double* __restrict__ dinv,
                                                           double* __restrict__ part) {
    const int64_t lb = logical_block();
    constexpr int RPB = kThreadsM / KP;
    __shared__ double sh[kThreadsM];
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    const double a = cs[j].alpha;
    double rz = 0.0, rr = 0.0;
    for (int64_t row = lb * chunk + rl; row < min(nf, (lb + 1) * chunk); row += RPB) {
        const int64_t i = row * KP + j;
        X[i] = fma(a, P[i], X[i]);
        const double r = fma(-a, Q[i], R[i]);
        const double z = dinv[row] * r;
        R[i] = r;
        Z[i] = z;
        rz = fma(r, z, rz);
        rr = fma(r, r, rr);
    }
    block_partial<KP>(rz, sh, part + lb * 2 * KP);
    block_partial<KP>(rr, sh, part + lb * 2 * KP + KP);
}

template <int KP>
__global__ __launch_bounds__(kThreadsM) void p_update_kernel(int64_t nf, int64_t chunk, const ColumnState* __restrict__ cs,
                                                             const double* __restrict__ Z, double* __restrict__ P) {
    const int64_t lb = logical_block();
    constexpr int RPB = kThreadsM / KP;
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    const double b = cs[j].beta;
    for (int64_t row = lb * chunk + rl; row < min(nf, (lb + 1) * chunk); row += RPB) {
        const int64_t i = row * KP + j;
        P[i] = fma(b, P[i], Z[i]);
    }
}

// R = src, Z = R / diag, P = Z for every column (start) or for the columns
// being restarted (only_restart); partials of r.z (those columns) and rhs.rhs.
template <int KP>
__global__ __launch_bounds__(kThreadsM) void init_kernel(int64_t nf, int64_t chunk, const ColumnState* __restrict__ cs, int only_restart,
                                                         const double* __restrict__ src, double* __restrict__ R,
                                                         double* __restrict__ Z, double* __restrict__ P,
                                                         const double* __restrict__ dinv,
                                                         const double* __restrict__ rhs, double* __restrict__ part) {
    const int64_t lb = logical_block();
    constexpr int RPB = kThreadsM / KP;
    __shared__ double sh[kThreadsM];
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    const bool sel = !only_restart || cs[j].state == kRestart;
    double rz = 0.0, bb = 0.0;
    for (int64_t row = lb * chunk + rl; row < min(nf, (lb + 1) * chunk); row += RPB) {
        const int64_t i = row * KP + j;
        if (sel) {
            const double r = src[i];
            const double z = dinv[row] * r;
            R[i] = r;
            Z[i] = z;
            P[i] = z;
            rz = fma(r, z, rz);
        }
        bb = fma(rhs[i], rhs[i], bb);
    }
    block_partial<KP>(rz, sh, part + lb * 2 * KP);
    block_partial<KP>(bb, sh, part + lb * 2 * KP + KP);
}

// T = RHS - Q (true residual of X when Q = K X); partial of T.T.  T may be Q.
template <int KP>
__global__ __launch_bounds__(kThreadsM) void true_resid_kernel(int64_t nf, int64_t chunk, const double* __restrict__ rhs,
                                                               const double* Q, double* T,
                                                               double* __restrict__ part) {
    const int64_t lb = logical_block();
    constexpr int RPB = kThreadsM / KP;
    __shared__ double sh[kThreadsM];
    const int j = threadIdx.x % KP, rl = threadIdx.x / KP;
    double rr = 0.0;
    for (int64_t row = lb * chunk + rl; row < min(nf, (lb + 1) * chunk); row += RPB) {
        const int64_t i = row * KP + j;
        const double t = rhs[i] - Q[i];
        T[i] = t;
        rr = fma(t, t, rr);
    }
    block_partial<KP>(rr, sh, part + lb * KP);
}

// Sum over the blocks' partials of one column by a 256-thread block: thread t
// adds blocks t, t + 256, ... in order, then a fixed tree; valid in thread 0.
__device__ __forceinline__ double sum_blocks(const double* part, int nb, int stride, int j) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[(int64_t)b * stride + j];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---- per-column scalars (one thread per column) --------------------------
__global__ void fin_init_kernel(ColumnState* cs, int k, int kp, const double* part, int nb, double rtol,
                                int first) {
    const int j = blockIdx.x;
    const double rz = sum_blocks(part, nb, 2 * kp, j);
    const double bb = sum_blocks(part, nb, 2 * kp, kp + j);
    if (threadIdx.x != 0) return;
    ColumnState c = cs[j];
    if (first) {
        c = ColumnState{};
        c.bb = bb;
        c.tol2 = rtol * rtol * bb;
        c.rz = rz;
        c.rr_true = bb;
        c.state = (j >= k || bb == 0.0) ? kDone : kActive;
    } else if (c.state == kRestart) {
        c.rz = rz;
        c.restarts += 1;
        c.state = kActive;
    }
    c.alpha = 0.0;
    c.beta = 0.0;
    cs[j] = c;
}

__global__ void fin_alpha_kernel(ColumnState* cs, int kp, const double* part, int nb) {
    const int j = blockIdx.x;
    const double pq = sum_blocks(part, nb, kp, j);
    if (threadIdx.x != 0) return;
    ColumnState c = cs[j];
    c.alpha = 0.0;
    if (c.state == kActive) {
        if (!(pq > 0.0) || !isfinite(pq)) c.state = kBreakdown;
        else c.alpha = c.rz / pq;
    }
    cs[j] = c;
}

__global__ void fin_beta_kernel(ColumnState* cs, int kp, const double* part, int nb, long long max_iter) {
    const int j = blockIdx.x;
    const double rz = sum_blocks(part, nb, 2 * kp, j);
    const double rr = sum_blocks(part, nb, 2 * kp, kp + j);
    if (threadIdx.x != 0) return;
    ColumnState c = cs[j];
    c.beta = 0.0;
    if (c.state == kActive) {
        c.iters += 1;
        c.rr_rec = rr;
        c.beta = rz / c.rz;
        c.rz = rz;
        if (!isfinite(rr)) c.state = kBreakdown;
        else if (rr <= c.tol2) c.state = kPending;
        else if (c.iters >= max_iter) { c.state = kPending; c.capped = 1; }
    }
    cs[j] = c;
}

__global__ void fin_true_kernel(ColumnState* cs, int kp, const double* part, int nb, int max_restarts) {
    const int j = blockIdx.x;
    const double rr = sum_blocks(part, nb, kp, j);
    if (threadIdx.x != 0) return;
    ColumnState c = cs[j];
    if (c.state == kPending) {
        c.rr_true = rr;
        if (rr <= c.tol2) c.state = kDone;
        else if (c.capped) c.state = kCapped;
        else if (c.restarts >= max_restarts) c.state = kStalled;
        else c.state = kRestart;
    }
    cs[j] = c;
}

__global__ void recombine_many_kernel(const int32_t* __restrict__ free_of, const double* __restrict__ X, int kp,
                                      const double* __restrict__ qd, int64_t n, int k, double* __restrict__ q) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    for (int j = 0; j < k; ++j)
        q[(int64_t)j * n + s] = row >= 0 ? X[(int64_t)row * kp + j] : (qd ? qd[(int64_t)j * n + s] : 0.0);
}

template <int KP>
struct ManyRun {
    static void run(sfem_ctx* ctx, sfem_system* S, int k, const double* dqd, const double* df,
                    const sfem_solve_opts& o, double* dq, std::vector<ColumnState>& hcs, double* ms) {
        const int64_t nf = S->nf, n = S->n;
        constexpr int RPB = kThreadsM / KP;
        // 16 blocks per CU keep every CU's resident slots filled through
        // the row chunks; the partials (nb x KP) are reduced by one block
        // per column, whose cost grows with nb
        const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * 16, (nf + RPB - 1) / RPB));
        // contiguous row chunks (a multiple of the rows per block pass); the
        // blocks of one XCD take consecutive chunks (logical_block)
        const int64_t chunk = ((nf + nb - 1) / nb + RPB - 1) / RPB * RPB;
        // node-block SpMV when K is an assembled (bitwise symmetric) 3x3-block matrix
        const sfem_csr* K = S->K;
        const bool blk = K->bs == 3 && !K->transposed;
        if (!blk) ensure_csr(ctx, S);
        const int64_t nunits = K->nunits;
        const int64_t uchunk = ((nunits + nb - 1) / nb + RPB - 1) / RPB * RPB;
        auto spmm = [&](const double* src, double* dst, double* prt) {
            if (blk)
                hipLaunchKernelGGL(spmm_blk3_kernel<KP>, dim3(nb), dim3(kThreadsM), 0, ctx->stream, K->bcolptr.get(),
                                   K->browidx.get(), K->bval.get(), S->free_of.get(), nunits, uchunk, src, dst, prt);
            else
                hipLaunchKernelGGL(spmm_kernel<KP>, dim3(nb), dim3(kThreadsM), 0, ctx->stream, S->rowptr.get(),
                                   S->col.get(), S->val64.get(), nf, chunk, src, dst, prt);
        };
        const size_t nv = (size_t)std::max<int64_t>(nf, 1) * KP;
        DevBuf<double> X(ctx, nv), R(ctx, nv), Z(ctx, nv), P(ctx, nv), Q(ctx, nv), RHS(ctx, nv);
        DevBuf<double> part(ctx, (size_t)nb * 2 * KP);
        DevBuf<ColumnState> cs(ctx, KP);
        memset_d(ctx, X.get(), 0, sizeof(double) * nv);
        launch_rhs_many(ctx, S, k, KP, dqd, df, RHS.get());
        ensure_dinv(ctx, S);
        const double* dinv = S->dinv64.get();
        hipEvent_t e0, e1;
        SFEM_HIP(hipEventCreate(&e0));
        SFEM_HIP(hipEventCreate(&e1));
        SFEM_HIP(hipEventRecord(e0, ctx->stream));
        const dim3 g(nb), b(kThreadsM);
        hipLaunchKernelGGL(init_kernel<KP>, g, b, 0, ctx->stream, nf, chunk, cs.get(), 0, RHS.get(), R.get(), Z.get(),
                           P.get(), dinv, RHS.get(), part.get());
        hipLaunchKernelGGL(fin_init_kernel, dim3(KP), dim3(256), 0, ctx->stream, cs.get(), k, KP, part.get(), nb,
                           o.rtol, 1);
        SFEM_LAUNCH_CHECK();
        const int every = o.check_every > 0 ? o.check_every : 32;
        const long long max_iter = std::max<int64_t>(o.max_iter, 1);
        const int max_restarts = 64;
        const int64_t max_checks = max_iter / every + 2 * max_restarts + 8;
        hcs.assign(KP, ColumnState{});
        for (int64_t check = 0;; ++check) {
            for (int it = 0; it < every; ++it) {
                spmm(P.get(), Q.get(), part.get());
                hipLaunchKernelGGL(fin_alpha_kernel, dim3(KP), dim3(256), 0, ctx->stream, cs.get(), KP, part.get(), nb);
                hipLaunchKernelGGL(update_kernel<KP>, g, b, 0, ctx->stream, nf, chunk, cs.get(), X.get(), R.get(), Z.get(),
                                   P.get(), Q.get(), dinv, part.get());
                hipLaunchKernelGGL(fin_beta_kernel, dim3(KP), dim3(256), 0, ctx->stream, cs.get(), KP, part.get(), nb,
                                   max_iter);
                hipLaunchKernelGGL(p_update_kernel<KP>, g, b, 0, ctx->stream, nf, chunk, cs.get(), Z.get(), P.get());
            }
            SFEM_LAUNCH_CHECK();
            d2h(ctx, hcs.data(), cs.get(), sizeof(ColumnState) * KP);
            bool pending = false, active = false;
            for (int j = 0; j < KP; ++j) {
                pending |= hcs[j].state == kPending;
                active |= hcs[j].state == kActive;
            }
            if (pending) {
                // true residuals of every column; decisions only for pending ones
                spmm(X.get(), Q.get(), nullptr);
                hipLaunchKernelGGL(true_resid_kernel<KP>, g, b, 0, ctx->stream, nf, chunk, RHS.get(), Q.get(), Q.get(),
                                   part.get());
                hipLaunchKernelGGL(fin_true_kernel, dim3(KP), dim3(256), 0, ctx->stream, cs.get(), KP, part.get(), nb,
                                   max_restarts);
                hipLaunchKernelGGL(init_kernel<KP>, g, b, 0, ctx->stream, nf, chunk, cs.get(), 1, Q.get(), R.get(),
                                   Z.get(), P.get(), dinv, RHS.get(), part.get());
                hipLaunchKernelGGL(fin_init_kernel, dim3(KP), dim3(256), 0, ctx->stream, cs.get(), k, KP, part.get(),
                                   nb, o.rtol, 0);
                SFEM_LAUNCH_CHECK();
                d2h(ctx, hcs.data(), cs.get(), sizeof(ColumnState) * KP);
                active = false;
                for (int j = 0; j < KP; ++j) active |= hcs[j].state == kActive;
            }
            if (!active || check >= max_checks) break;
        }
        SFEM_HIP(hipEventRecord(e1, ctx->stream));
        hipLaunchKernelGGL(recombine_many_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                           S->free_of.get(), X.get(), KP, dqd, n, k, dq);
        SFEM_LAUNCH_CHECK();
        sync(ctx);
        float t = 0;
        SFEM_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

int pow2_at_least(int k) {
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

}  // namespace
}  // namespace sfem

using namespace sfem;

extern "C" {

int sfem_system_solve_many(sfem_ctx* ctx, sfem_system* S, int k, const double* q_d, const double* f,
                           const sfem_solve_opts* opts, double* q, sfem_batch_problem_stats* per, double* solve_ms) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S || !q) { ctx->last_error = "solve_many: NULL argument"; return SFEM_ERR_ARGUMENT; }
    sfem_solve_opts o;
    sfem_solve_opts_default(&o);
    if (opts) o = *opts;
    return guard(ctx, [&] {
        if (k < 1 || k > 64) fail(SFEM_ERR_ARGUMENT, "solve_many: 1 <= k <= 64 right-hand sides");
        if (o.precision != SFEM_PREC_FP64) fail(SFEM_ERR_UNSUPPORTED, "solve_many: FP64 only");
        if (o.preconditioner != SFEM_PC_JACOBI) fail(SFEM_ERR_UNSUPPORTED, "solve_many: point Jacobi only");
        if (!(o.rtol >= 0.0)) fail(SFEM_ERR_ARGUMENT, "solve_many: rtol must be >= 0");
        const int64_t n = S->n;
        DevBuf<double> dqd, df, dq(ctx, (size_t)n * k);
        if (q_d) {
            dqd.reset(ctx, (size_t)n * k);
            h2d(ctx, dqd.get(), q_d, sizeof(double) * n * k);
        }
        if (f) {
            df.reset(ctx, (size_t)n * k);
            h2d(ctx, df.get(), f, sizeof(double) * n * k);
        }
        std::vector<ColumnState> hcs;
        double ms = 0;
        const int kp = pow2_at_least(k);
        switch (kp) {
            case 1: ManyRun<1>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            case 2: ManyRun<2>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            case 4: ManyRun<4>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            case 8: ManyRun<8>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            case 16: ManyRun<16>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            case 32: ManyRun<32>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
            default: ManyRun<64>::run(ctx, S, k, dqd.get(), df.get(), o, dq.get(), hcs, &ms); break;
        }
        d2h(ctx, q, dq.get(), sizeof(double) * n * k);
        int64_t bad = 0;
        for (int j = 0; j < k; ++j) {
            const ColumnState& c = hcs[(size_t)j];
            const int reason = c.state == kDone ? 1 : c.state == kBreakdown ? 2 : c.state == kStalled ? 3 : 4;
            if (per) {
                per[j].iterations = c.iters;
                per[j].fp64_updates = c.restarts;
                per[j].rel_residual = c.bb > 0 ? std::sqrt(c.rr_true / c.bb) : 0.0;
                per[j].n_free = S->nf;
                per[j].stop_reason = reason;
                per[j].floor_residual = 0.0;
            }
            if (reason != 1) ++bad;
        }
        if (solve_ms) *solve_ms = ms;
        if (bad) fail(SFEM_ERR_NOT_CONVERGED, "solve_many: " + std::to_string(bad) + " columns did not converge");
    });
}

}  // extern "C"
