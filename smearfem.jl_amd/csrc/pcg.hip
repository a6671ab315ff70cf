// pcg.hip -- Jacobi-preconditioned CG on K_ff for gfx950.
//
// Replaces the reference's dense inverse (smearFEM examples/vector3D.jl:318,
// inv(Matrix(K_free)) * C' * (-K q_d)) with an iterative solve that keeps K_ff
// sparse and resident in HBM.  Three precisions share one driver:
//
//   FP64        : operator, vectors and reductions in FP64.
//   MIXED       : FP32 operator values and FP32 CG vectors; the solution
//                 increment is folded into an FP64 solution.
//   FP32_MATRIX : FP32 operator values, FP64 CG vectors.
// In every mode the true residual rhs - K_ff x is recomputed with the FP64
// operator whenever the recurrence residual has dropped by `reliable_delta`
// since the last replacement, or claims convergence ("reliable updates"), and
// the search direction is re-orthogonalised against the new residual.  In the
// low-precision modes this is what carries the solve to FP64 accuracy; in
// FP64 it costs one extra SpMV per 10x drop and keeps the recurrence honest.
// Every mode stops only on that FP64 residual.
//
// Per iteration four kernels run, with no host round trip:
//   p_update: p = zcoef z + beta p (streaming)
//   spmv_cg : q = K_ff p, partial p.q -> alpha
//   update  : x += alpha p, r -= alpha q, z = D^-1 r, partial r.z, r.r
//             -> beta, replacement flag
//   resid   : exits at once unless a replacement is due (FP64 residual);
//             replacements that stop making progress end the solve
//             ("stagnation": the tolerance is below what FP64 can reach)
// The host reads the 100-byte state every `check_every` iterations.
//
// Reductions are deterministic: wave64 butterfly shuffles, a fixed 4-wave
// block combine, per-block partials, and the last block to arrive sums the
// partials in a fixed order.  The grid is a fixed function of the problem
// size, so results are bitwise reproducible run to run.
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "sfem_objects.hpp"
#include "solver_common.hpp"

namespace sfem {

namespace {

// ---------------------------------------------------------- p_update ---
// p = zcoef*z + beta*p in place (zcoef = 1 except right after a residual
// replacement, where it also makes p orthogonal to the new residual).
template <typename T>
__global__ __launch_bounds__(kThreads) void p_update_kernel(T* __restrict__ p, const T* __restrict__ z, int64_t n,
                                                            int64_t per_block, const CgState* st) {
    if (st->done) return;
    const T beta = (T)st->beta, zc = (T)st->zcoef;
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(n, i0 + per_block);
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) p[i] = fmaT<T>(beta, p[i], zc * z[i]);
}

// ---------------------------------------------------------- spmv_cg ----
// q = K_ff p (accumulated in T), partial p.q -> last block: alpha = rz / pq.
template <int G, typename TV, typename T>
__global__ __launch_bounds__(kThreads) void csr_spmv_cg_kernel(const int32_t* __restrict__ rowptr,
                                                               const int32_t* __restrict__ col,
                                                               const TV* __restrict__ val, int64_t nrows,
                                                               int64_t rows_per_block, const T* __restrict__ p,
                                                               T* __restrict__ q, CgState* st,
                                                               double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const int64_t r0 = logical_block() * rows_per_block;
    const int64_t r1 = min(nrows, r0 + rows_per_block);
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    double dot = 0.0;
    auto gx = [=](int64_t c) { return p[c]; };
    for (int64_t row = r0 + g; row < r1; row += kThreads / G) {
        const T s = csr_row<G, TV, T>(rowptr, col, val, row, lane, gx);
        if (lane == 0) {
            q[row] = s;
            dot += (double)p[row] * (double)s;
        }
    }
    const double bsum = block_sum_256(dot, sh);
    if (threadIdx.x == 0) st_coherent(partials + 3 * blockIdx.x, bsum);
    if (last_block_arrive(&st->arrive)) {
        double pq, u1, u2;
        sum_partials3(partials, gridDim.x, &pq, &u1, &u2, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz / pq;
            if (!(pq > 0.0) || !std::isfinite(st->alpha)) st->done = 2;  // breakdown
        }
    }
}

template <int BS, typename TV, typename T, bool NT>
__global__ __launch_bounds__(kThreads) void sell_spmv_cg_kernel(
    const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w, const int32_t* __restrict__ scol,
    Vals<TV> sval, int64_t nslices, int64_t nrows, int64_t slices_per_block, const T* __restrict__ p,
    T* __restrict__ q, CgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double dot = 0.0;
    auto gx = [=](int64_t c) { return p[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        T out[BS];
        sell_row<BS, TV, T, NT>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                q[d] = out[j];
                dot += (double)p[d] * (double)out[j];
            }
        }
    }
    const double bsum = block_sum_256(dot, sh);
    if (threadIdx.x == 0) st_coherent(partials + 3 * blockIdx.x, bsum);
    if (last_block_arrive(&st->arrive)) {
        double pq, u1, u2;
        sum_partials3(partials, gridDim.x, &pq, &u1, &u2, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz / pq;
            if (!(pq > 0.0) || !std::isfinite(st->alpha)) st->done = 2;
        }
    }
}

// Last block of an update: beta, the replacement triggers and the stopping
// backstops from the partial sums of r.z and r.r.
__device__ __forceinline__ void cg_update_finish(CgState* st, double* partials, double* sh) {
    double trz, trr, u;
    sum_partials3(partials, gridDim.x, &trz, &trr, &u, sh);
    if (threadIdx.x == 0) {
        st->rz_prev = st->rz;
        st->rz = trz;
        st->beta = trz / st->rz_prev;
        st->rr = trr;
        st->iters += 1;
        st->fold_pending = 0;
        st->zcoef = 1.0;
        // replacement when the recurrence converged, dropped by delta, or
        // made no such drop for max(256, iteration of the best true residual)
        const long long since = st->iters - st->last_upd;
        const long long every = st->best_true_it > 256 ? st->best_true_it : 256;
        if (trr <= st->tol2bb || trr < st->delta2 * st->maxrr || since >= every) st->need_update = 1;
        // backstop for a recurrence that stops decreasing: no 2x drop of |r|
        // within max(10000, 4x the iterations it took to reach the best)
        if (trr < 0.25 * st->best_rr) {
            st->best_rr = trr;
            st->best_it = st->iters;
        } else if (st->iters - st->best_it > (4 * st->best_it > 10000 ? 4 * st->best_it : 10000)) {
            st->done = 3;
        }
        if (!std::isfinite(trr)) st->done = 2;
    }
}

// ----------------------------------------------------------- update ----
// x += alpha p, r -= alpha q, z = D^-1 r, partial r.z and r.r.  With FOLD
// (FP32 solution increment xs) a pending fold moves xs into x64 first.  The
// last block sets beta and decides whether the FP64 residual is due.
template <typename T, bool FOLD>
__global__ __launch_bounds__(kThreads) void cg_update_kernel(T* __restrict__ x, double* __restrict__ x64,
                                                             T* __restrict__ r, T* __restrict__ z,
                                                             const T* __restrict__ p, const T* __restrict__ q,
                                                             const T* __restrict__ dinv, int64_t n,
                                                             int64_t per_block, CgState* st,
                                                             double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const T a = (T)st->alpha;
    const bool fold = FOLD && st->fold_pending;
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(n, i0 + per_block);
    double rz = 0.0, rr = 0.0;
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        T xi = x[i];
        if (fold) {
            x64[i] += (double)xi;
            xi = 0;
        }
        const T ri = fmaT<T>(-a, q[i], r[i]);
        const T zi = dinv[i] * ri;
        x[i] = fmaT<T>(a, p[i], xi);
        r[i] = ri;
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += (double)ri * (double)ri;
    }
    const double b0 = block_sum_256(rz, sh);
    const double b1 = block_sum_256(rr, sh + 4);
    if (threadIdx.x == 0) {
        st_coherent(partials + 3 * blockIdx.x, b0);
        st_coherent(partials + 3 * blockIdx.x + 1, b1);
    }
    if (last_block_arrive(&st->arrive)) cg_update_finish(st, partials, sh);
}

// Node-block Jacobi form of cg_update: one thread per node; z = B^-1 r with
// B^-1 the node's inverse diagonal block (planar: binv[(j*BS+i)*nnodes + n]).
template <int BS, typename T, bool FOLD>
__global__ __launch_bounds__(kThreads) void cg_update_block_kernel(T* __restrict__ x, double* __restrict__ x64,
                                                                   T* __restrict__ r, T* __restrict__ z,
                                                                   const T* __restrict__ p, const T* __restrict__ q,
                                                                   const T* __restrict__ binv, int64_t nnodes,
                                                                   int64_t per_block, CgState* st,
                                                                   double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const T a = (T)st->alpha;
    const bool fold = FOLD && st->fold_pending;
    const int64_t n0 = (int64_t)blockIdx.x * per_block;
    const int64_t n1 = min(nnodes, n0 + per_block);
    double rz = 0.0, rr = 0.0;
    for (int64_t n = n0 + threadIdx.x; n < n1; n += kThreads) {
        T rv[BS];
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            const int64_t i = n * BS + j;
            T xi = x[i];
            if (fold) {
                x64[i] += (double)xi;
                xi = 0;
            }
            x[i] = fmaT<T>(a, p[i], xi);
            rv[j] = fmaT<T>(-a, q[i], r[i]);
            r[i] = rv[j];
            rr += (double)rv[j] * (double)rv[j];
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            T zj = 0;
#pragma unroll
            for (int i = 0; i < BS; ++i) zj = fmaT<T>(binv[(j * BS + i) * nnodes + n], rv[i], zj);
            z[n * BS + j] = zj;
            rz += (double)rv[j] * (double)zj;
        }
    }
    const double b0 = block_sum_256(rz, sh);
    const double b1 = block_sum_256(rr, sh + 4);
    if (threadIdx.x == 0) {
        st_coherent(partials + 3 * blockIdx.x, b0);
        st_coherent(partials + 3 * blockIdx.x + 1, b1);
    }
    if (last_block_arrive(&st->arrive)) cg_update_finish(st, partials, sh);
}

// ------------------------------------------- FP64 residual replacement --
// Runs after every update; exits at once unless the FP64 residual is due.
// r = rhs - K_ff x with x = x64 (+ xs, the FP32 increment, when FOLD), in
// FP64; writes r and z = D^-1 r in the CG precision.  The search direction
// is made orthogonal to the new residual (p <- p - gamma z, gamma = r.p/r.z),
// folded into zcoef of the next spmv_cg.
__global__ __launch_bounds__(kThreads) void fold_kernel(double* __restrict__ x64, float* __restrict__ xs, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        x64[i] += (double)xs[i];
        xs[i] = 0.0f;
    }
}

template <bool FOLD>
__device__ __forceinline__ void resid_finish(CgState* st, double* partials, double* sh) {
    double sums[4];
    sum_partials4(partials, gridDim.x, sums, sh);
    const double trr = sums[0], trz = sums[1], trp = sums[2];
    // 16x the attainable |r|^2, (eps * || |rhs| + |K||x| ||)^2
    const double floor_rr = 16.0 * kEps64 * kEps64 * sums[3];
    if (threadIdx.x == 0) {
        const double rec = st->rr;  // recurrence |r|^2 that triggered this replacement
        st->rr_true = trr;
        st->floor_rr = floor_rr;
        st->rr = trr;
        st->last_upd = st->iters;
        st->updates += 1;
        st->need_update = 0;
        st->fold_pending = FOLD ? 1 : 0;
        if (trr <= st->tol2bb) {
            st->done = 1;
        } else {
            const double beta = trz / st->rz_prev;
            st->beta = beta;
            st->zcoef = 1.0 - beta * (trp / trz);
            st->rz = trz;
            st->maxrr = trr;
            // Attainable-floor test.  A replacement whose true residual does
            // not beat the best by 10% ends the solve when it is within 4x of
            // the rounding floor eps*|| |rhs| + |K||x| ||, and counts once
            // when the recurrence ran 5x below it (the residual gap of finite
            // precision; twice in a row ends the solve).  Slow progress away
            // from the floor never stops it: the CG residual norm can plateau
            // for thousands of iterations.
            const bool gap = st->iters > 0 && 25.0 * rec < trr;
            if (trr <= 0.81 * st->best_true) {
                st->stall = 0;
                st->best_true_it = st->iters;
            } else if (trr <= floor_rr) {
                st->stall += 2;
            } else if (gap) {
                st->stall += 1;
            }
            if (st->stall >= 2) st->done = 3;
        }
        if (trr < st->best_true) st->best_true = trr;
        if (!std::isfinite(trr)) st->done = 2;
    }
}

template <int G, typename T, bool FOLD>
__global__ __launch_bounds__(kThreads) void csr_resid_kernel(const int32_t* __restrict__ rowptr,
                                                             const int32_t* __restrict__ col,
                                                             const double* __restrict__ val, int64_t nrows,
                                                             int64_t rows_per_block, const double* __restrict__ x64,
                                                             const T* __restrict__ xs, const double* __restrict__ rhs,
                                                             const T* __restrict__ dinv, const T* __restrict__ p,
                                                             T* __restrict__ r, T* __restrict__ z, CgState* st,
                                                             double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done || !st->need_update) return;
    const int64_t r0 = logical_block() * rows_per_block;
    const int64_t r1 = min(nrows, r0 + rows_per_block);
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    double rr = 0.0, rz = 0.0, rp = 0.0, ff = 0.0;
    auto gx = [=](int64_t c) { return FOLD ? x64[c] + (double)xs[c] : x64[c]; };
    for (int64_t row = r0 + g; row < r1; row += kThreads / G) {
        double a;
        const double s = csr_row_abs<G>(rowptr, col, val, row, lane, gx, &a);
        if (lane == 0) {
            const double res = rhs[row] - s;
            const double sc = fabs(rhs[row]) + a;
            ff += sc * sc;
            const T rt = (T)res;
            const T zt = dinv[row] * rt;
            r[row] = rt;
            z[row] = zt;
            rr += res * res;
            rz += (double)rt * (double)zt;
            rp += (double)rt * (double)p[row];
        }
    }
    const double b0 = block_sum_256(rr, sh);
    const double b1 = block_sum_256(rz, sh + 4);
    const double b2 = block_sum_256(rp, sh);
    const double b3 = block_sum_256(ff, sh + 4);
    if (threadIdx.x == 0) {
        st_coherent(partials + 4 * blockIdx.x, b0);
        st_coherent(partials + 4 * blockIdx.x + 1, b1);
        st_coherent(partials + 4 * blockIdx.x + 2, b2);
        st_coherent(partials + 4 * blockIdx.x + 3, b3);
    }
    if (last_block_arrive(&st->arrive)) resid_finish<FOLD>(st, partials, sh);
}

// BPC: z = B^-1 r with the node's inverse diagonal block (binv, planar with
// stride nrows) instead of the point Jacobi dinv.
template <int BS, typename T, bool FOLD, bool BPC>
__global__ __launch_bounds__(kThreads) void sell_resid_kernel(
    const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w, const int32_t* __restrict__ scol,
    const double* __restrict__ sval, int64_t nslices, int64_t nrows, int64_t slices_per_block,
    const double* __restrict__ x64, const T* __restrict__ xs, const double* __restrict__ rhs,
    const T* __restrict__ dinv, const T* __restrict__ binv, const T* __restrict__ p, T* __restrict__ r,
    T* __restrict__ z, CgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done || !st->need_update) return;
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double rr = 0.0, rz = 0.0, rp = 0.0, ff = 0.0;
    auto gx = [=](int64_t c) { return FOLD ? x64[c] + (double)xs[c] : x64[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        double out[BS], aout[BS];
        sell_row_abs<BS>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out, aout);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
            T rt[BS];
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                const double res = rhs[d] - out[j];
                const double sc = fabs(rhs[d]) + aout[j];
                ff += sc * sc;
                rt[j] = (T)res;
                r[d] = rt[j];
                rr += res * res;
                rp += (double)rt[j] * (double)p[d];
            }
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                T zt;
                if constexpr (BPC) {
                    zt = 0;
#pragma unroll
                    for (int i = 0; i < BS; ++i) zt = fmaT<T>(binv[(j * BS + i) * nrows + row], rt[i], zt);
                } else {
                    zt = dinv[d] * rt[j];
                }
                z[d] = zt;
                rz += (double)rt[j] * (double)zt;
            }
        }
    }
    const double b0 = block_sum_256(rr, sh);
    const double b1 = block_sum_256(rz, sh + 4);
    const double b2 = block_sum_256(rp, sh);
    const double b3 = block_sum_256(ff, sh + 4);
    if (threadIdx.x == 0) {
        st_coherent(partials + 4 * blockIdx.x, b0);
        st_coherent(partials + 4 * blockIdx.x + 1, b1);
        st_coherent(partials + 4 * blockIdx.x + 2, b2);
        st_coherent(partials + 4 * blockIdx.x + 3, b3);
    }
    if (last_block_arrive(&st->arrive)) resid_finish<FOLD>(st, partials, sh);
}

// bb = rhs.rhs (FP64), once per solve
__global__ __launch_bounds__(kThreads) void norm2_kernel(const double* __restrict__ v, int64_t n, int64_t per_block,
                                                         CgState* st, double* __restrict__ partials, double rtol,
                                                         double delta) {
    __shared__ double sh[8];
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(n, i0 + per_block);
    double s = 0.0;
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s += v[i] * v[i];
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) {
        st_coherent(partials + 3 * blockIdx.x, b);
        st_coherent(partials + 3 * blockIdx.x + 1, 0.0);
        st_coherent(partials + 3 * blockIdx.x + 2, 0.0);
    }
    if (last_block_arrive(&st->arrive)) {
        double bb, u1, u2;
        sum_partials3(partials, gridDim.x, &bb, &u1, &u2, sh);
        if (threadIdx.x == 0) {
            st->bb = bb;
            st->tol2bb = rtol * rtol * bb;
            st->delta2 = delta * delta;
            st->maxrr = bb;
        }
    }
}

// --------------------------------------------------- plain y = A x -----
template <int G, typename TV, typename T>
__global__ __launch_bounds__(kThreads) void csr_spmv_kernel(const int32_t* __restrict__ rowptr,
                                                            const int32_t* __restrict__ col,
                                                            const TV* __restrict__ val, int64_t nrows,
                                                            int64_t rows_per_block, const T* __restrict__ x,
                                                            T* __restrict__ y) {
    const int64_t r0 = logical_block() * rows_per_block;
    const int64_t r1 = min(nrows, r0 + rows_per_block);
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    auto gx = [=](int64_t c) { return x[c]; };
    for (int64_t row = r0 + g; row < r1; row += kThreads / G) {
        const T s = csr_row<G, TV, T>(rowptr, col, val, row, lane, gx);
        if (lane == 0) y[row] = s;
    }
}

template <int BS, typename TV, typename T, bool NT>
__global__ __launch_bounds__(kThreads) void sell_spmv_kernel(const int64_t* __restrict__ sl_ptr,
                                                             const int32_t* __restrict__ sl_w,
                                                             const int32_t* __restrict__ scol, Vals<TV> sval,
                                                             int64_t nslices,
                                                             int64_t nrows, int64_t slices_per_block,
                                                             const T* __restrict__ x, T* __restrict__ y) {
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto gx = [=](int64_t c) { return x[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        T out[BS];
        sell_row<BS, TV, T, NT>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < nrows)
#pragma unroll
            for (int j = 0; j < BS; ++j) y[row * BS + j] = out[j];
    }
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// Streaming copy used as the measured HBM ceiling: UNROLL 16-byte loads in
// flight per thread before the stores, optionally non-temporal.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int UNROLL, bool NT>
__global__ __launch_bounds__(kThreads) void copy_kernel(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kThreads * UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * kThreads * UNROLL + threadIdx.x; base < n; base += stride) {
        v4f v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t i = base + (int64_t)u * kThreads;
            if (i < n) v[u] = ld_stream<NT>(a + i);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t i = base + (int64_t)u * kThreads;
            if (i < n) {
                if constexpr (NT) __builtin_nontemporal_store(v[u], b + i);
                else b[i] = v[u];
            }
        }
    }
}

// --------------------------------------------------------- host side ---
constexpr int kCsrLanes = 16;  // lanes per CSR row (rows of K_ff hold up to 81 entries)

struct Launch {
    unsigned grid;
    int64_t per_block;  // rows / slices / elements per block
};

// grid: a multiple of 8 (XCD mapping), at most `cap` blocks, each block a
// contiguous range that is a multiple of `quantum` items
Launch plan(int64_t items, int quantum, unsigned cap) {
    int64_t need = (items + quantum - 1) / quantum;
    int64_t grid = std::min<int64_t>(need, cap);
    grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
    int64_t per = (items + grid - 1) / grid;
    per = (per + quantum - 1) / quantum * quantum;
    if (per == 0) per = quantum;
    return {(unsigned)grid, per};
}

// Sliced-operator launches: `wave_slices` > 0 gives every wavefront that
// many slices (grid = slices / (4 * wave_slices)); 0 spreads them over the cap.
Launch plan_sell(int64_t nslices, unsigned cap, int wave_slices) {
    if (wave_slices <= 0) return plan(nslices, kThreads / 64, cap);
    const int64_t per = (int64_t)wave_slices * (kThreads / 64);
    int64_t grid = (nslices + per - 1) / per;
    grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
    return {(unsigned)grid, per};
}

struct Ops {
    sfem_ctx* ctx;
    sfem_system* S;
    int fmt;  // SFEM_FMT_CSR / SFEM_FMT_BSR
    unsigned cap;
    Launch csr, sell, vec;
    bool nt;        // non-temporal matrix loads in the sliced SpMV
    bool block_pc;  // node-block Jacobi preconditioner
    Launch node;    // one thread per node (block update)
    // vectors of the solve: the free dofs, or the padded block rows when
    // the node-block format carries partially fixed nodes
    double* x64;
    const double* rhs;
    int64_t n;
};

template <typename TV>
const TV* csr_vals(const sfem_system* S);
template <>
const double* csr_vals<double>(const sfem_system* S) { return S->val64.get(); }
template <>
const float* csr_vals<float>(const sfem_system* S) { return S->val32.get(); }
template <typename TV>
Vals<TV> sell_vals(const sfem_system* S);
template <>
Vals<double> sell_vals<double>(const sfem_system* S) { return {S->sval64.get()}; }
template <>
Vals<float> sell_vals<float>(const sfem_system* S) { return {S->sval32.get()}; }
template <>
Vals<SplitVal> sell_vals<SplitVal>(const sfem_system* S) { return {S->sval32.get(), S->sval16.get()}; }

// bytes per stored operator value
template <typename TV>
constexpr double val_bytes() { return std::is_same<TV, SplitVal>::value ? 6.0 : (double)sizeof(TV); }

template <typename TV, typename T>
void launch_spmv_cg(const Ops& o, const T* p, T* q, CgState* st, double* partials) {
    sfem_system* S = o.S;
    hipStream_t s = o.ctx->stream;
    if (o.fmt == SFEM_FMT_BSR) {
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(o.sell.grid), dim3(kThreads), 0, s, S->sl_ptr.get(), S->sl_w.get(),
                               S->scol.get(), sell_vals<TV>(S), S->nslices, S->nfb, o.sell.per_block, p, q, st,
                               partials);
        };
        if (S->bs == 3) {
            if (o.nt) go(sell_spmv_cg_kernel<3, TV, T, true>);
            else go(sell_spmv_cg_kernel<3, TV, T, false>);
        } else {
            if (o.nt) go(sell_spmv_cg_kernel<2, TV, T, true>);
            else go(sell_spmv_cg_kernel<2, TV, T, false>);
        }
    } else if constexpr (!std::is_same<TV, SplitVal>::value) {
        hipLaunchKernelGGL((csr_spmv_cg_kernel<kCsrLanes, TV, T>), dim3(o.csr.grid), dim3(kThreads), 0, s,
                           S->rowptr.get(), S->col.get(), csr_vals<TV>(S), S->nf, o.csr.per_block, p, q, st,
                           partials);
    } else {
        fail(SFEM_ERR_UNSUPPORTED, "split precision needs the node-block format");
    }
    SFEM_LAUNCH_CHECK();
}

template <typename T, bool FOLD>
void launch_resid(const Ops& o, const T* xs, const T* dinv, const T* binv, const T* p, T* r, T* z, CgState* st,
                  double* partials) {
    sfem_system* S = o.S;
    hipStream_t s = o.ctx->stream;
    if (o.fmt == SFEM_FMT_BSR) {
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(o.sell.grid), dim3(kThreads), 0, s, S->sl_ptr.get(), S->sl_w.get(),
                               S->scol.get(), S->sval64.get(), S->nslices, S->nfb, o.sell.per_block, o.x64, xs,
                               o.rhs, dinv, binv, p, r, z, st, partials);
        };
        if (S->bs == 3) {
            if (o.block_pc) go(sell_resid_kernel<3, T, FOLD, true>);
            else go(sell_resid_kernel<3, T, FOLD, false>);
        } else {
            if (o.block_pc) go(sell_resid_kernel<2, T, FOLD, true>);
            else go(sell_resid_kernel<2, T, FOLD, false>);
        }
    } else {
        hipLaunchKernelGGL((csr_resid_kernel<kCsrLanes, T, FOLD>), dim3(o.csr.grid), dim3(kThreads), 0, s,
                           S->rowptr.get(), S->col.get(), S->val64.get(), S->nf, o.csr.per_block, S->x.get(), xs,
                           S->rhs.get(), dinv, p, r, z, st, partials);
    }
    SFEM_LAUNCH_CHECK();
}

double spmv_bytes_for(const sfem_system* S, int fmt, double sv, double sx, int nvec) {
    const int64_t nv = (fmt == SFEM_FMT_BSR && S->partial) ? S->nfb * S->bs : S->nf;
    const double vec = (double)nvec * (double)nv * sx;
    if (fmt == SFEM_FMT_BSR)
        return (double)S->nnzb * (S->bs * S->bs * sv + 4.0) + (double)(S->nslices + 1) * 8.0 +
               (double)S->nslices * 4.0 + vec;
    return (double)S->nnz * (sv + 4.0) + (double)(S->nf + 1) * 4.0 + vec;
}

void ensure_val32(sfem_ctx* ctx, sfem_system* S) {
    ensure_csr(ctx, S);
    if (S->val32.size() == (size_t)S->nnz || S->nnz == 0) return;
    S->val32.reset(ctx, (size_t)S->nnz);
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(2048), dim3(256), 0, ctx->stream, S->val64.get(), S->val32.get(),
                       S->nnz);
    SFEM_LAUNCH_CHECK();
}

int resolve_format(const sfem_system* S, int fmt) {
    if (fmt == SFEM_FMT_BSR && !S->has_block) return SFEM_FMT_CSR;
    if (fmt == SFEM_FMT_AUTO) return S->has_block ? SFEM_FMT_BSR : SFEM_FMT_CSR;
    return fmt;
}

// Device workspace of one solve when the caller provides it (the solve that
// allocates nothing): four CG vectors, the FP32 solution increment, the
// per-block partial sums and the state.
struct CgWork {
    void *r, *z, *q, *p, *xs;
    double* partials;
    size_t max_partials;  // doubles
    CgState* st;
};

// Zeroes n elements of a vector (a kernel, not hipMemsetAsync, so that the
// solve that allocates nothing issues kernels only).
template <typename T>
__global__ __launch_bounds__(kThreads) void zero_kernel(T* __restrict__ v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        v[i] = T(0);
}

__global__ void cg_state_init_kernel(CgState* st) {
    if (threadIdx.x != 0) return;
    CgState init{};
    init.rr = HUGE_VAL;
    init.rz_prev = HUGE_VAL;
    init.zcoef = 1.0;
    init.best_rr = HUGE_VAL;
    init.best_true = HUGE_VAL;
    init.need_update = 1;
    *st = init;
}

// The iteration cap was reached: the true residual of the final iterate is due.
__global__ void cg_state_flag_update_kernel(CgState* st) {
    if (threadIdx.x == 0) st->need_update = 1;
}

template <typename T>
void zero_dev(sfem_ctx* ctx, T* v, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_kernel<T>, dim3(std::min<unsigned>(grid_for(n, kThreads), 2048)), dim3(kThreads), 0,
                       ctx->stream, v, n);
    SFEM_LAUNCH_CHECK();
}

// TV: operator value type, T: CG vector type.  FOLD (T == float): the
// solution increment lives in FP32 and is folded into the FP64 solution at
// every residual replacement.  ws: the caller's workspace (no device
// allocation, no events, no host-to-device copies), or NULL.
template <typename TV, typename T>
void run_pcg(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* stats, int fmt,
             const CgWork* ws = nullptr) {
    constexpr bool FOLD = std::is_same<T, float>::value;
    const bool padded = fmt == SFEM_FMT_BSR && S->partial;
    const int64_t nf = padded ? S->nfb * S->bs : S->nf;  // length of the solve's vectors
    const bool block_pc = o->preconditioner == SFEM_PC_BLOCK_JACOBI;
    if (block_pc && fmt != SFEM_FMT_BSR)
        fail(SFEM_ERR_UNSUPPORTED, "block-Jacobi preconditioning needs the node-block format");
    if (o->preconditioner != SFEM_PC_JACOBI && !block_pc) fail(SFEM_ERR_ARGUMENT, "solve: unknown preconditioner");
    double* x64 = padded ? S->px.get() : S->x.get();
    Ops ops{ctx, S, fmt, (unsigned)std::max(8, ctx->n_cu * 8), {}, {}, {}, ctx->tune_nt, block_pc, {},
            x64, padded ? S->prhs.get() : S->rhs.get(), nf};
    if (block_pc) ops.node = plan(S->nfb, kThreads, std::max(8, ctx->n_cu * 4));
    ops.csr = plan(S->nf, kThreads / kCsrLanes, ops.cap);
    if (S->has_block) ops.sell = plan_sell(S->nslices, ops.cap, ctx->tune_wave_slices);
    ops.vec = plan(nf, kThreads, std::max(8, ctx->n_cu * 4));
    const unsigned maxgrid = std::max({ops.csr.grid, ops.sell.grid, ops.vec.grid, ops.node.grid});

    DevBuf<T> rb, zb, qb, pb, xsb;
    DevBuf<double> partb;
    DevBuf<CgState> dst;
    T *r, *z, *q, *p, *xsv = nullptr;
    double* partials;
    CgState* st;
    if (ws) {
        if (4 * (size_t)maxgrid > ws->max_partials) fail(SFEM_ERR_UNSUPPORTED, "solve: workspace too small");
        r = (T*)ws->r, z = (T*)ws->z, q = (T*)ws->q, p = (T*)ws->p;
        if (FOLD) xsv = (T*)ws->xs;
        partials = ws->partials;
        st = ws->st;
    } else {
        rb.reset(ctx, nf), zb.reset(ctx, nf), qb.reset(ctx, nf), pb.reset(ctx, nf);
        r = rb.get(), z = zb.get(), q = qb.get(), p = pb.get();
        if (FOLD) {
            xsb.reset(ctx, nf);
            xsv = xsb.get();
        }
        partb.reset(ctx, 4 * (size_t)maxgrid);
        partials = partb.get();
        dst.reset(ctx, 1);
        st = dst.get();
    }
    const T* dinv = padded ? (FOLD ? (const T*)S->pdinv32.get() : (const T*)S->pdinv64.get())
                           : (FOLD ? (const T*)S->dinv32.get() : (const T*)S->dinv64.get());
    const T* binv = FOLD ? (const T*)S->binv32.get() : (const T*)S->binv64.get();
    T* x = (T*)x64;  // FP64 paths update the FP64 solution directly
    if (FOLD) {
        zero_dev(ctx, xsv, nf);
        x = xsv;
    }
    const T* xs = FOLD ? xsv : nullptr;
    if (o->x0_zero) zero_dev(ctx, S->x.get(), S->nf);
    if (padded) pad_gather(ctx, S, S->x.get(), x64);  // start from the current solution
    zero_dev(ctx, p, nf);
    hipLaunchKernelGGL(cg_state_init_kernel, dim3(1), dim3(64), 0, ctx->stream, st);
    SFEM_LAUNCH_CHECK();

    const bool timed_run = ws == nullptr;  // events are runtime objects: none in the no-allocation solve
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timed_run) {
        SFEM_HIP(hipEventCreate(&t0));
        SFEM_HIP(hipEventCreate(&t1));
    }
    const auto w0 = std::chrono::steady_clock::now();
    const int check = std::max(1, o->check_every);
    std::vector<hipEvent_t> ev;
    if (o->time_kernels && timed_run) {
        ev.resize(2 * (size_t)check);
        for (auto& e : ev) SFEM_HIP(hipEventCreate(&e));
    }
    if (timed_run) SFEM_HIP(hipEventRecord(t0, ctx->stream));
    hipLaunchKernelGGL(norm2_kernel, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, ops.rhs, nf,
                       ops.vec.per_block, st, partials, o->rtol, o->reliable_delta);
    SFEM_LAUNCH_CHECK();
    launch_resid<T, FOLD>(ops, xs, dinv, binv, p, r, z, st, partials);

    CgState host{};
    int64_t enq = 0, counted = 0, timed = 0;
    double spmv_ms = 0.0;
    d2h(ctx, &host, st, sizeof(host));
    while (!host.done && enq < o->max_iter) {
        const int64_t chunk = std::min<int64_t>(check, o->max_iter - enq);
        for (int64_t k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(p_update_kernel<T>, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, p, z, nf,
                               ops.vec.per_block, st);
            SFEM_LAUNCH_CHECK();
            if (!ev.empty()) SFEM_HIP(hipEventRecord(ev[2 * k], ctx->stream));
            launch_spmv_cg<TV, T>(ops, p, q, st, partials);
            if (!ev.empty()) SFEM_HIP(hipEventRecord(ev[2 * k + 1], ctx->stream));
            if (!block_pc)
                hipLaunchKernelGGL((cg_update_kernel<T, FOLD>), dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream,
                                   x, x64, r, z, p, q, dinv, nf, ops.vec.per_block, st, partials);
            else if (S->bs == 3)
                hipLaunchKernelGGL((cg_update_block_kernel<3, T, FOLD>), dim3(ops.node.grid), dim3(kThreads), 0,
                                   ctx->stream, x, x64, r, z, p, q, binv, S->nfb, ops.node.per_block, st, partials);
            else
                hipLaunchKernelGGL((cg_update_block_kernel<2, T, FOLD>), dim3(ops.node.grid), dim3(kThreads), 0,
                                   ctx->stream, x, x64, r, z, p, q, binv, S->nfb, ops.node.per_block, st, partials);
            SFEM_LAUNCH_CHECK();
            launch_resid<T, FOLD>(ops, xs, dinv, binv, p, r, z, st, partials);
        }
        enq += chunk;
        d2h(ctx, &host, st, sizeof(host));
        if (!ev.empty()) {
            // launches after the stop exit at once; time only the real ones
            const int64_t real = std::min<int64_t>(chunk, std::max<int64_t>(0, host.iters - counted));
            for (int64_t k = 0; k < real; ++k) {
                float ms = 0;
                SFEM_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                spmv_ms += ms;
            }
            timed += real;
        }
        counted = host.iters;
    }
    if (!host.done) {  // iteration cap: evaluate the true residual of the final iterate
        hipLaunchKernelGGL(cg_state_flag_update_kernel, dim3(1), dim3(64), 0, ctx->stream, st);
        SFEM_LAUNCH_CHECK();
        launch_resid<T, FOLD>(ops, xs, dinv, binv, p, r, z, st, partials);
        d2h(ctx, &host, st, sizeof(host));
    }
    if (FOLD) {  // x = x64 + xs
        hipLaunchKernelGGL(fold_kernel, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, x64, (float*)xsv, nf);
        SFEM_LAUNCH_CHECK();
    }
    if (padded) pad_scatter(ctx, S, x64, S->x.get());
    float total = 0;
    if (timed_run) {
        SFEM_HIP(hipEventRecord(t1, ctx->stream));
        sync(ctx);
        SFEM_HIP(hipEventElapsedTime(&total, t0, t1));
        for (auto& e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    } else {
        sync(ctx);
        total = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }

    stats->iterations = host.iters;
    stats->fp64_updates = host.updates;
    stats->spmv_launches = host.iters;
    stats->spmv64_launches = host.updates;
    stats->rel_residual = host.bb > 0 ? std::sqrt(host.rr_true / host.bb) : 0.0;
    stats->solve_ms = total;
    stats->spmv_ms = spmv_ms;
    stats->spmv_timed = timed;
    // spmv_cg: matrix once, p read once, q written
    stats->spmv_bytes = spmv_bytes_for(S, fmt, val_bytes<TV>(), sizeof(T), 2);
    stats->converged = host.done == 1 ? 1 : 0;
    stats->format = fmt;
    stats->stop_reason = host.done == 0 ? 4 : host.done;
    stats->floor_residual = host.bb > 0 ? std::sqrt(host.floor_rr / 16.0 / host.bb) : 0.0;
}

}  // namespace

// ---------------------------------------------- the no-allocation solve --
// Workspace of sfem_system_solve_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Layout: five CG vectors of
// kNoAllocMaxDofs doubles (r, z, q, p, FP32 increment), the per-block
// partial sums, the CG state.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (sfem_ctx::scratch).
namespace {
constexpr size_t kNoAllocVec = sizeof(double) * (size_t)kNoAllocMaxDofs;
constexpr size_t kNoAllocPartials = 4 * 16384;  // doubles: 4 per block of the largest plan
constexpr size_t kNoAllocBytes = 5 * kNoAllocVec + sizeof(double) * kNoAllocPartials + (sizeof(CgState) + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
}  // namespace

void* noalloc_scratch_address() {
    void* ptr = nullptr;
    SFEM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    return ptr;
}

size_t noalloc_scratch_bytes() { return kNoAllocBytes; }

void system_solve_noalloc(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st) {
    *st = sfem_solve_stats{};
    if (!ctx->scratch) fail(SFEM_ERR_STATE, "solve_noalloc: no workspace address for this device");
    if (!(o->rtol >= 0.0) || o->max_iter < 0) fail(SFEM_ERR_ARGUMENT, "solve: bad rtol or max_iter");
    if (o->preconditioner == SFEM_PC_MULTIGRID)
        fail(SFEM_ERR_UNSUPPORTED, "solve_noalloc: point or block Jacobi only (the multigrid V-cycle allocates)");
    const int fmt = resolve_format(S, o->format);
    const int64_t nv = (fmt == SFEM_FMT_BSR && S->partial) ? S->nfb * S->bs : S->nf;
    if (nv > kNoAllocMaxDofs)
        fail(SFEM_ERR_UNSUPPORTED, "solve_noalloc: " + std::to_string(nv) + " solver dofs, at most " +
                                       std::to_string(kNoAllocMaxDofs));
    if (S->nf == 0) {
        st->converged = 1;
        st->stop_reason = 1;
        return;
    }
    // the operator formats are built by the first allocating solve; this one
    // only reads them
    const bool f32 = o->precision == SFEM_PREC_MIXED || o->precision == SFEM_PREC_FP32_MATRIX;
    const bool ready = S->dinv64.size() &&
                       (fmt == SFEM_FMT_BSR ? S->sell_ready
                                            : (S->col.size() == (size_t)S->nnz && (!f32 || S->val32.size())));
    if (!ready)
        fail(SFEM_ERR_STATE, "solve_noalloc: the operator format these options read is not built yet; one "
                             "sfem_system_solve with the same precision and format builds it");
    unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
    CgWork ws{base, base + kNoAllocVec, base + 2 * kNoAllocVec, base + 3 * kNoAllocVec, base + 4 * kNoAllocVec,
              reinterpret_cast<double*>(base + 5 * kNoAllocVec), kNoAllocPartials,
              reinterpret_cast<CgState*>(base + 5 * kNoAllocVec + sizeof(double) * kNoAllocPartials)};
    sfem_solve_opts oo = *o;
    oo.time_kernels = 0;
    switch (o->precision) {
        case SFEM_PREC_FP64: run_pcg<double, double>(ctx, S, &oo, st, fmt, &ws); break;
        case SFEM_PREC_MIXED: run_pcg<float, float>(ctx, S, &oo, st, fmt, &ws); break;
        case SFEM_PREC_FP32_MATRIX: run_pcg<float, double>(ctx, S, &oo, st, fmt, &ws); break;
        case SFEM_PREC_SPLIT:
            if (fmt != SFEM_FMT_BSR) fail(SFEM_ERR_UNSUPPORTED, "split precision needs the node-block format");
            run_pcg<SplitVal, double>(ctx, S, &oo, st, fmt, &ws);
            break;
        default: fail(SFEM_ERR_ARGUMENT, "solve: unknown precision");
    }
    if (!st->converged) {
        static const char* why[] = {"", "", "breakdown", "stagnation", "iteration cap reached"};
        char buf[200];
        std::snprintf(buf, sizeof buf, "solve: %s after %lld iterations (relative residual %.3e)",
                      why[st->stop_reason], (long long)st->iterations, st->rel_residual);
        fail(SFEM_ERR_NOT_CONVERGED, buf);
    }
}

void system_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st) {
    *st = sfem_solve_stats{};
    if (S->nf == 0) {
        st->converged = 1;
        st->stop_reason = 1;
        return;
    }
    if (!(o->rtol >= 0.0) || o->max_iter < 0) fail(SFEM_ERR_ARGUMENT, "solve: bad rtol or max_iter");
    const int fmt = resolve_format(S, o->format);
    if (fmt == SFEM_FMT_CSR) ensure_csr(ctx, S);
    if (o->preconditioner != SFEM_PC_MULTIGRID) {  // multigrid builds what its levels read
        if (fmt == SFEM_FMT_BSR) ensure_sell(ctx, S);
        ensure_dinv(ctx, S);
    }
    if (o->preconditioner == SFEM_PC_MULTIGRID) {
        if (fmt != SFEM_FMT_BSR) fail(SFEM_ERR_UNSUPPORTED, "multigrid: needs the node-block format");
        mg_solve(ctx, S, o, st);
    } else switch (o->precision) {
        case SFEM_PREC_FP64: run_pcg<double, double>(ctx, S, o, st, fmt); break;
        case SFEM_PREC_MIXED:
            if (fmt == SFEM_FMT_CSR) ensure_val32(ctx, S);
            run_pcg<float, float>(ctx, S, o, st, fmt);
            break;
        case SFEM_PREC_FP32_MATRIX:
            if (fmt == SFEM_FMT_CSR) ensure_val32(ctx, S);
            run_pcg<float, double>(ctx, S, o, st, fmt);
            break;
        case SFEM_PREC_SPLIT:
            if (fmt != SFEM_FMT_BSR) fail(SFEM_ERR_UNSUPPORTED, "split precision needs the node-block format");
            run_pcg<SplitVal, double>(ctx, S, o, st, fmt);
            break;
        default: fail(SFEM_ERR_ARGUMENT, "solve: unknown precision");
    }
    if (!st->converged) {
        static const char* why[] = {"", "", "breakdown", "stagnation", "iteration cap reached"};
        char buf[200];
        std::snprintf(buf, sizeof buf, "solve: %s after %lld iterations (relative residual %.3e)",
                      why[st->stop_reason], (long long)st->iterations, st->rel_residual);
        fail(SFEM_ERR_NOT_CONVERGED, buf);
    }
}

namespace {
template <typename TV, typename T>
double spmv_bench_impl(sfem_ctx* ctx, sfem_system* S, int fmt, int reps, bool nt) {
    const bool padded = fmt == SFEM_FMT_BSR && S->partial;
    const int64_t nf = padded ? S->nfb * S->bs : S->nf;
    const unsigned cap = (unsigned)std::max(8, ctx->n_cu * 8);
    const Launch lc = plan(nf, kThreads / kCsrLanes, cap);
    const Launch ls = S->has_block ? plan_sell(S->nslices, cap, ctx->tune_wave_slices) : Launch{8, 1};
    DevBuf<double> x64(ctx, nf);
    DevBuf<T> x(ctx, nf), y(ctx, nf);
    SFEM_HIP(hipMemcpyAsync(x64.get(), padded ? S->prhs.get() : S->rhs.get(), sizeof(double) * nf,
                            hipMemcpyDeviceToDevice, ctx->stream));
    if (std::is_same<T, float>::value)
        hipLaunchKernelGGL(f64_to_f32_kernel, dim3(1024), dim3(256), 0, ctx->stream, x64.get(), (float*)x.get(), nf);
    else
        SFEM_HIP(hipMemcpyAsync(x.get(), x64.get(), sizeof(double) * nf, hipMemcpyDeviceToDevice, ctx->stream));
    auto launch = [&]() {
        hipStream_t s = ctx->stream;
        if (fmt == SFEM_FMT_BSR) {
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(ls.grid), dim3(kThreads), 0, s, S->sl_ptr.get(), S->sl_w.get(),
                                   S->scol.get(), sell_vals<TV>(S), S->nslices, S->nfb, ls.per_block, x.get(), y.get());
            };
            if (S->bs == 3) {
                if (nt) go(sell_spmv_kernel<3, TV, T, true>);
                else go(sell_spmv_kernel<3, TV, T, false>);
            } else {
                if (nt) go(sell_spmv_kernel<2, TV, T, true>);
                else go(sell_spmv_kernel<2, TV, T, false>);
            }
        } else if constexpr (!std::is_same<TV, SplitVal>::value) {
            hipLaunchKernelGGL((csr_spmv_kernel<kCsrLanes, TV, T>), dim3(lc.grid), dim3(kThreads), 0, s,
                               S->rowptr.get(), S->col.get(), csr_vals<TV>(S), nf, lc.per_block, x.get(), y.get());
        }
        SFEM_LAUNCH_CHECK();
    };
    launch();  // warm
    hipEvent_t a, b;
    SFEM_HIP(hipEventCreate(&a));
    SFEM_HIP(hipEventCreate(&b));
    SFEM_HIP(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps; ++i) launch();
    SFEM_HIP(hipEventRecord(b, ctx->stream));
    sync(ctx);
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return reps > 0 ? ms / reps : 0.0;
}
}  // namespace

void system_spmv_bench(sfem_ctx* ctx, sfem_system* S, int precision, int format, int reps, double* avg_ms,
                       double* bytes) {
    const int fmt = resolve_format(S, format);
    if (fmt == SFEM_FMT_CSR) ensure_csr(ctx, S);
    if (fmt == SFEM_FMT_BSR) ensure_sell(ctx, S);
    if (precision != SFEM_PREC_FP64 && fmt == SFEM_FMT_CSR) ensure_val32(ctx, S);
    switch (precision) {
        case SFEM_PREC_FP64:
            *avg_ms = spmv_bench_impl<double, double>(ctx, S, fmt, reps, ctx->tune_nt);
            *bytes = spmv_bytes_for(S, fmt, 8, 8, 2);
            break;
        case SFEM_PREC_MIXED:
            *avg_ms = spmv_bench_impl<float, float>(ctx, S, fmt, reps, ctx->tune_nt);
            *bytes = spmv_bytes_for(S, fmt, 4, 4, 2);
            break;
        case SFEM_PREC_FP32_MATRIX:
            *avg_ms = spmv_bench_impl<float, double>(ctx, S, fmt, reps, ctx->tune_nt);
            *bytes = spmv_bytes_for(S, fmt, 4, 8, 2);
            break;
        case SFEM_PREC_SPLIT:
            if (fmt != SFEM_FMT_BSR) fail(SFEM_ERR_UNSUPPORTED, "split precision needs the node-block format");
            *avg_ms = spmv_bench_impl<SplitVal, double>(ctx, S, fmt, reps, ctx->tune_nt);
            *bytes = spmv_bytes_for(S, fmt, 6, 8, 2);
            break;
        default: fail(SFEM_ERR_ARGUMENT, "spmv bench: unknown precision");
    }
}

void stream_copy_bench(sfem_ctx* ctx, int64_t bytes, int reps, double* avg_ms) {
    const int64_t n4 = bytes / 16;
    if (n4 <= 0) fail(SFEM_ERR_ARGUMENT, "copy bench: size too small");
    DevBuf<v4f> a(ctx, n4), b(ctx, n4);
    memset_d(ctx, a.get(), 0, sizeof(v4f) * n4);
    hipEvent_t e0, e1;
    SFEM_HIP(hipEventCreate(&e0));
    SFEM_HIP(hipEventCreate(&e1));
    double best = 1e30;
    // best of a few launch shapes and of plain vs non-temporal access
    auto trial = [&](auto kern, unsigned grid) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, ctx->stream, a.get(), b.get(), n4);
        SFEM_LAUNCH_CHECK();
        SFEM_HIP(hipEventRecord(e0, ctx->stream));
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, ctx->stream, a.get(), b.get(), n4);
        SFEM_LAUNCH_CHECK();
        SFEM_HIP(hipEventRecord(e1, ctx->stream));
        sync(ctx);
        float ms = 0;
        SFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms / std::max(1, reps));
    };
    for (unsigned mult : {4u, 8u, 16u}) {
        const unsigned grid = (unsigned)std::max(8, ctx->n_cu * (int)mult);
        trial(copy_kernel<1, false>, grid);
        trial(copy_kernel<4, false>, grid);
        trial(copy_kernel<4, true>, grid);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = best;
}

}  // namespace sfem
