// pcg.hip -- Jacobi-preconditioned CG on K_ff for gfx950.
//
// Replaces the reference's dense inverse (smearFEM examples/vector3D.jl:318,
// inv(Matrix(K_free)) * C' * (-K q_d)) with an iterative solve that keeps K_ff
// sparse and resident in HBM.  Two precisions share one driver:
//
//   FP64  : vectors, operator and reductions in FP64 on the scalar CSR.
//   MIXED : FP32 operator (block-sliced or CSR) and FP32 vectors for the CG
//           recurrences, FP64 reductions; the solution accumulates in FP64 and
//           the true residual rhs - K_ff x is recomputed in FP64 whenever the
//           FP32 residual has dropped by `reliable_delta` since the last
//           replacement, or claims convergence ("reliable updates").  The
//           stopping test is always on that FP64 residual.
//
// Per iteration two kernels run, with no host round trip:
//   spmv_cg : p_new = z + beta p_old (formed on the fly, also while gathering),
//             q = K_ff p_new, partial p_new.q  -> last block: alpha
//   update  : x += alpha p, r -= alpha q, z = D^-1 r, partial r.z, r.r
//             -> last block: beta, flags
// Every `check_every` iterations the FP64 residual kernels run (they exit at
// once unless a flag is set) and the host reads the state once.
//
// Reductions are deterministic: wave64 butterfly shuffles, a fixed 4-wave
// block combine, per-block partials, and the last block to arrive sums the
// partials in a fixed order.  The grid is a fixed function of the problem
// size, so results are bitwise reproducible run to run.
#include <cmath>
#include <vector>

#include "sfem_objects.hpp"

namespace sfem {

namespace {

constexpr int kThreads = 256;

// ------------------------------------------------------------ helpers --
__device__ __forceinline__ double ld_coherent(const double* p) {
    unsigned long long v =
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

// True in every thread of the block that arrives last; its partial results
// (written before the call) are then visible to that block.
__device__ __forceinline__ bool last_block_arrive(unsigned int* counter) {
    __shared__ int is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) is_last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (is_last) __threadfence();
    return is_last;
}

// Fixed-order sum of nparts partial pairs by one block.
__device__ __forceinline__ void sum_partials2(const double* partials, int nparts, double* a, double* b, double* sh) {
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        s0 += ld_coherent(partials + 2 * i);
        s1 += ld_coherent(partials + 2 * i + 1);
    }
    *a = block_sum_256(s0, sh);
    *b = block_sum_256(s1, sh + 4);
}

// Logical block id: consecutive logical blocks land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so each XCD's L2
// sees a contiguous part of the rows and of the gathered vector.
__device__ __forceinline__ int64_t logical_block() {
    const int64_t g = gridDim.x, b = blockIdx.x;
    if (g % 8) return b;
    return (b % 8) * (g / 8) + b / 8;
}

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return fma(a, b, c); }

// ------------------------------------------------------- operator rows --
// CSR, G lanes per row.  Returns the row sum in every lane of the group.
template <int G, typename TV, typename ACC, typename Gather>
__device__ __forceinline__ ACC csr_row(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                       const TV* __restrict__ val, int64_t row, int lane, Gather gx) {
    ACC s = 0;
    const int32_t e = rowptr[row + 1];
    for (int32_t k = rowptr[row] + lane; k < e; k += G) s = fmaT<ACC>((ACC)val[k], gx(col[k]), s);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Node-block sliced format: this lane's block row of slice `sl`.
template <int BS, typename TV, typename ACC, typename Gather>
__device__ __forceinline__ void sell_row(const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                         const int32_t* __restrict__ scol, const TV* __restrict__ sval, int64_t sl,
                                         int lane, Gather gx, ACC out[BS]) {
    constexpr int B2 = BS * BS;
    const int64_t base = sl_ptr[sl];
    const int w = sl_w[sl];
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = 0;
#pragma unroll 2
    for (int k = 0; k < w; ++k) {
        const int64_t slot = base + (int64_t)k * 64;
        const int64_t c = scol[slot + lane];
        ACC xv[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) xv[i] = gx(c * BS + i);
        const TV* v = sval + slot * B2 + lane;
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) out[j] = fmaT<ACC>((ACC)v[(j * BS + i) * 64], xv[i], out[j]);
    }
}

// ---------------------------------------------------------- spmv_cg ----
template <int G, typename TV, typename T>
__global__ __launch_bounds__(kThreads) void csr_spmv_cg_kernel(const int32_t* __restrict__ rowptr,
                                                               const int32_t* __restrict__ col,
                                                               const TV* __restrict__ val, int64_t nrows,
                                                               int64_t rows_per_block, const T* __restrict__ z,
                                                               const T* __restrict__ pold, T* __restrict__ pnew,
                                                               T* __restrict__ q, CgState* st,
                                                               double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const T beta = (T)st->beta;
    const int64_t r0 = logical_block() * rows_per_block;
    const int64_t r1 = min(nrows, r0 + rows_per_block);
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    double dot = 0.0;
    auto gx = [=](int64_t c) { return fmaT<T>(beta, pold[c], z[c]); };
    for (int64_t row = r0 + g; row < r1; row += kThreads / G) {
        const T s = csr_row<G, TV, T>(rowptr, col, val, row, lane, gx);
        if (lane == 0) {
            const T pn = fmaT<T>(beta, pold[row], z[row]);
            pnew[row] = pn;
            q[row] = s;
            dot += (double)pn * (double)s;
        }
    }
    const double bsum = block_sum_256(dot, sh);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = bsum;
    if (last_block_arrive(&st->counter)) {
        double pq, unused;
        sum_partials2(partials, gridDim.x, &pq, &unused, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz / pq;
            if (!(pq > 0.0) || !std::isfinite(st->alpha)) st->done = 2;  // breakdown
            st->counter = 0;
        }
    }
}

template <int BS, typename TV, typename T>
__global__ __launch_bounds__(kThreads) void sell_spmv_cg_kernel(
    const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w, const int32_t* __restrict__ scol,
    const TV* __restrict__ sval, int64_t nslices, int64_t nrows, int64_t slices_per_block, const T* __restrict__ z,
    const T* __restrict__ pold, T* __restrict__ pnew, T* __restrict__ q, CgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const T beta = (T)st->beta;
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double dot = 0.0;
    auto gx = [=](int64_t c) { return fmaT<T>(beta, pold[c], z[c]); };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        T out[BS];
        sell_row<BS, TV, T>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                const T pn = fmaT<T>(beta, pold[d], z[d]);
                pnew[d] = pn;
                q[d] = out[j];
                dot += (double)pn * (double)out[j];
            }
        }
    }
    const double bsum = block_sum_256(dot, sh);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = bsum;
    if (last_block_arrive(&st->counter)) {
        double pq, unused;
        sum_partials2(partials, gridDim.x, &pq, &unused, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz / pq;
            if (!(pq > 0.0) || !std::isfinite(st->alpha)) st->done = 2;
            st->counter = 0;
        }
    }
}

// ----------------------------------------------------------- update ----
template <typename T>
__global__ __launch_bounds__(kThreads) void cg_update_kernel(T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                             const T* __restrict__ p, const T* __restrict__ q,
                                                             const T* __restrict__ dinv, int64_t n,
                                                             int64_t per_block, int mixed, CgState* st,
                                                             double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const T a = (T)st->alpha;
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(n, i0 + per_block);
    double rz = 0.0, rr = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        const T xi = fmaT<T>(a, p[i], x[i]);
        const T ri = fmaT<T>(-a, q[i], r[i]);
        const T zi = dinv[i] * ri;
        x[i] = xi;
        r[i] = ri;
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += (double)ri * (double)ri;
    }
    const double b0 = block_sum_256(rz, sh);
    const double b1 = block_sum_256(rr, sh + 4);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = b0;
        partials[2 * blockIdx.x + 1] = b1;
    }
    if (last_block_arrive(&st->counter)) {
        double trz, trr;
        sum_partials2(partials, gridDim.x, &trz, &trr, sh);
        if (threadIdx.x == 0) {
            st->rz_prev = st->rz;
            st->rz = trz;
            st->beta = trz / st->rz_prev;
            st->rr = trr;
            st->iters += 1;
            if (trr <= st->tol2bb || (mixed && trr < st->delta2 * st->maxrr)) st->need_update = 1;
            if (!std::isfinite(trr)) st->done = 2;
            st->counter = 0;
        }
    }
}

// ------------------------------------------- FP64 residual replacement --
// x64 += xs, xs = 0 (mixed only), when an update is pending.
__global__ __launch_bounds__(kThreads) void fold_kernel(double* __restrict__ x64, float* __restrict__ xs, int64_t n,
                                                        const CgState* st) {
    if (st->done || !st->need_update) return;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        x64[i] += (double)xs[i];
        xs[i] = 0.0f;
    }
}

__device__ __forceinline__ void resid_finish(CgState* st, double* partials, double* sh) {
    double trr, trz;
    sum_partials2(partials, gridDim.x, &trr, &trz, sh);
    if (threadIdx.x == 0) {
        st->rr_true = trr;
        st->rr = trr;
        st->updates += 1;
        st->need_update = 0;
        if (trr <= st->tol2bb) {
            st->done = 1;
        } else {
            st->rz = trz;
            st->beta = trz / st->rz_prev;
            st->maxrr = trr;
        }
        if (!std::isfinite(trr)) st->done = 2;
        st->counter = 0;
    }
}

// r = rhs - K_ff x64 in FP64; writes r, z = D^-1 r in the CG precision T.
template <int G, typename T>
__global__ __launch_bounds__(kThreads) void csr_resid_kernel(const int32_t* __restrict__ rowptr,
                                                             const int32_t* __restrict__ col,
                                                             const double* __restrict__ val, int64_t nrows,
                                                             int64_t rows_per_block, const double* __restrict__ x64,
                                                             const double* __restrict__ rhs,
                                                             const T* __restrict__ dinv, T* __restrict__ r,
                                                             T* __restrict__ z, CgState* st,
                                                             double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done || !st->need_update) return;
    const int64_t r0 = logical_block() * rows_per_block;
    const int64_t r1 = min(nrows, r0 + rows_per_block);
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    double rr = 0.0, rz = 0.0;
    auto gx = [=](int64_t c) { return x64[c]; };
    for (int64_t row = r0 + g; row < r1; row += kThreads / G) {
        const double s = csr_row<G, double, double>(rowptr, col, val, row, lane, gx);
        if (lane == 0) {
            const double res = rhs[row] - s;
            const T rt = (T)res;
            const T zt = dinv[row] * rt;
            r[row] = rt;
            z[row] = zt;
            rr += res * res;
            rz += (double)rt * (double)zt;
        }
    }
    const double b0 = block_sum_256(rr, sh);
    const double b1 = block_sum_256(rz, sh + 4);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = b0;
        partials[2 * blockIdx.x + 1] = b1;
    }
    if (last_block_arrive(&st->counter)) resid_finish(st, partials, sh);
}

template <int BS, typename T>
__global__ __launch_bounds__(kThreads) void sell_resid_kernel(
    const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w, const int32_t* __restrict__ scol,
    const double* __restrict__ sval, int64_t nslices, int64_t nrows, int64_t slices_per_block,
    const double* __restrict__ x64, const double* __restrict__ rhs, const T* __restrict__ dinv, T* __restrict__ r,
    T* __restrict__ z, CgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done || !st->need_update) return;
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double rr = 0.0, rz = 0.0;
    auto gx = [=](int64_t c) { return x64[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        double out[BS];
        sell_row<BS, double, double>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                const double res = rhs[d] - out[j];
                const T rt = (T)res;
                const T zt = dinv[d] * rt;
                r[d] = rt;
                z[d] = zt;
                rr += res * res;
                rz += (double)rt * (double)zt;
            }
        }
    }
    const double b0 = block_sum_256(rr, sh);
    const double b1 = block_sum_256(rz, sh + 4);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = b0;
        partials[2 * blockIdx.x + 1] = b1;
    }
    if (last_block_arrive(&st->counter)) resid_finish(st, partials, sh);
}

// bb = rhs.rhs (FP64), once per solve
__global__ __launch_bounds__(kThreads) void norm2_kernel(const double* __restrict__ v, int64_t n, int64_t per_block,
                                                         CgState* st, double* __restrict__ partials, double rtol,
                                                         double delta) {
    __shared__ double sh[8];
    const int64_t i0 = (int64_t)blockIdx.x * per_block;
    const int64_t i1 = min(n, i0 + per_block);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s += v[i] * v[i];
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = b;
        partials[2 * blockIdx.x + 1] = 0.0;
    }
    if (last_block_arrive(&st->counter)) {
        double bb, unused;
        sum_partials2(partials, gridDim.x, &bb, &unused, sh);
        if (threadIdx.x == 0) {
            st->bb = bb;
            st->tol2bb = rtol * rtol * bb;
            st->delta2 = delta * delta;
            st->maxrr = bb;
            st->counter = 0;
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

template <int BS, typename TV, typename T>
__global__ __launch_bounds__(kThreads) void sell_spmv_kernel(const int64_t* __restrict__ sl_ptr,
                                                             const int32_t* __restrict__ sl_w,
                                                             const int32_t* __restrict__ scol,
                                                             const TV* __restrict__ sval, int64_t nslices,
                                                             int64_t nrows, int64_t slices_per_block,
                                                             const T* __restrict__ x, T* __restrict__ y) {
    const int64_t s0 = logical_block() * slices_per_block;
    const int64_t s1 = min(nslices, s0 + slices_per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto gx = [=](int64_t c) { return x[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        T out[BS];
        sell_row<BS, TV, T>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
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

__global__ __launch_bounds__(kThreads) void copy_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        b[i] = a[i];
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

struct Ops {
    sfem_ctx* ctx;
    sfem_system* S;
    int fmt;   // SFEM_FMT_CSR / SFEM_FMT_BSR
    bool mixed;
    unsigned cap;
    Launch csr, sell, vec;
};

template <typename T>
void launch_spmv_cg(const Ops& o, const T* z, const T* pold, T* pnew, T* q, CgState* st, double* partials) {
    sfem_system* S = o.S;
    hipStream_t s = o.ctx->stream;
    if (o.fmt == SFEM_FMT_BSR) {
        const void* v = o.mixed ? (const void*)S->sval32.get() : (const void*)S->sval64.get();
        using TV = T;
        if (S->bs == 3)
            hipLaunchKernelGGL((sell_spmv_cg_kernel<3, TV, T>), dim3(o.sell.grid), dim3(kThreads), 0, s,
                               S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), (const TV*)v, S->nslices, S->nfb,
                               o.sell.per_block, z, pold, pnew, q, st, partials);
        else
            hipLaunchKernelGGL((sell_spmv_cg_kernel<2, TV, T>), dim3(o.sell.grid), dim3(kThreads), 0, s,
                               S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), (const TV*)v, S->nslices, S->nfb,
                               o.sell.per_block, z, pold, pnew, q, st, partials);
    } else {
        const void* v = o.mixed ? (const void*)S->val32.get() : (const void*)S->val64.get();
        hipLaunchKernelGGL((csr_spmv_cg_kernel<kCsrLanes, T, T>), dim3(o.csr.grid), dim3(kThreads), 0, s,
                           S->rowptr.get(), S->col.get(), (const T*)v, S->nf, o.csr.per_block, z, pold, pnew, q, st,
                           partials);
    }
    SFEM_LAUNCH_CHECK();
}

template <typename T>
void launch_resid(const Ops& o, const T* dinv, T* r, T* z, CgState* st, double* partials) {
    sfem_system* S = o.S;
    hipStream_t s = o.ctx->stream;
    if (o.fmt == SFEM_FMT_BSR) {
        if (S->bs == 3)
            hipLaunchKernelGGL((sell_resid_kernel<3, T>), dim3(o.sell.grid), dim3(kThreads), 0, s, S->sl_ptr.get(),
                               S->sl_w.get(), S->scol.get(), S->sval64.get(), S->nslices, S->nfb, o.sell.per_block,
                               S->x.get(), S->rhs.get(), dinv, r, z, st, partials);
        else
            hipLaunchKernelGGL((sell_resid_kernel<2, T>), dim3(o.sell.grid), dim3(kThreads), 0, s, S->sl_ptr.get(),
                               S->sl_w.get(), S->scol.get(), S->sval64.get(), S->nslices, S->nfb, o.sell.per_block,
                               S->x.get(), S->rhs.get(), dinv, r, z, st, partials);
    } else {
        hipLaunchKernelGGL((csr_resid_kernel<kCsrLanes, T>), dim3(o.csr.grid), dim3(kThreads), 0, s, S->rowptr.get(),
                           S->col.get(), S->val64.get(), S->nf, o.csr.per_block, S->x.get(), S->rhs.get(), dinv, r, z,
                           st, partials);
    }
    SFEM_LAUNCH_CHECK();
}

double spmv_cg_bytes(const sfem_system* S, int fmt, bool mixed) {
    const double sv = mixed ? 4.0 : 8.0;
    const double vec = 4.0 * (double)S->nf * sv;  // z, p_old read; p_new, q written
    if (fmt == SFEM_FMT_BSR)
        return (double)S->nnzb * (S->bs * S->bs * sv + 4.0) + (double)(S->nslices + 1) * 8.0 +
               (double)S->nslices * 4.0 + vec;
    return (double)S->nnz * (sv + 4.0) + (double)(S->nf + 1) * 4.0 + vec;
}

double spmv_plain_bytes(const sfem_system* S, int fmt, bool mixed) {
    const double sv = mixed ? 4.0 : 8.0;
    const double vec = 2.0 * (double)S->nf * sv;  // x read once, y written once
    if (fmt == SFEM_FMT_BSR)
        return (double)S->nnzb * (S->bs * S->bs * sv + 4.0) + (double)(S->nslices + 1) * 8.0 +
               (double)S->nslices * 4.0 + vec;
    return (double)S->nnz * (sv + 4.0) + (double)(S->nf + 1) * 4.0 + vec;
}

void ensure_val32(sfem_ctx* ctx, sfem_system* S) {
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

template <typename T>
void run_pcg(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* stats, int fmt) {
    const bool mixed = std::is_same<T, float>::value;
    const int64_t nf = S->nf;
    Ops ops{ctx, S, fmt, mixed, (unsigned)std::max(8, ctx->n_cu * 8), {}, {}, {}};
    ops.csr = plan(nf, kThreads / kCsrLanes, ops.cap);
    if (S->has_block) ops.sell = plan(S->nslices, kThreads / 64, ops.cap);
    ops.vec = plan(nf, kThreads, std::max(8, ctx->n_cu * 4));
    const unsigned maxgrid = std::max({ops.csr.grid, ops.sell.grid, ops.vec.grid});

    DevBuf<T> r(ctx, nf), z(ctx, nf), q(ctx, nf), p0(ctx, nf), p1(ctx, nf);
    DevBuf<float> xs;
    DevBuf<double> partials(ctx, 2 * (size_t)maxgrid);
    DevBuf<CgState> dst(ctx, 1);
    const T* dinv = mixed ? (const T*)S->dinv32.get() : (const T*)S->dinv64.get();
    T* x = mixed ? nullptr : (T*)S->x.get();
    if (mixed) {
        xs.reset(ctx, nf);
        memset_d(ctx, xs.get(), 0, sizeof(float) * nf);
        x = (T*)xs.get();
    }
    if (o->x0_zero) memset_d(ctx, S->x.get(), 0, sizeof(double) * nf);
    memset_d(ctx, p0.get(), 0, sizeof(T) * nf);
    memset_d(ctx, p1.get(), 0, sizeof(T) * nf);
    CgState init{};
    init.rz = 0.0;
    init.rz_prev = HUGE_VAL;
    init.need_update = 1;
    h2d(ctx, dst.get(), &init, sizeof(init));
    CgState* st = dst.get();

    hipEvent_t t0, t1;
    SFEM_HIP(hipEventCreate(&t0));
    SFEM_HIP(hipEventCreate(&t1));
    const int check = std::max(1, o->check_every);
    std::vector<hipEvent_t> ev;
    if (o->time_kernels) {
        ev.resize(2 * (size_t)check);
        for (auto& e : ev) SFEM_HIP(hipEventCreate(&e));
    }
    SFEM_HIP(hipEventRecord(t0, ctx->stream));
    hipLaunchKernelGGL(norm2_kernel, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, S->rhs.get(), nf,
                       ops.vec.per_block, st, partials.get(), o->rtol, mixed ? o->reliable_delta : 0.0);
    SFEM_LAUNCH_CHECK();
    launch_resid<T>(ops, dinv, r.get(), z.get(), st, partials.get());

    CgState host{};
    int64_t enq = 0, spmv = 0, timed = 0, resid_launches = 1;
    double spmv_ms = 0.0;
    T* pc = p0.get();
    T* pn = p1.get();
    int done = 0;
    d2h(ctx, &host, st, sizeof(host));
    done = host.done;
    while (!done && enq < o->max_iter) {
        const int64_t chunk = std::min<int64_t>(check, o->max_iter - enq);
        for (int64_t k = 0; k < chunk; ++k) {
            if (o->time_kernels) SFEM_HIP(hipEventRecord(ev[2 * k], ctx->stream));
            launch_spmv_cg<T>(ops, z.get(), pc, pn, q.get(), st, partials.get());
            if (o->time_kernels) SFEM_HIP(hipEventRecord(ev[2 * k + 1], ctx->stream));
            std::swap(pc, pn);
            hipLaunchKernelGGL(cg_update_kernel<T>, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, x, r.get(),
                               z.get(), pc, q.get(), dinv, nf, ops.vec.per_block, mixed ? 1 : 0, st,
                               partials.get());
            SFEM_LAUNCH_CHECK();
        }
        enq += chunk;
        if (mixed) {
            hipLaunchKernelGGL(fold_kernel, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, S->x.get(), xs.get(),
                               nf, st);
            SFEM_LAUNCH_CHECK();
        }
        launch_resid<T>(ops, dinv, r.get(), z.get(), st, partials.get());
        ++resid_launches;
        d2h(ctx, &host, st, sizeof(host));
        done = host.done;
        if (o->time_kernels) {
            // launches after convergence exit at once; count only the real ones
            const int64_t real = std::min<int64_t>(chunk, std::max<int64_t>(0, host.iters - spmv));
            for (int64_t k = 0; k < real; ++k) {
                float ms = 0;
                SFEM_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                spmv_ms += ms;
            }
            timed += real;
        }
        spmv = host.iters;
    }
    if (!done) {  // iteration cap: fold and evaluate the true residual once more
        if (mixed) {
            CgState h2 = host;
            h2.need_update = 1;
            h2d(ctx, st, &h2, sizeof(h2));
            hipLaunchKernelGGL(fold_kernel, dim3(ops.vec.grid), dim3(kThreads), 0, ctx->stream, S->x.get(),
                               xs.get(), nf, st);
            SFEM_LAUNCH_CHECK();
        } else {
            CgState h2 = host;
            h2.need_update = 1;
            h2d(ctx, st, &h2, sizeof(h2));
        }
        launch_resid<T>(ops, dinv, r.get(), z.get(), st, partials.get());
        ++resid_launches;
        d2h(ctx, &host, st, sizeof(host));
    }
    SFEM_HIP(hipEventRecord(t1, ctx->stream));
    sync(ctx);
    float total = 0;
    SFEM_HIP(hipEventElapsedTime(&total, t0, t1));
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);

    stats->iterations = host.iters;
    stats->fp64_updates = host.updates;
    stats->spmv_launches = host.iters;
    stats->spmv64_launches = host.updates;
    stats->rel_residual = host.bb > 0 ? std::sqrt(host.rr_true / host.bb) : 0.0;
    stats->solve_ms = total;
    stats->spmv_ms = spmv_ms;
    stats->spmv_timed = timed;
    stats->spmv_bytes = spmv_cg_bytes(S, fmt, mixed);
    stats->converged = host.done == 1;
    stats->format = fmt;
    (void)resid_launches;
}

}  // namespace

void system_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st) {
    *st = sfem_solve_stats{};
    if (S->nf == 0) {
        st->converged = 1;
        return;
    }
    if (!(o->rtol >= 0.0) || o->max_iter < 0) fail(SFEM_ERR_ARGUMENT, "solve: bad rtol or max_iter");
    const bool mixed = o->precision == SFEM_PREC_MIXED;
    if (!mixed && o->precision != SFEM_PREC_FP64) fail(SFEM_ERR_ARGUMENT, "solve: unknown precision");
    const int fmt = resolve_format(S, o->format);
    if (mixed) {
        if (fmt == SFEM_FMT_CSR) ensure_val32(ctx, S);
        run_pcg<float>(ctx, S, o, st, fmt);
    } else {
        run_pcg<double>(ctx, S, o, st, fmt);
    }
    if (!st->converged) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "solve: %s after %lld iterations (relative residual %.3e)",
                      st->iterations >= o->max_iter ? "iteration cap reached" : "breakdown",
                      (long long)st->iterations, st->rel_residual);
        fail(SFEM_ERR_NOT_CONVERGED, buf);
    }
}

void system_spmv_bench(sfem_ctx* ctx, sfem_system* S, int precision, int format, int reps, double* avg_ms,
                       double* bytes) {
    const bool mixed = precision == SFEM_PREC_MIXED;
    const int fmt = resolve_format(S, format);
    if (mixed && fmt == SFEM_FMT_CSR) ensure_val32(ctx, S);
    const int64_t nf = S->nf;
    const unsigned cap = (unsigned)std::max(8, ctx->n_cu * 8);
    const Launch lc = plan(nf, kThreads / kCsrLanes, cap);
    const Launch ls = S->has_block ? plan(S->nslices, kThreads / 64, cap) : Launch{8, 1};
    DevBuf<double> x64(ctx, nf), y64(ctx, nf);
    DevBuf<float> x32(ctx, nf), y32(ctx, nf);
    SFEM_HIP(hipMemcpyAsync(x64.get(), S->rhs.get(), sizeof(double) * nf, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(1024), dim3(256), 0, ctx->stream, x64.get(), x32.get(), nf);
    SFEM_LAUNCH_CHECK();
    auto launch = [&]() {
        hipStream_t s = ctx->stream;
        if (fmt == SFEM_FMT_BSR) {
            if (mixed) {
                if (S->bs == 3)
                    hipLaunchKernelGGL((sell_spmv_kernel<3, float, float>), dim3(ls.grid), dim3(kThreads), 0, s,
                                       S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval32.get(), S->nslices,
                                       S->nfb, ls.per_block, x32.get(), y32.get());
                else
                    hipLaunchKernelGGL((sell_spmv_kernel<2, float, float>), dim3(ls.grid), dim3(kThreads), 0, s,
                                       S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval32.get(), S->nslices,
                                       S->nfb, ls.per_block, x32.get(), y32.get());
            } else {
                if (S->bs == 3)
                    hipLaunchKernelGGL((sell_spmv_kernel<3, double, double>), dim3(ls.grid), dim3(kThreads), 0, s,
                                       S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval64.get(), S->nslices,
                                       S->nfb, ls.per_block, x64.get(), y64.get());
                else
                    hipLaunchKernelGGL((sell_spmv_kernel<2, double, double>), dim3(ls.grid), dim3(kThreads), 0, s,
                                       S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval64.get(), S->nslices,
                                       S->nfb, ls.per_block, x64.get(), y64.get());
            }
        } else if (mixed) {
            hipLaunchKernelGGL((csr_spmv_kernel<kCsrLanes, float, float>), dim3(lc.grid), dim3(kThreads), 0, s,
                               S->rowptr.get(), S->col.get(), S->val32.get(), nf, lc.per_block, x32.get(), y32.get());
        } else {
            hipLaunchKernelGGL((csr_spmv_kernel<kCsrLanes, double, double>), dim3(lc.grid), dim3(kThreads), 0, s,
                               S->rowptr.get(), S->col.get(), S->val64.get(), nf, lc.per_block, x64.get(),
                               y64.get());
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
    *avg_ms = reps > 0 ? ms / reps : 0.0;
    *bytes = spmv_plain_bytes(S, fmt, mixed);
}

void stream_copy_bench(sfem_ctx* ctx, int64_t bytes, int reps, double* avg_ms) {
    const int64_t n4 = bytes / 16;
    if (n4 <= 0) fail(SFEM_ERR_ARGUMENT, "copy bench: size too small");
    DevBuf<float4> a(ctx, n4), b(ctx, n4);
    memset_d(ctx, a.get(), 0, sizeof(float4) * n4);
    const unsigned grid = (unsigned)std::max(8, ctx->n_cu * 8);
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kThreads), 0, ctx->stream, a.get(), b.get(), n4);
    SFEM_LAUNCH_CHECK();
    hipEvent_t e0, e1;
    SFEM_HIP(hipEventCreate(&e0));
    SFEM_HIP(hipEventCreate(&e1));
    SFEM_HIP(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kThreads), 0, ctx->stream, a.get(), b.get(), n4);
    SFEM_LAUNCH_CHECK();
    SFEM_HIP(hipEventRecord(e1, ctx->stream));
    sync(ctx);
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / std::max(1, reps);
}

}  // namespace sfem
