// batch.hip -- many independent meshes assembled together and solved by
// persistent (cooperative) kernel launches; the same engine runs the request
// table of the serving loop (serve.hip).
//
// A batch is a block-diagonal system: the meshes are concatenated (node ids
// offset per problem) and assembled by the single-mesh sort-by-key pipeline
// (duplicates never cross problems, so each problem's K is bitwise the one it
// gets alone), then Dirichlet-eliminated into one node-block sliced operator
// in which every problem starts on a fresh 64-row slice ("padded layout":
// padding rows have zero matrix rows, zero rhs and unit diagonal, so they
// stay exactly zero through the iteration).
//
// The solver keeps one CG state per problem.  A problem's rows are cut into
// work items of `item_slices` slices; each per-problem dot product is formed
// from that problem's item partials by one block in a fixed order.  A
// problem's arithmetic therefore depends only on its own data and on
// item_slices: it is bitwise the same alone, in this batch, in any other
// batch, for any grid size and for any split of its iterations over launches.
//
// Reference: the per-problem maths is smearFEM's assemble_system
// (src/fem.jl:135-256) and the example's elimination and solve
// (examples/vector3D.jl:314-322); batching has no reference counterpart.

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "batch.hpp"
#include "block_ops.hpp"
#include "grid_barrier.hpp"
#include "solver_common.hpp"


using namespace sfem;

sfem_batch::~sfem_batch() {
    if (K) sfem_csr_destroy(K);
    if (bc) sfem_bc_destroy(bc);
    if (mesh) sfem_mesh_destroy(mesh);
}

namespace sfem {
namespace {

// ------------------------------------------------ padded elimination ---
// A node with any free dof is a block row; its fixed dofs (a partially fixed
// node) get a zero row / column and a unit diagonal, zero rhs and zero start,
// so they stay exactly 0.
__global__ void unit_free_kernel(const uint8_t* __restrict__ fixed, int64_t nunits, int bs,
                                 uint32_t* __restrict__ ufree) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int c = 0;
    for (int i = 0; i < bs; ++i) c += fixed[u * bs + i] ? 1 : 0;
    ufree[u] = c < bs ? 1u : 0u;
}

// out[i] = upos[idx[i]] (idx[i] == nunits gives the total)
__global__ void gather_pos_kernel(const uint32_t* __restrict__ upos, const int64_t* __restrict__ idx, int64_t n,
                                  int64_t nunits, const uint32_t* __restrict__ total, uint32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = idx[i] < nunits ? upos[idx[i]] : *total;
}

__global__ void pad_map_kernel(const uint32_t* __restrict__ ufree, const uint32_t* __restrict__ upos,
                               const int64_t* __restrict__ node_off, const uint32_t* __restrict__ first_free,
                               const int64_t* __restrict__ pad_base, int64_t nprob, int64_t nunits,
                               int32_t* __restrict__ pad_of) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    if (!ufree[u]) { pad_of[u] = -1; return; }
    int64_t lo = 0, hi = nprob - 1;  // last problem whose first node is <= u
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (node_off[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    pad_of[u] = (int32_t)(pad_base[lo] + (int64_t)(upos[u] - first_free[lo]));
}

__global__ void pad_blen_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                const int32_t* __restrict__ pad_of, int64_t nunits, int32_t* __restrict__ blen) {
    const int64_t cu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cu >= nunits) return;
    const int32_t row = pad_of[cu];
    if (row < 0) return;
    int cnt = 0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) cnt += pad_of[browidx[u]] >= 0 ? 1 : 0;
    blen[row] = cnt;
}

__global__ void slice_w_kernel(const int32_t* __restrict__ blen, int64_t nslices, int32_t* __restrict__ sl_w,
                               int64_t* __restrict__ cnt) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int w = 0;
    for (int l = 0; l < 64; ++l) w = max(w, blen[s * 64 + l]);
    sl_w[s] = w;
    cnt[s] = (int64_t)w * 64;
}

// Row `row` of the sliced operator from column unit cu of K (K symmetric):
// its free blocks in ascending order, then padding slots (own column, zeros).
template <int BS>
__global__ void sell_fill_pad_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                     const double* __restrict__ bval, const int32_t* __restrict__ pad_of,
                                     int64_t nunits, const int64_t* __restrict__ sl_ptr,
                                     const int32_t* __restrict__ sl_w, int32_t* __restrict__ scol,
                                     double* __restrict__ sval64, float* __restrict__ sval32,
                                     int16_t* __restrict__ sval16, int64_t nrows, double* __restrict__ binv64,
                                     float* __restrict__ binv32, const uint8_t* __restrict__ fixed) {
    constexpr int B2 = BS * BS;
    const int64_t cu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cu >= nunits) return;
    const int32_t row = pad_of[cu];
    if (row < 0) return;
    const int64_t s = row / 64;
    const int lane = row % 64;
    int k = 0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t ru = browidx[u];
        const int32_t c = pad_of[ru];
        if (c < 0) continue;
        const int64_t slot = sl_ptr[s] + (int64_t)k * 64;
        scol[slot + lane] = c;
        // plane j*BS+i holds M[j][i] = K[(row,j),(c,i)] = K[(c,i),(row,j)];
        // fixed dofs of a partially fixed node: zero row / column, unit diagonal
        double m[B2];
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                const bool masked = fixed[cu * BS + j] || fixed[ru * BS + i];
                m[j * BS + i] = masked ? ((ru == cu && i == j) ? 1.0 : 0.0) : bval[u * B2 + i * BS + j];
            }
#pragma unroll
        for (int v = 0; v < B2; ++v) {
            sval64[slot * B2 + v * 64 + lane] = m[v];
            sval32[slot * B2 + v * 64 + lane] = (float)m[v];
            sval16[slot * B2 + v * 64 + lane] = split_correction(m[v]);
        }
        if (c == row) {  // diagonal block: its inverse for block-Jacobi preconditioning
            double inv[B2];
            invert_block<BS>(m, inv);
#pragma unroll
            for (int v = 0; v < B2; ++v) {
                binv64[v * nrows + row] = inv[v];
                binv32[v * nrows + row] = (float)inv[v];
            }
        }
        ++k;
    }
    for (; k < sl_w[s]; ++k) {
        const int64_t slot = sl_ptr[s] + (int64_t)k * 64;
        scol[slot + lane] = row;
        for (int v = 0; v < B2; ++v) {
            sval64[slot * B2 + v * 64 + lane] = 0.0;
            sval32[slot * B2 + v * 64 + lane] = 0.0f;
            sval16[slot * B2 + v * 64 + lane] = 0;
        }
    }
}

// identity blocks for every row (padding rows keep them)
__global__ void binv_identity_kernel(double* __restrict__ binv64, float* __restrict__ binv32, int64_t nrows,
                                     int bs) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    for (int j = 0; j < bs; ++j)
        for (int i = 0; i < bs; ++i) {
            binv64[(j * bs + i) * nrows + row] = i == j ? 1.0 : 0.0;
            binv32[(j * bs + i) * nrows + row] = i == j ? 1.0f : 0.0f;
        }
}

// padding rows (no source column unit, blen < 0): every slot is padding
template <int BS>
__global__ void sell_pad_gap_kernel(const int32_t* __restrict__ blen, int64_t nrows,
                                    const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                    int32_t* __restrict__ scol, double* __restrict__ sval64,
                                    float* __restrict__ sval32, int16_t* __restrict__ sval16) {
    constexpr int B2 = BS * BS;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows || blen[row] >= 0) return;
    const int64_t s = row / 64;
    const int lane = (int)(row % 64);
    for (int k = 0; k < sl_w[s]; ++k) {
        const int64_t slot = sl_ptr[s] + (int64_t)k * 64;
        scol[slot + lane] = (int32_t)row;
        for (int v = 0; v < B2; ++v) {
            sval64[slot * B2 + v * 64 + lane] = 0.0;
            sval32[slot * B2 + v * 64 + lane] = 0.0f;
            sval16[slot * B2 + v * 64 + lane] = 0;
        }
    }
}

__global__ void fill_pad_vectors_kernel(double* __restrict__ rhs, double* __restrict__ dinv64,
                                        float* __restrict__ dinv32, double* __restrict__ x, int64_t npad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    rhs[i] = 0.0;
    dinv64[i] = 1.0;
    dinv32[i] = 1.0f;
    x[i] = 0.0;
}

// rhs = f - K q_d and 1/diag for every free dof, in padded numbering
__global__ void pad_rhs_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                               const double* __restrict__ bval, const int32_t* __restrict__ pad_of,
                               const uint8_t* __restrict__ fixed, const double* __restrict__ qd,
                               const double* __restrict__ f, int64_t n, int bs, double* __restrict__ rhs,
                               double* __restrict__ dinv64, float* __restrict__ dinv32) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t cu = s / bs;
    const int j = (int)(s - cu * bs);
    const int32_t prow = pad_of[cu];
    if (prow < 0) return;
    if (fixed[s]) {  // fixed dof of a partially fixed node: stays 0
        const int64_t pd = (int64_t)prow * bs + j;
        rhs[pd] = 0.0;
        dinv64[pd] = 1.0;
        dinv32[pd] = 1.0f;
        return;
    }
    double lift = 0.0, diag = 0.0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t r = browidx[u];
        for (int i = 0; i < bs; ++i) {
            const int64_t d = r * bs + i;
            const double v = bval[u * bs * bs + i * bs + j];
            if (fixed[d]) lift = fma(v, qd[d], lift);
            else if (d == s) diag = v;
        }
    }
    const int64_t pd = (int64_t)prow * bs + j;
    rhs[pd] = f[s] - lift;
    dinv64[pd] = diag != 0.0 ? 1.0 / diag : 1.0;
    dinv32[pd] = (float)dinv64[pd];
}

__global__ void batch_recombine_kernel(const int32_t* __restrict__ pad_of, const uint8_t* __restrict__ fixed,
                                       const double* __restrict__ x, const double* __restrict__ qd, int64_t n, int bs,
                                       double* __restrict__ q) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const int32_t pr = pad_of[d / bs];
    q[d] = (pr >= 0 && !fixed[d]) ? x[(int64_t)pr * bs + d % bs] : qd[d];
}

__global__ void fold_all_kernel(double* __restrict__ x64, float* __restrict__ xs, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x64[i] += (double)xs[i];
        xs[i] = 0.0f;
    }
}

// ----------------------------------------------------- persistent PCG ---
// Block-uniform OR of pred over the table's problems.
template <typename F>
__device__ __forceinline__ bool block_any(int nprob, F pred) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < nprob; i += kThreads) mine |= pred(i) ? 1 : 0;
    if (mine) atomicOr(&flag, 1);
    __syncthreads();
    const bool r = flag != 0;
    __syncthreads();
    return r;
}

// Sum of partial slot k over items [i0, i1), fixed order for a 256-thread block.
__device__ __forceinline__ double items_sum(const double* part, int stride, int k, int i0, int i1, double* sh) {
    double s = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += kThreads) s += glb(part)[(int64_t)stride * i + k];
    return block_sum_256(s, sh);
}

// The operator values of a descriptor as the launch's value source.
template <typename TV>
__device__ __forceinline__ Vals<TV> desc_vals(const ProbDesc& D) {
    if constexpr (std::is_same<TV, SplitVal>::value) return Vals<SplitVal>{(const float*)D.sval, D.sval16};
    else return Vals<TV>{(const TV*)D.sval};
}

// z = preconditioner(r) for the BS dofs of node row `row`
template <int BS, typename T>
__device__ __forceinline__ void apply_pc(const ProbDesc& D, int64_t row, const T (&rv)[BS], T (&zv)[BS]) {
    if (D.block_pc) {
        const SFEM_GLOBAL T* binv = glb((const T*)D.binv);
        T m[BS * BS];
#pragma unroll
        for (int v = 0; v < BS * BS; ++v) m[v] = binv[v * D.binv_stride + row];
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            T acc = 0;
#pragma unroll
            for (int i = 0; i < BS; ++i) acc = fmaT<T>(m[j * BS + i], rv[i], acc);
            zv[j] = acc;
        }
    } else {
        const SFEM_GLOBAL T* dinv = glb((const T*)D.dinv);
#pragma unroll
        for (int j = 0; j < BS; ++j) zv[j] = dinv[row * BS + j] * rv[j];
    }
}

// Update phase of one work item: x += al p, r -= al q, z = M^-1 r, one node
// per thread; returns the thread's r.z and r.r.
template <int BS, typename T, bool FOLD>
__device__ __forceinline__ void update_item(const ProbDesc& D, const WorkItem& I, T al, bool fold, double& rz,
                                            double& rr) {
    SFEM_GLOBAL double* x64 = glb(D.x64);
    SFEM_GLOBAL T* xs = glb((T*)D.xs);
    SFEM_GLOBAL T* r = glb((T*)D.r);
    SFEM_GLOBAL T* z = glb((T*)D.z);
    const SFEM_GLOBAL T* p = glb((const T*)D.p);
    const SFEM_GLOBAL T* q = glb((const T*)D.q);
    for (int64_t n = I.slice0 * 64 + threadIdx.x; n < I.slice1 * 64; n += kThreads) {
        T pv[BS], qv[BS], rv[BS], zv[BS];
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            pv[j] = p[n * BS + j];
            qv[j] = q[n * BS + j];
            rv[j] = r[n * BS + j];
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            const int64_t i = n * BS + j;
            if constexpr (FOLD) {
                T xi = xs[i];
                if (fold) {
                    x64[i] += (double)xi;
                    xi = 0;
                }
                xs[i] = fmaT<T>(al, pv[j], xi);
            } else {
                x64[i] = fma((double)al, (double)pv[j], x64[i]);
            }
            rv[j] = fmaT<T>(-al, qv[j], rv[j]);
            r[i] = rv[j];
            rr += (double)rv[j] * (double)rv[j];
        }
        apply_pc<BS, T>(D, n, rv, zv);
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            z[n * BS + j] = zv[j];
            rz += (double)rv[j] * (double)zv[j];
        }
    }
}

// FP64 residual of one work item: r = rhs - K x, z = M^-1 r; accumulates
// |r|^2, r.z, r.p and the rounding-floor scale.
template <int BS, typename T, bool FOLD>
__device__ __forceinline__ void resid_item(const ProbDesc& D, const WorkItem& I, int w, int lane, double& rr,
                                           double& rz, double& rp, double& ff) {
    const SFEM_GLOBAL double* x64 = glb((const double*)D.x64);
    const SFEM_GLOBAL T* xs = glb((const T*)D.xs);
    const SFEM_GLOBAL T* p = glb((const T*)D.p);
    const SFEM_GLOBAL double* rhs = glb(D.rhs);
    SFEM_GLOBAL T* r = glb((T*)D.r);
    SFEM_GLOBAL T* z = glb((T*)D.z);
    auto gx = [&](int64_t c) { return FOLD ? x64[c] + (double)xs[c] : x64[c]; };
    for (int64_t sl = I.slice0 + w; sl < I.slice1; sl += kThreads / 64) {
        double out[BS], aout[BS];
        sell_row_abs<BS>(D.sl_ptr, D.sl_w, D.scol, D.sval64, sl, lane, gx, out, aout);
        const int64_t row = sl * 64 + lane;
        T rt[BS], zt[BS];
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
        apply_pc<BS, T>(D, row, rt, zt);
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            z[row * BS + j] = zt[j];
            rz += (double)rt[j] * (double)zt[j];
        }
    }
}

template <typename TV, typename T, bool FOLD>
__global__ __launch_bounds__(kThreads) void table_pcg_kernel(TableLaunch a) {
    GridBarrier grid{a.bar_ws, a.bar_timeout, 0u, false};
    if (!grid.start(a.bar_grouped != 0)) return;
    __shared__ double sh[8];
    const int G = gridDim.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const ProbDesc* probs = a.probs;

    // ---- I: first launch of a problem: |rhs|^2 and the initial state ----
    if (block_any(a.nprob, [&](int i) { return probs[i].st->init_pending != 0; })) {
        for (int it = blockIdx.x; it < a.nitems; it += G) {
            const WorkItem I = a.items[it];
            const ProbDesc& D = probs[I.prob];
            if (!D.st->init_pending) continue;
            double s = 0.0;
            const SFEM_GLOBAL double* rhs = glb(D.rhs);
            for (int64_t i = I.slice0 * 64 * D.bs + threadIdx.x; i < I.slice1 * 64 * D.bs; i += kThreads)
                s = fma(rhs[i], rhs[i], s);
            const double b = block_sum_256(s, sh);
            if (threadIdx.x == 0) glb(a.part)[4 * (int64_t)it + 3] = b;
        }
        if (!grid.sync()) return;
        for (int pr = blockIdx.x; pr < a.nprob; pr += G) {
            ProbState* st = probs[pr].st;
            if (!st->init_pending) continue;
            const double bb = items_sum(a.part, 4, 3, probs[pr].item0, probs[pr].item1, sh);
            if (threadIdx.x == 0) {
                ProbState S = *st;
                S.bb = bb;
                S.tol2bb = S.rtol * S.rtol * bb;
                S.delta2 = S.delta * S.delta;
                S.maxrr = bb;
                S.rz = 0.0;
                S.rz_prev = HUGE_VAL;
                S.beta = 0.0;
                S.zcoef = 1.0;
                S.best_rr = HUGE_VAL;
                S.best_true = HUGE_VAL;
                S.rr_true = bb;
                S.need_update = 1;
                S.capped = S.max_iter <= 0 ? 1 : 0;
                S.init_pending = 0;
                *st = S;
            }
        }
        if (!grid.sync()) return;
    }

    for (int64_t k = 0; k < a.iters; ++k) {
        // ---- R: FP64 residual replacement for the problems that are due ----
        if (block_any(a.nprob, [&](int i) { return probs[i].st->need_update && !probs[i].st->done; })) {
            for (int it = blockIdx.x; it < a.nitems; it += G) {
                const WorkItem I = a.items[it];
                const ProbDesc& D = probs[I.prob];
                if (!D.st->need_update || D.st->done) continue;
                double rr = 0.0, rz = 0.0, rp = 0.0, ff = 0.0;
                if (D.bs == 3) resid_item<3, T, FOLD>(D, I, w, lane, rr, rz, rp, ff);
                else resid_item<2, T, FOLD>(D, I, w, lane, rr, rz, rp, ff);
                const double b0 = block_sum_256(rr, sh), b1 = block_sum_256(rz, sh + 4);
                const double b2 = block_sum_256(rp, sh), b3 = block_sum_256(ff, sh + 4);
                if (threadIdx.x == 0) {
                    glb(a.partr)[4 * (int64_t)it] = b0;
                    glb(a.partr)[4 * (int64_t)it + 1] = b1;
                    glb(a.partr)[4 * (int64_t)it + 2] = b2;
                    glb(a.partr)[4 * (int64_t)it + 3] = b3;
                }
            }
            if (!grid.sync()) return;
            for (int pr = blockIdx.x; pr < a.nprob; pr += G) {
                ProbState* st = probs[pr].st;
                if (!st->need_update || st->done) continue;
                const int i0 = probs[pr].item0, i1 = probs[pr].item1;
                const double trr = items_sum(a.partr, 4, 0, i0, i1, sh);
                const double trz = items_sum(a.partr, 4, 1, i0, i1, sh);
                const double trp = items_sum(a.partr, 4, 2, i0, i1, sh);
                // 16x the attainable |r|^2, (eps * || |rhs| + |K||x| ||)^2
                const double floor_rr = 16.0 * kEps64 * kEps64 * items_sum(a.partr, 4, 3, i0, i1, sh);
                if (threadIdx.x == 0) {
                    ProbState S = *st;
                    S.rr_true = trr;
                    S.floor_rr = floor_rr;
                    S.updates += 1;
                    S.last_upd = S.iters;
                    S.need_update = 0;
                    S.fold_pending = FOLD ? 1 : 0;
                    if (trr <= S.tol2bb) {
                        S.done = 1;
                    } else if (S.capped) {
                        S.done = 4;
                    } else {
                        S.beta = trz / S.rz_prev;
                        S.zcoef = 1.0 - S.beta * (trp / trz);
                        S.rz = trz;
                        S.maxrr = trr;
                        // attainable floor: no 10% gain on the best true residual
                        // within 4x of the rounding floor, or twice in a row while
                        // the recurrence ran 5x below it
                        const bool gap = S.iters > 0 && 25.0 * S.rr_rec < trr;
                        if (trr <= 0.81 * S.best_true) {
                            S.stall = 0;
                            S.best_true_it = S.iters;
                        } else if (trr <= floor_rr) {
                            S.stall += 2;
                        } else if (gap) {
                            S.stall += 1;
                        }
                        if (S.stall >= 2) S.done = 3;
                    }
                    if (trr < S.best_true) S.best_true = trr;
                    if (!isfinite(trr)) S.done = 2;
                    *st = S;
                }
            }
            if (!grid.sync()) return;
        }
        if (!block_any(a.nprob, [&](int i) { return !probs[i].st->done; })) break;

        // ---- P: p = zcoef z + beta p ----
        for (int it = blockIdx.x; it < a.nitems; it += G) {
            const WorkItem I = a.items[it];
            const ProbDesc& D = probs[I.prob];
            if (D.st->done) continue;
            const T beta = (T)D.st->beta, zc = (T)D.st->zcoef;
            SFEM_GLOBAL T* p = glb((T*)D.p);
            const SFEM_GLOBAL T* z = glb((const T*)D.z);
            for (int64_t i = I.slice0 * 64 * D.bs + threadIdx.x; i < I.slice1 * 64 * D.bs; i += kThreads)
                p[i] = fmaT<T>(beta, p[i], zc * z[i]);
        }
        if (!grid.sync()) return;

        // ---- S: q = K p and the item partials of p.q ----
        for (int it = blockIdx.x; it < a.nitems; it += G) {
            const WorkItem I = a.items[it];
            const ProbDesc& D = probs[I.prob];
            if (D.st->done) continue;
            const SFEM_GLOBAL T* p = glb((const T*)D.p);
            SFEM_GLOBAL T* q = glb((T*)D.q);
            const Vals<TV> vals = desc_vals<TV>(D);
            const int bs = D.bs;
            double dot = 0.0;
            auto gx = [&](int64_t c) { return p[c]; };
            for (int64_t sl = I.slice0 + w; sl < I.slice1; sl += kThreads / 64) {
                T out[3];
                if (bs == 3) sell_row_staged<3, TV, T, true>(D.sl_ptr, D.sl_w, D.scol, vals, sl, lane, gx, out);
                else sell_row_staged<2, TV, T, true>(D.sl_ptr, D.sl_w, D.scol, vals, sl, lane, gx, out);
                for (int j = 0; j < bs; ++j) {
                    const int64_t d = (sl * 64 + lane) * bs + j;
                    q[d] = out[j];
                    dot += (double)p[d] * (double)out[j];
                }
            }
            const double b = block_sum_256(dot, sh);
            if (threadIdx.x == 0) glb(a.part)[4 * (int64_t)it] = b;
        }
        if (!grid.sync()) return;

        // ---- U: alpha; x += alpha p; r -= alpha q; z = M^-1 r (per node) ----
        for (int it = blockIdx.x; it < a.nitems; it += G) {
            const WorkItem I = a.items[it];
            const ProbDesc& D = probs[I.prob];
            if (D.st->done) continue;
            const double pq = items_sum(a.part, 4, 0, D.item0, D.item1, sh);
            const T al = (T)(D.st->rz / pq);
            const bool fold = FOLD && D.st->fold_pending;
            double rz = 0.0, rr = 0.0;
            if (D.bs == 3) update_item<3, T, FOLD>(D, I, al, fold, rz, rr);
            else update_item<2, T, FOLD>(D, I, al, fold, rz, rr);
            const double b0 = block_sum_256(rz, sh), b1 = block_sum_256(rr, sh + 4);
            if (threadIdx.x == 0) {
                glb(a.part)[4 * (int64_t)it + 1] = b0;
                glb(a.part)[4 * (int64_t)it + 2] = b1;
            }
        }
        if (!grid.sync()) return;

        // ---- F: per-problem scalars and stopping tests ----
        for (int pr = blockIdx.x; pr < a.nprob; pr += G) {
            ProbState* st = probs[pr].st;
            if (st->done) continue;
            const int i0 = probs[pr].item0, i1 = probs[pr].item1;
            const double pq = items_sum(a.part, 4, 0, i0, i1, sh);
            const double trz = items_sum(a.part, 4, 1, i0, i1, sh);
            const double trr = items_sum(a.part, 4, 2, i0, i1, sh);
            if (threadIdx.x == 0) {
                ProbState S = *st;
                S.rz_prev = S.rz;
                S.rz = trz;
                S.beta = trz / S.rz_prev;
                S.iters += 1;
                S.fold_pending = 0;
                S.zcoef = 1.0;
                S.rr_rec = trr;
                const int64_t since = S.iters - S.last_upd;
                if (trr <= S.tol2bb || trr < S.delta2 * S.maxrr || since >= (S.best_true_it > 256 ? S.best_true_it : 256))
                    S.need_update = 1;
                // backstop for a recurrence that stops decreasing: no 2x drop of |r|
                // within max(10000, 4x the iterations it took to reach the best)
                if (trr < 0.25 * S.best_rr) {
                    S.best_rr = trr;
                    S.best_it = S.iters;
                } else if (S.iters - S.best_it > (4 * S.best_it > 10000 ? 4 * S.best_it : 10000)) {
                    S.done = 3;
                }
                if (S.iters >= S.max_iter && !S.done) {  // final FP64 residual, then stop
                    S.need_update = 1;
                    S.capped = 1;
                }
                if (!(pq > 0.0) || !isfinite(trr)) S.done = 2;
                *st = S;
            }
        }
        if (!grid.sync()) return;
    }
}

template <typename TV, typename T, bool FOLD>
void launch_table(sfem_ctx* ctx, const TableLaunch& a, int grid) {
    TableLaunch arg = a;
    void* args[] = {&arg};
    SFEM_HIP(hipLaunchCooperativeKernel((const void*)table_pcg_kernel<TV, T, FOLD>, dim3(grid), dim3(kThreads), args,
                                        0, ctx->stream));
}

template <typename TV, typename T, bool FOLD>
int resident_blocks() {
    int per_cu = 0;
    SFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, table_pcg_kernel<TV, T, FOLD>, kThreads, 0));
    return per_cu;
}

}  // namespace

size_t prec_vec_bytes(int precision) { return precision == SFEM_PREC_MIXED ? sizeof(float) : sizeof(double); }

int table_max_blocks_per_cu(int precision) {
    switch (precision) {
        case SFEM_PREC_FP64: return resident_blocks<double, double, false>();
        case SFEM_PREC_MIXED: return resident_blocks<float, float, true>();
        case SFEM_PREC_FP32_MATRIX: return resident_blocks<float, double, false>();
        case SFEM_PREC_SPLIT: return resident_blocks<SplitVal, double, false>();
        default: fail(SFEM_ERR_ARGUMENT, "unknown precision");
    }
}

void table_launch(sfem_ctx* ctx, int precision, const TableLaunch& a0, int grid) {
    if (a0.nitems <= 0 || a0.nprob <= 0) return;
    if (!ctx->bar_ws) ctx->bar_ws = static_cast<unsigned*>(ctx->alloc->alloc(sizeof(unsigned) * kBarWords));
    TableLaunch a = a0;
    a.bar_ws = ctx->bar_ws;
    a.bar_timeout = ctx->bar_timeout_ticks;
    a.bar_grouped = ctx->bar_grouped ? 1 : 0;
    SFEM_HIP(hipMemsetAsync(a.bar_ws, 0, sizeof(unsigned) * kBarWords, ctx->stream));
    switch (precision) {
        case SFEM_PREC_FP64: launch_table<double, double, false>(ctx, a, grid); break;
        case SFEM_PREC_MIXED: launch_table<float, float, true>(ctx, a, grid); break;
        case SFEM_PREC_FP32_MATRIX: launch_table<float, double, false>(ctx, a, grid); break;
        case SFEM_PREC_SPLIT: launch_table<SplitVal, double, false>(ctx, a, grid); break;
        default: fail(SFEM_ERR_ARGUMENT, "unknown precision");
    }
    unsigned flags[2] = {0u, 0u};
    d2h(ctx, flags, a.bar_ws + kBarFlags, sizeof(flags));
    ctx->bar_last_grouped = a.bar_grouped && flags[0] == 0u;
    if (flags[1]) fail(SFEM_ERR_DEVICE, "persistent solver: a grid barrier wait exceeded its time limit");
}

void table_fold(sfem_ctx* ctx, double* x64, void* xs, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fold_all_kernel, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, ctx->stream,
                       x64, (float*)xs, n);
    SFEM_LAUNCH_CHECK();
}

int table_grid(sfem_ctx* ctx, int precision, int blocks_per_cu) {
    const int maxb = table_max_blocks_per_cu(precision);
    if (maxb < 1) fail(SFEM_ERR_DEVICE, "persistent solver kernel cannot be resident");
    const int b = blocks_per_cu > 0 ? std::min(blocks_per_cu, maxb) : std::min(maxb, 2);
    return b * ctx->n_cu;
}

void batch_build(sfem_ctx* ctx, sfem_batch* B) {
    const sfem_csr* K = B->K;
    const sfem_bc* bc = B->bc;
    const int bs = K->bs;
    if (bs != B->ndof || bs < 2 || bs > 3) fail(SFEM_ERR_UNSUPPORTED, "batch: needs node-blocked 2D or 3D dofs");
    const int64_t nunits = K->nunits;
    B->bs = bs;
    B->n = K->n;
    hipEvent_t e0, e1;
    SFEM_HIP(hipEventCreate(&e0));
    SFEM_HIP(hipEventCreate(&e1));
    SFEM_HIP(hipEventRecord(e0, ctx->stream));
    DevBuf<uint32_t> ufree(ctx, (size_t)nunits), upos(ctx, (size_t)nunits), utot(ctx, 1);
    const dim3 gu(grid_for(nunits, 256));
    hipLaunchKernelGGL(unit_free_kernel, gu, dim3(256), 0, ctx->stream, bc->fixed.get(), nunits, bs, ufree.get());
    SFEM_LAUNCH_CHECK();
    exclusive_scan_u32_dev(ctx, ufree.get(), upos.get(), nunits, utot.get());
    const int64_t np = B->nprob;
    DevBuf<int64_t> doff(ctx, (size_t)np + 1);
    h2d(ctx, doff.get(), B->node_off.data(), sizeof(int64_t) * (np + 1));
    DevBuf<uint32_t> dfirst(ctx, (size_t)np + 1);
    hipLaunchKernelGGL(gather_pos_kernel, dim3(grid_for(np + 1, 256)), dim3(256), 0, ctx->stream, upos.get(),
                       doff.get(), np + 1, nunits, utot.get(), dfirst.get());
    SFEM_LAUNCH_CHECK();
    std::vector<uint32_t> first((size_t)np + 1);
    d2h(ctx, first.data(), dfirst.get(), sizeof(uint32_t) * (np + 1));
    // padded layout: each problem starts on a new slice
    std::vector<int64_t> pad_base((size_t)np + 1);
    B->hprobs.resize((size_t)np);
    int64_t base = 0;
    for (int64_t p = 0; p < np; ++p) {
        const int64_t nfu = (int64_t)first[p + 1] - (int64_t)first[p];
        const int64_t ns = (nfu + 63) / 64;
        pad_base[p] = base;
        BatchProb& P = B->hprobs[(size_t)p];
        P.slice0 = base / 64;
        P.slice1 = P.slice0 + ns;
        P.n_free = nfu * bs;
        base += ns * 64;
    }
    pad_base[(size_t)np] = base;
    if (base * bs >= (int64_t(1) << 31)) fail(SFEM_ERR_UNSUPPORTED, "batch: 2^31 or more padded dofs");
    if (base == 0) fail(SFEM_ERR_ARGUMENT, "batch: no free dofs");
    B->npad = base * bs;
    B->nslices = base / 64;
    DevBuf<int64_t> dpb(ctx, (size_t)np + 1);
    h2d(ctx, dpb.get(), pad_base.data(), sizeof(int64_t) * (np + 1));
    B->pad_of.reset(ctx, (size_t)nunits);
    hipLaunchKernelGGL(pad_map_kernel, gu, dim3(256), 0, ctx->stream, ufree.get(), upos.get(), doff.get(),
                       dfirst.get(), dpb.get(), np, nunits, B->pad_of.get());
    SFEM_LAUNCH_CHECK();
    // slice widths; blen < 0 marks padding rows
    DevBuf<int32_t> blen(ctx, (size_t)base);
    memset_d(ctx, blen.get(), 0xFF, sizeof(int32_t) * base);
    hipLaunchKernelGGL(pad_blen_kernel, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                       B->pad_of.get(), nunits, blen.get());
    SFEM_LAUNCH_CHECK();
    B->sl_w.reset(ctx, (size_t)B->nslices);
    B->sl_ptr.reset(ctx, (size_t)B->nslices + 1);
    DevBuf<int64_t> cnt(ctx, (size_t)B->nslices + 1);
    memset_d(ctx, cnt.get(), 0, sizeof(int64_t) * (B->nslices + 1));
    hipLaunchKernelGGL(slice_w_kernel, dim3(grid_for(B->nslices, 256)), dim3(256), 0, ctx->stream, blen.get(),
                       B->nslices, B->sl_w.get(), cnt.get());
    SFEM_LAUNCH_CHECK();
    const int64_t slots = exclusive_scan_i64(ctx, cnt.get(), B->sl_ptr.get(), B->nslices + 1);
    B->sell_blocks = slots;
    B->scol.reset(ctx, (size_t)slots);
    B->sval64.reset(ctx, (size_t)slots * bs * bs);
    B->sval32.reset(ctx, (size_t)slots * bs * bs);
    B->sval16.reset(ctx, (size_t)slots * bs * bs);
    B->binv64.reset(ctx, (size_t)base * bs * bs);
    B->binv32.reset(ctx, (size_t)base * bs * bs);
    const dim3 gr(grid_for(base, 256));
    hipLaunchKernelGGL(binv_identity_kernel, gr, dim3(256), 0, ctx->stream, B->binv64.get(), B->binv32.get(), base,
                       bs);
    SFEM_LAUNCH_CHECK();
    if (bs == 3) {
        hipLaunchKernelGGL(sell_fill_pad_kernel<3>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), B->pad_of.get(), nunits, B->sl_ptr.get(), B->sl_w.get(), B->scol.get(),
                           B->sval64.get(), B->sval32.get(), B->sval16.get(), base, B->binv64.get(),
                           B->binv32.get(), bc->fixed.get());
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(sell_pad_gap_kernel<3>, gr, dim3(256), 0, ctx->stream, blen.get(), base, B->sl_ptr.get(),
                           B->sl_w.get(), B->scol.get(), B->sval64.get(), B->sval32.get(), B->sval16.get());
    } else {
        hipLaunchKernelGGL(sell_fill_pad_kernel<2>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), B->pad_of.get(), nunits, B->sl_ptr.get(), B->sl_w.get(), B->scol.get(),
                           B->sval64.get(), B->sval32.get(), B->sval16.get(), base, B->binv64.get(),
                           B->binv32.get(), bc->fixed.get());
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(sell_pad_gap_kernel<2>, gr, dim3(256), 0, ctx->stream, blen.get(), base, B->sl_ptr.get(),
                           B->sl_w.get(), B->scol.get(), B->sval64.get(), B->sval32.get(), B->sval16.get());
    }
    SFEM_LAUNCH_CHECK();
    B->rhs.reset(ctx, (size_t)B->npad);
    B->x.reset(ctx, (size_t)B->npad);
    B->dinv64.reset(ctx, (size_t)B->npad);
    B->dinv32.reset(ctx, (size_t)B->npad);
    hipLaunchKernelGGL(fill_pad_vectors_kernel, dim3(grid_for(B->npad, 256)), dim3(256), 0, ctx->stream,
                       B->rhs.get(), B->dinv64.get(), B->dinv32.get(), B->x.get(), B->npad);
    SFEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(pad_rhs_kernel, dim3(grid_for(B->n, 256)), dim3(256), 0, ctx->stream, K->bcolptr.get(),
                       K->browidx.get(), K->bval.get(), B->pad_of.get(), bc->fixed.get(), bc->qd.get(), bc->f.get(),
                       B->n, bs, B->rhs.get(), B->dinv64.get(), B->dinv32.get());
    SFEM_LAUNCH_CHECK();
    SFEM_HIP(hipEventRecord(e1, ctx->stream));
    sync(ctx);
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    B->t_eliminate_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void batch_items(const sfem_batch* B, int item_slices, std::vector<WorkItem>* items,
                 std::vector<std::pair<int, int>>* ranges) {
    items->clear();
    ranges->clear();
    for (size_t p = 0; p < B->hprobs.size(); ++p) {
        const BatchProb& P = B->hprobs[p];
        const int i0 = (int)items->size();
        for (int64_t s = P.slice0; s < P.slice1; s += item_slices) {
            WorkItem I{};
            I.prob = (int32_t)p;
            I.slice0 = s;
            I.slice1 = std::min(P.slice1, s + item_slices);
            items->push_back(I);
        }
        ranges->push_back({i0, (int)items->size()});
    }
}

void batch_ensure_work(sfem_ctx* ctx, sfem_batch* B, int precision) {
    const size_t vb = prec_vec_bytes(precision);
    if (B->work_vb != vb) {
        B->wr.reset(ctx, vb * B->npad);
        B->wz.reset(ctx, vb * B->npad);
        B->wp.reset(ctx, vb * B->npad);
        B->wq.reset(ctx, vb * B->npad);
        B->work_vb = vb;
    }
    if (precision == SFEM_PREC_MIXED && B->wxs.size() != sizeof(float) * B->npad)
        B->wxs.reset(ctx, sizeof(float) * B->npad);
}

ProbDesc batch_desc(sfem_batch* B, int precision, int preconditioner, ProbState* st, int item0, int item1) {
    ProbDesc D{};
    D.sl_ptr = B->sl_ptr.get();
    D.sl_w = B->sl_w.get();
    D.scol = B->scol.get();
    D.sval = precision == SFEM_PREC_FP64 ? (const void*)B->sval64.get() : (const void*)B->sval32.get();
    D.sval16 = B->sval16.get();
    D.sval64 = B->sval64.get();
    D.rhs = B->rhs.get();
    D.x64 = B->x.get();
    D.xs = precision == SFEM_PREC_MIXED ? (void*)B->wxs.get() : nullptr;
    D.r = B->wr.get();
    D.z = B->wz.get();
    D.p = B->wp.get();
    D.q = B->wq.get();
    D.dinv = precision == SFEM_PREC_MIXED ? (const void*)B->dinv32.get() : (const void*)B->dinv64.get();
    D.binv = precision == SFEM_PREC_MIXED ? (const void*)B->binv32.get() : (const void*)B->binv64.get();
    D.binv_stride = B->nslices * 64;
    D.block_pc = preconditioner == SFEM_PC_BLOCK_JACOBI ? 1 : 0;
    D.st = st;
    D.bs = B->bs;
    D.item0 = item0;
    D.item1 = item1;
    return D;
}

ProbState initial_state(const sfem_solve_opts& o) {
    ProbState S{};
    S.rtol = o.rtol;
    S.delta = o.reliable_delta;
    S.max_iter = o.max_iter;
    S.init_pending = 1;
    return S;
}

void batch_reset(sfem_ctx* ctx, sfem_batch* B, int precision, const sfem_solve_opts& o, ProbState* d_st) {
    batch_ensure_work(ctx, B, precision);
    const size_t vb = prec_vec_bytes(precision);
    memset_d(ctx, B->wp.get(), 0, vb * B->npad);
    if (precision == SFEM_PREC_MIXED) memset_d(ctx, B->wxs.get(), 0, sizeof(float) * B->npad);
    if (o.x0_zero) memset_d(ctx, B->x.get(), 0, sizeof(double) * B->npad);
    if (d_st) {
        std::vector<ProbState> hst((size_t)B->nprob, initial_state(o));
        h2d(ctx, d_st, hst.data(), sizeof(ProbState) * B->nprob);
    }
}

double batch_run(sfem_ctx* ctx, sfem_batch* B, const sfem_solve_opts& o, int item_slices, int blocks_per_cu,
                 int64_t* launches) {
    const int prec = o.precision;
    if (prec != SFEM_PREC_FP64 && prec != SFEM_PREC_MIXED && prec != SFEM_PREC_FP32_MATRIX && prec != SFEM_PREC_SPLIT)
        fail(SFEM_ERR_ARGUMENT, "batch solve: unknown precision");
    if (o.preconditioner != SFEM_PC_JACOBI && o.preconditioner != SFEM_PC_BLOCK_JACOBI)
        fail(SFEM_ERR_ARGUMENT, "batch solve: unknown preconditioner");
    if (item_slices < 1) fail(SFEM_ERR_ARGUMENT, "batch solve: item_slices must be >= 1");
    if (o.max_iter < 0) fail(SFEM_ERR_ARGUMENT, "batch solve: max_iter must be >= 0");
    if (B->st.size() != (size_t)B->nprob) B->st.reset(ctx, (size_t)B->nprob);
    batch_reset(ctx, B, prec, o, B->st.get());
    std::vector<WorkItem> items;
    std::vector<std::pair<int, int>> ranges;
    batch_items(B, item_slices, &items, &ranges);
    std::vector<ProbDesc> descs((size_t)B->nprob);
    for (int64_t p = 0; p < B->nprob; ++p)
        descs[(size_t)p] = batch_desc(B, prec, o.preconditioner, B->st.get() + p, ranges[(size_t)p].first,
                                      ranges[(size_t)p].second);
    DevBuf<ProbDesc> ddesc(ctx, descs.size());
    DevBuf<WorkItem> ditems(ctx, items.size());
    DevBuf<double> part(ctx, 4 * items.size()), partr(ctx, 4 * items.size());
    h2d(ctx, ddesc.get(), descs.data(), sizeof(ProbDesc) * descs.size());
    h2d(ctx, ditems.get(), items.data(), sizeof(WorkItem) * items.size());
    TableLaunch a{};
    a.probs = ddesc.get();
    a.nprob = (int)B->nprob;
    a.items = ditems.get();
    a.nitems = (int)items.size();
    a.part = part.get();
    a.partr = partr.get();
    const int grid = table_grid(ctx, prec, blocks_per_cu);
    // One launch holds the whole solve unless a per-launch iteration budget
    // (check_every > 0) splits it; the state lives in memory either way.  A
    // problem needs at most max_iter + 1 loop passes (the last one is the
    // final FP64 residual of an iteration-capped problem).
    const int64_t need = o.max_iter + 2;
    const int64_t budget = o.check_every > 0 ? std::min<int64_t>(o.check_every, need) : need;
    hipEvent_t e0, e1;
    SFEM_HIP(hipEventCreate(&e0));
    SFEM_HIP(hipEventCreate(&e1));
    SFEM_HIP(hipEventRecord(e0, ctx->stream));
    std::vector<ProbState> hst((size_t)B->nprob);
    int64_t nl = 0, run = 0;
    while (true) {
        a.iters = budget;
        table_launch(ctx, prec, a, grid);
        ++nl;
        run += budget;
        d2h(ctx, hst.data(), B->st.get(), sizeof(ProbState) * B->nprob);
        const bool all = std::all_of(hst.begin(), hst.end(), [](const ProbState& s) { return s.done != 0; });
        if (all || run >= need) break;
    }
    if (prec == SFEM_PREC_MIXED) table_fold(ctx, B->x.get(), B->wxs.get(), B->npad);
    SFEM_HIP(hipEventRecord(e1, ctx->stream));
    sync(ctx);
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (launches) *launches = nl;
    return ms;
}

void batch_solution(sfem_ctx* ctx, const sfem_batch* B, double* q) {
    DevBuf<double> dq(ctx, (size_t)B->n);
    hipLaunchKernelGGL(batch_recombine_kernel, dim3(grid_for(B->n, 256)), dim3(256), 0, ctx->stream,
                       B->pad_of.get(), B->fixed.get(), B->x.get(), B->qd.get(), B->n, B->bs, dq.get());
    SFEM_LAUNCH_CHECK();
    d2h(ctx, q, dq.get(), sizeof(double) * B->n);
}

// The padded operator of problems already on the device as one
// concatenated mesh (node ids offset per problem) with its loads and
// constraints.  Takes ownership of mesh and bc (both are destroyed once the
// operator exists, or on failure).  nfree: free dofs per problem.
sfem_batch* batch_from_device(sfem_ctx* ctx, int64_t nprob, sfem_mesh* mesh, sfem_bc* bc,
                              const std::vector<int64_t>& node_offsets, const std::vector<int64_t>& nfree) {
    std::unique_ptr<sfem_batch> B(new sfem_batch());
    B->ctx = ctx;
    B->mesh = mesh;
    B->bc = bc;
    B->nprob = nprob;
    B->ndim = mesh->ndim;
    B->ndof = mesh->ndof;
    B->node_off = node_offsets;
    const int64_t N = mesh->n_nodes;
    const int n_dof = mesh->ndof;
    sfem_csr* K = nullptr;
    int s = sfem_assemble_mesh(ctx, mesh, SFEM_ASM_DEFAULT, &K);
    if (s) fail(s, ctx->last_error);
    B->K = K;
    B->t_assemble_ms = K->t_total;
    if (K->n != N * n_dof || K->bs != n_dof)
        fail(SFEM_ERR_ARGUMENT, "batch_create: every node must belong to an element");
    if (bc->n != K->n) fail(SFEM_ERR_ARGUMENT, "batch_create: loads do not match the mesh");
    batch_build(ctx, B.get());
    B->qd.reset(ctx, (size_t)B->n);
    SFEM_HIP(hipMemcpyAsync(B->qd.get(), bc->qd.get(), sizeof(double) * B->n, hipMemcpyDeviceToDevice, ctx->stream));
    B->fixed.reset(ctx, (size_t)B->n);
    SFEM_HIP(hipMemcpyAsync(B->fixed.get(), bc->fixed.get(), (size_t)B->n, hipMemcpyDeviceToDevice, ctx->stream));
    // free dofs per problem (a partially fixed node keeps only its free dofs)
    for (int64_t p = 0; p < nprob; ++p) B->hprobs[(size_t)p].n_free = nfree[(size_t)p];
    sync(ctx);
    sfem_csr_destroy(B->K);
    B->K = nullptr;
    sfem_bc_destroy(B->bc);
    B->bc = nullptr;
    sfem_mesh_destroy(B->mesh);
    B->mesh = nullptr;
    return B.release();
}

sfem_batch* batch_create(sfem_ctx* ctx, int64_t nprob, int ndim, int n_dof, const double* node_list,
                         const int64_t* node_offsets, const int64_t* ien, int nen, const int64_t* elem_offsets,
                         const double* young, const double* nu, const double* f, const uint8_t* fixed,
                         const double* q_d) {
    if (nprob <= 0 || !node_list || !node_offsets || !ien || !elem_offsets || !young || !nu || !fixed)
        fail(SFEM_ERR_ARGUMENT, "batch_create: NULL array or empty batch");
    if (n_dof != ndim || (ndim != 2 && ndim != 3) || nen != (1 << ndim))
        fail(SFEM_ERR_UNSUPPORTED, "batch_create: Q4 (2D) or Hex8 (3D) vector problems only");
    if (node_offsets[0] != 0 || elem_offsets[0] != 0)
        fail(SFEM_ERR_ARGUMENT, "batch_create: offsets must start at 0");
    for (int64_t p = 0; p < nprob; ++p)
        if (node_offsets[p + 1] <= node_offsets[p] || elem_offsets[p + 1] <= elem_offsets[p])
            fail(SFEM_ERR_ARGUMENT, "batch_create: every problem needs nodes and elements (problem " +
                                        std::to_string(p) + ")");
    const int64_t N = node_offsets[nprob], E = elem_offsets[nprob];
    std::vector<int64_t> gien((size_t)(E * nen)), gid((size_t)(N * n_dof));
    for (int64_t p = 0; p < nprob; ++p)
        for (int64_t e = elem_offsets[p]; e < elem_offsets[p + 1]; ++e)
            for (int a = 0; a < nen; ++a) {
                const int64_t v = ien[(int64_t)a * E + e];
                if (v < 1 || v > node_offsets[p + 1] - node_offsets[p])
                    fail(SFEM_ERR_BOUNDS, "batch_create: IEN entry outside its problem's nodes (problem " +
                                              std::to_string(p) + ")");
                gien[(size_t)((int64_t)a * E + e)] = v + node_offsets[p];
            }
    for (int64_t m = 0; m < N; ++m)
        for (int l = 0; l < n_dof; ++l) gid[(size_t)((int64_t)l * N + m)] = (int64_t)n_dof * m + l + 1;
    sfem_mesh* mesh = nullptr;
    int s = sfem_mesh_create(ctx, ndim, SFEM_Q1, n_dof, node_list, N, gien.data(), E, nen, E, gid.data(), N, young, E,
                             0.3, &mesh);
    if (s) fail(s, ctx->last_error);
    try {
        mesh->nu_e.reset(ctx, (size_t)E);
        h2d(ctx, mesh->nu_e.get(), nu, sizeof(double) * E);
    } catch (...) {
        sfem_mesh_destroy(mesh);
        throw;
    }
    sfem_bc* bc = nullptr;
    s = sfem_bc_create(ctx, N * n_dof, f, fixed, q_d, &bc);
    if (s) {
        sfem_mesh_destroy(mesh);
        fail(s, ctx->last_error);
    }
    std::vector<int64_t> noff(node_offsets, node_offsets + nprob + 1), nfree((size_t)nprob, 0);
    for (int64_t p = 0; p < nprob; ++p)
        for (int64_t d = node_offsets[p] * n_dof; d < node_offsets[p + 1] * n_dof; ++d)
            nfree[(size_t)p] += fixed[d] ? 0 : 1;
    return batch_from_device(ctx, nprob, mesh, bc, noff, nfree);
}

int64_t batch_device_bytes(const sfem_batch* B, int precision) {
    const int64_t b2 = (int64_t)B->bs * B->bs;
    const int64_t vb = (int64_t)prec_vec_bytes(precision);
    return B->sell_blocks * (4 + b2 * 14) + B->nslices * 12 + B->npad * (8 * 3 + 4 + 4 * vb + (vb == 4 ? 4 : 0)) +
           B->nslices * 64 * b2 * 12 + (int64_t)B->pad_of.size() * 4 + B->n * 9;
}

}  // namespace sfem

extern "C" {

int sfem_batch_create(sfem_ctx* ctx, int64_t nprob, int ndim, int n_dof, const double* node_list,
                      const int64_t* node_offsets, const int64_t* ien, int nen, const int64_t* elem_offsets,
                      const double* young, const double* nu, const double* f, const uint8_t* fixed,
                      const double* q_d, sfem_batch** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out) { ctx->last_error = "batch_create: NULL out"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_batch* B = nullptr;
    const int st = guard(ctx, [&] {
        B = batch_create(ctx, nprob, ndim, n_dof, node_list, node_offsets, ien, nen, elem_offsets, young, nu, f,
                         fixed, q_d);
    });
    if (st) return st;
    ctx_retain(ctx);
    *out = B;
    return SFEM_OK;
}

int sfem_batch_info(const sfem_batch* B, int64_t* n_total, int64_t* n_padded, int64_t* slices, double* assemble_ms,
                    double* eliminate_ms) {
    if (!B) return SFEM_ERR_ARGUMENT;
    if (n_total) *n_total = B->n;
    if (n_padded) *n_padded = B->npad;
    if (slices) *slices = B->nslices;
    if (assemble_ms) *assemble_ms = B->t_assemble_ms;
    if (eliminate_ms) *eliminate_ms = B->t_eliminate_ms;
    return SFEM_OK;
}

int sfem_batch_solve(sfem_ctx* ctx, sfem_batch* B, const sfem_solve_opts* opts, int item_slices,
                     int blocks_per_cu, sfem_batch_problem_stats* per, sfem_batch_stats* stats) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!B) { ctx->last_error = "batch_solve: NULL batch"; return SFEM_ERR_ARGUMENT; }
    sfem_solve_opts o;
    sfem_solve_opts_default(&o);
    if (opts) o = *opts;
    return guard(ctx, [&] {
        int64_t nl = 0;
        const double ms = batch_run(ctx, B, o, item_slices, blocks_per_cu, &nl);
        std::vector<ProbState> hst((size_t)B->nprob);
        d2h(ctx, hst.data(), B->st.get(), sizeof(ProbState) * B->nprob);
        int64_t bad = 0, max_it = 0, sum_it = 0;
        for (int64_t p = 0; p < B->nprob; ++p) {
            const ProbState& S = hst[(size_t)p];
            if (per) {
                per[p].iterations = S.iters;
                per[p].fp64_updates = S.updates;
                per[p].rel_residual = S.bb > 0 ? std::sqrt(S.rr_true / S.bb) : 0.0;
                per[p].n_free = B->hprobs[(size_t)p].n_free;
                per[p].stop_reason = S.done == 0 ? 4 : S.done;
                per[p].floor_residual = S.bb > 0 ? std::sqrt(S.floor_rr / 16.0 / S.bb) : 0.0;
            }
            max_it = std::max<int64_t>(max_it, S.iters);
            sum_it += S.iters;
            if (S.done != 1) ++bad;
        }
        if (stats) {
            stats->solve_ms = ms;
            stats->launches = nl;
            stats->max_iterations = max_it;
            stats->sum_iterations = sum_it;
            stats->not_converged = bad;
            stats->grid = table_grid(ctx, o.precision, blocks_per_cu);
            stats->device_bytes = batch_device_bytes(B, o.precision);
            stats->barrier_grouped = ctx->bar_last_grouped ? 1 : 0;
        }
        if (bad) fail(SFEM_ERR_NOT_CONVERGED, "batch_solve: " + std::to_string(bad) + " problems did not converge");
    });
}

int sfem_batch_solution(sfem_ctx* ctx, const sfem_batch* B, double* q) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!B || !q) { ctx->last_error = "batch_solution: NULL argument"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { batch_solution(ctx, B, q); });
}

int sfem_batch_max_blocks_per_cu(sfem_ctx* ctx, int precision, int* blocks) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!blocks) { ctx->last_error = "NULL blocks"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { *blocks = table_max_blocks_per_cu(precision); });
}

void sfem_batch_destroy(sfem_batch* B) {
    if (!B) return;
    sfem_ctx* ctx = B->ctx;
    delete B;
    ctx_release(ctx);
}

}  // extern "C"
