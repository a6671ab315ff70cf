// eliminate.hip -- Dirichlet elimination K_ff = C' K C, rhs = C'(f - K q_d)
// and recombination q = q_d + C q_f on the device, plus the solver formats.
//
// Reference: smearFEM examples/vector3D.jl:151,170 (C = I[:, free]), :315-316
// (K_free = C' K_bar C, which picks the submatrix K[free, free] exactly),
// :320 (rhs = C'(-K_bar q_d); a load vector f adds in), :322 (q = q_d + C q_f).
//
// K arrives as a block CSC (column unit c, blocks (r, bs x bs)).  Because K is
// bitwise symmetric, column c of K is row c, so every format below is built
// by walking columns and treating them as rows:
//   * scalar CSR of K_ff, int32 columns (always);
//   * when bs > 1 and every node is fully fixed or fully free: a node-block
//     sliced format ("SELL"): 64 block rows per slice, one lane per block row,
//     slot k of all 64 rows stored contiguously, values as bs*bs planes of 64.
#include "block_ops.hpp"
#include "sfem_objects.hpp"

namespace sfem {

namespace {

__global__ void free_flags_kernel(const uint8_t* __restrict__ fixed, int64_t n, uint32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = fixed[i] ? 0u : 1u;
}

__global__ void free_map_kernel(const uint8_t* __restrict__ fixed, const uint32_t* __restrict__ pos, int64_t n,
                                int32_t* __restrict__ free_of) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) free_of[i] = fixed[i] ? -1 : (int32_t)pos[i];
}

// per unit: 1 if all bs dofs free; clears *complete if a unit is partly fixed
// unit_free[u] = 1 when any dof of node u is free (the node is a block row);
// *partial = 1 when some node has both fixed and free dofs.
__global__ void unit_status_kernel(const uint8_t* __restrict__ fixed, int64_t nunits, int bs,
                                   uint32_t* __restrict__ unit_free, int* partial) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int c = 0;
    for (int i = 0; i < bs; ++i) c += fixed[u * bs + i] ? 1 : 0;
    unit_free[u] = c < bs ? 1u : 0u;
    if (c != 0 && c != bs) atomicOr(partial, 1);
}

// Padded block vectors <-> compact free dofs: pad_of[row*bs + i] is the free
// index of dof i of block row `row`, or -1 for a fixed dof of a partially
// fixed node (kept at exactly 0 by a unit diagonal and zero couplings).
__global__ void pad_map_kernel(const int32_t* __restrict__ ufree_of, const int32_t* __restrict__ free_of,
                               int64_t nunits, int bs, int32_t* __restrict__ pad_of) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const int32_t row = ufree_of[u];
    if (row < 0) return;
    for (int i = 0; i < bs; ++i) pad_of[(int64_t)row * bs + i] = free_of[u * bs + i];
}

__global__ void unit_fixed_kernel(const uint32_t* __restrict__ ufree, uint8_t* __restrict__ ufixed, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) ufixed[i] = ufree[i] ? 0 : 1;
}

__global__ void csr_count_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                 const int32_t* __restrict__ free_of, int64_t nunits, int bs,
                                 uint32_t* __restrict__ rowcnt) {
    const int64_t cu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cu >= nunits) return;
    uint32_t cnt = 0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t r = browidx[u];
        for (int i = 0; i < bs; ++i) cnt += free_of[r * bs + i] < 0 ? 0u : 1u;
    }
    for (int j = 0; j < bs; ++j) {
        const int32_t row = free_of[cu * bs + j];
        if (row >= 0) rowcnt[row] = cnt;
    }
}

// One thread per dof of K (a column of K = a row of K_ff when free).
__global__ void csr_fill_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                const double* __restrict__ bval, const int32_t* __restrict__ free_of,
                                const double* __restrict__ qd, const double* __restrict__ f, int64_t n, int bs,
                                const uint32_t* __restrict__ rowptr, int32_t* __restrict__ col,
                                double* __restrict__ val, double* __restrict__ rhs, double* __restrict__ dinv) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    if (row < 0) return;
    const int64_t cu = s / bs;
    const int j = (int)(s - cu * bs);
    const int bs2 = bs * bs;
    uint32_t pos = rowptr[row];
    double lift = 0.0, diag = 0.0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t r = browidx[u];
        for (int i = 0; i < bs; ++i) {
            const int64_t d = r * bs + i;
            const double v = bval[u * bs2 + i * bs + j];
            const int32_t c = free_of[d];
            if (c >= 0) {
                col[pos] = c;
                val[pos] = v;
                ++pos;
                if (d == s) diag = v;
            } else {
                lift += v * qd[d];  // (K q_d)[s], columns ascending as the CSC matvec sums
            }
        }
    }
    rhs[row] = f[s] - lift;
    dinv[row] = diag != 0.0 ? 1.0 / diag : 1.0;
}

// rhs and 1/diag of csr_fill_kernel without the CSR arrays (same loop, same
// order, so the same bits); the arrays are built on first use (ensure_csr).
// rhs = f on the free dofs (no prescribed value is nonzero)
__global__ void rhs_gather_kernel(const int32_t* __restrict__ free_of, const double* __restrict__ f, int64_t n,
                                  double* __restrict__ rhs) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    if (row >= 0) rhs[row] = f[s];
}

// point-Jacobi 1 / K_ss of every free dof (1 where the diagonal is zero)
__global__ void dinv_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                            const double* __restrict__ bval, const int32_t* __restrict__ free_of, int64_t n, int bs,
                            double* __restrict__ dinv) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    if (row < 0) return;
    const int64_t cu = s / bs;
    const int j = (int)(s - cu * bs);
    double diag = 0.0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u)
        if (browidx[u] == cu) diag = bval[u * bs * bs + j * bs + j];
    dinv[row] = diag != 0.0 ? 1.0 / diag : 1.0;
}

// Inverse node diagonal blocks of the block rows (and, for padded vectors,
// the point-Jacobi 1/diag with a unit diagonal on fixed dofs): the diagonal
// block masked as sell_fill_kernel masks it, so both give the same values.
template <int BS>
__global__ void diag_blocks_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                   const double* __restrict__ bval, const int32_t* __restrict__ ufree_of,
                                   const int32_t* __restrict__ free_of, int64_t nunits, int64_t nfb,
                                   double* __restrict__ binv64, float* __restrict__ binv32,
                                   double* __restrict__ pdinv64) {
    constexpr int B2 = BS * BS;
    const int64_t cu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cu >= nunits) return;
    const int32_t row = ufree_of[cu];
    if (row < 0) return;
    // assembled columns list their rows in ascending order: binary search for
    // the diagonal block, a scan when that misses (an uploaded, unsorted CSC)
    int64_t lo = bcolptr[cu], hi = bcolptr[cu + 1], ud = -1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        const int64_t rm = browidx[mid];
        if (rm == cu) {
            ud = mid;
            break;
        }
        if (rm < cu) lo = mid + 1;
        else hi = mid;
    }
    if (ud < 0)
        for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u)
            if (browidx[u] == cu) ud = u;
    double m[B2];
#pragma unroll
    for (int j = 0; j < BS; ++j)
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            const bool masked = free_of[cu * BS + j] < 0 || free_of[cu * BS + i] < 0;
            m[j * BS + i] = masked || ud < 0 ? (i == j && masked ? 1.0 : 0.0) : bval[ud * B2 + i * BS + j];
        }
    if (pdinv64)
#pragma unroll
        for (int j = 0; j < BS; ++j) pdinv64[(int64_t)row * BS + j] = m[j * BS + j] != 0.0 ? 1.0 / m[j * BS + j] : 1.0;
    double inv[B2];
    invert_block<BS>(m, inv);
#pragma unroll
    for (int v = 0; v < B2; ++v) {
        binv64[v * nfb + row] = inv[v];
        binv32[v * nfb + row] = (float)inv[v];
    }
}

// rhs[row] = f[s] - (K q_d)[s] over the fixed columns, the sum csr_fill_kernel
// forms (same code, same order), for new loads / prescribed values.
__global__ void rhs_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                           const double* __restrict__ bval, const int32_t* __restrict__ free_of,
                           const double* __restrict__ qd, const double* __restrict__ f, int64_t n, int bs,
                           double* __restrict__ rhs) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    if (row < 0) return;
    const int64_t cu = s / bs;
    const int j = (int)(s - cu * bs);
    const int bs2 = bs * bs;
    double lift = 0.0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t r = browidx[u];
        for (int i = 0; i < bs; ++i) {
            const int64_t d = r * bs + i;
            const double v = bval[u * bs2 + i * bs + j];
            if (free_of[d] < 0) lift += v * qd[d];  // (K q_d)[s], columns ascending as the CSC matvec sums
        }
    }
    rhs[row] = f[s] - lift;
}

// rhs_many[row*kp + j] = f_j[s] - (K q_d_j)[s] for k columns (q_d, f: n x k
// column-major, NULL = zeros); per column the same sum as rhs_kernel.
__global__ void rhs_many_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                const double* __restrict__ bval, const int32_t* __restrict__ free_of,
                                const double* __restrict__ qd, const double* __restrict__ f, int64_t n, int bs, int k,
                                int kp, double* __restrict__ rhs) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t row = free_of[s];
    if (row < 0) return;
    const int64_t cu = s / bs;
    const int jd = (int)(s - cu * bs);
    const int bs2 = bs * bs;
    for (int c = 0; c < k; ++c) {
        double lift = 0.0;
        if (qd) {
            const double* q = qd + (int64_t)c * n;
            for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
                const int64_t r = browidx[u];
                for (int i = 0; i < bs; ++i) {
                    const int64_t d = r * bs + i;
                    const double v = bval[u * bs2 + i * bs + jd];
                    if (free_of[d] < 0) lift += v * q[d];  // (K q_d)[s], columns ascending as the CSC matvec sums
                }
            }
        }
        rhs[(int64_t)row * kp + c] = (f ? f[(int64_t)c * n + s] : 0.0) - lift;
    }
    for (int c = k; c < kp; ++c) rhs[(int64_t)row * kp + c] = 0.0;
}

__global__ void mask_diff_kernel(const uint8_t* __restrict__ fixed, const int32_t* __restrict__ free_of, int64_t n,
                                 int* __restrict__ diff) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && ((fixed[s] != 0) != (free_of[s] < 0))) atomicOr(diff, 1);
}

// Free blocks per free column: a group of 32 lanes per column reads the
// column's row indices together (one contiguous run) and sums the count by a
// butterfly; every lane of a group runs the same trips.
__global__ void blk_count_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                 const int32_t* __restrict__ ufree_of, int64_t nunits, int32_t* __restrict__ blen) {
    constexpr int G = 32;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t cu = t / G;
    const int lane = (int)(t % G);
    int32_t row = -1;
    int cnt = 0;
    if (cu < nunits) {
        row = ufree_of[cu];
        if (row >= 0)
            for (int64_t u = bcolptr[cu] + lane; u < bcolptr[cu + 1]; u += G) cnt += ufree_of[browidx[u]] >= 0 ? 1 : 0;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, G);
    if (lane == 0 && row >= 0) blen[row] = cnt;
}

__global__ void slice_width_kernel(const int32_t* __restrict__ blen, int64_t nfb, int64_t nslices,
                                   int32_t* __restrict__ sl_w, int64_t* __restrict__ sl_cnt) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int w = 0;
    for (int64_t r = s * 64; r < min(nfb, s * 64 + 64); ++r) w = max(w, blen[r]);
    sl_w[s] = w;
    sl_cnt[s] = (int64_t)w * 64;
}

template <int BS>
__global__ void sell_fill_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                 const double* __restrict__ bval, const int32_t* __restrict__ ufree_of,
                                 int64_t nunits, const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                 int32_t* __restrict__ scol, double* __restrict__ sval64,
                                 float* __restrict__ sval32, int16_t* __restrict__ sval16, int64_t nfb,
                                 double* __restrict__ binv64, float* __restrict__ binv32,
                                 const int32_t* __restrict__ free_of, double* __restrict__ pdinv64) {
    constexpr int B2 = BS * BS;
    const int64_t cu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cu >= nunits) return;
    const int32_t row = ufree_of[cu];
    if (row < 0) return;
    const int64_t s = row / 64;
    const int lane = row % 64;
    const int64_t base = sl_ptr[s];
    int k = 0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t ru = browidx[u];
        const int32_t c = ufree_of[ru];
        if (c < 0) continue;
        const int64_t slot = base + (int64_t)k * 64;
        scol[slot + lane] = c;
        // M[j][i] = K[(row,j),(c,i)] = K[(c,i),(row,j)]; a fixed dof of a
        // partially fixed node gets a zero row / column and a unit diagonal
        double m[B2];
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                const bool masked = free_of[cu * BS + j] < 0 || free_of[ru * BS + i] < 0;
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
            if (pdinv64)
#pragma unroll
                for (int j = 0; j < BS; ++j) pdinv64[(int64_t)row * BS + j] = m[j * BS + j] != 0.0 ? 1.0 / m[j * BS + j] : 1.0;
            invert_block<BS>(m, inv);
#pragma unroll
            for (int v = 0; v < B2; ++v) {
                binv64[v * nfb + row] = inv[v];
                binv32[v * nfb + row] = (float)inv[v];
            }
        }
        ++k;
    }
    for (; k < sl_w[s]; ++k) {  // padding slots: own column, zero values
        const int64_t slot = base + (int64_t)k * 64;
        scol[slot + lane] = row;
#pragma unroll
        for (int v = 0; v < B2; ++v) {
            sval64[slot * B2 + v * 64 + lane] = 0.0;
            sval32[slot * B2 + v * 64 + lane] = 0.0f;
            sval16[slot * B2 + v * 64 + lane] = 0;
        }
    }
}

// Rows of a partially filled last slice (row >= nfb) still get full padding
// so the SpMV can run every lane without bounds checks on the values.
template <int BS>
__global__ void sell_pad_tail_kernel(int64_t nfb, int64_t nslices, const int64_t* __restrict__ sl_ptr,
                                     const int32_t* __restrict__ sl_w, int32_t* __restrict__ scol,
                                     double* __restrict__ sval64, float* __restrict__ sval32,
                                     int16_t* __restrict__ sval16) {
    constexpr int B2 = BS * BS;
    const int lane = threadIdx.x;
    const int64_t s = nslices - 1;
    const int64_t row = s * 64 + lane;
    if (lane >= 64 || row < nfb) return;
    for (int k = 0; k < sl_w[s]; ++k) {
        const int64_t slot = sl_ptr[s] + (int64_t)k * 64;
        scol[slot + lane] = (int32_t)(nfb - 1);
        for (int v = 0; v < B2; ++v) {
            sval64[slot * B2 + v * 64 + lane] = 0.0;
            sval32[slot * B2 + v * 64 + lane] = 0.0f;
            sval16[slot * B2 + v * 64 + lane] = 0;
        }
    }
}

__global__ void to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void recombine_kernel(const int32_t* __restrict__ free_of, const double* __restrict__ x,
                                 const double* __restrict__ qd, int64_t n, double* __restrict__ q) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n) q[d] = free_of[d] >= 0 ? x[free_of[d]] : qd[d];
}

// R = K q - f on the fixed dofs (zero on free dofs).
__global__ void reactions_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                 const double* __restrict__ bval, const int32_t* __restrict__ free_of,
                                 const double* __restrict__ q, const double* __restrict__ f, int64_t n, int bs,
                                 double* __restrict__ R) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (free_of[s] >= 0) { R[s] = 0.0; return; }
    const int64_t cu = s / bs;
    const int j = (int)(s - cu * bs);
    double acc = 0.0;
    for (int64_t u = bcolptr[cu]; u < bcolptr[cu + 1]; ++u) {
        const int64_t r = browidx[u];
        for (int i = 0; i < bs; ++i) acc += bval[u * bs * bs + i * bs + j] * q[r * bs + i];
    }
    R[s] = acc - f[s];
}

}  // namespace

namespace {
__global__ void pad_gather_kernel(const int32_t* __restrict__ pad_of, int64_t n, const double* __restrict__ src,
                                  double* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t f = pad_of[k];
    dst[k] = f >= 0 ? src[f] : 0.0;
}

__global__ void pad_scatter_kernel(const int32_t* __restrict__ pad_of, int64_t n, const double* __restrict__ src,
                                   double* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t f = pad_of[k];
    if (f >= 0) dst[f] = src[k];
}
}  // namespace

void pad_gather(sfem_ctx* ctx, const sfem_system* S, const double* compact, double* padded) {
    const int64_t n = S->nfb * S->bs;
    if (n == 0) return;
    hipLaunchKernelGGL(pad_gather_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, S->pad_of.get(), n,
                       compact, padded);
    SFEM_LAUNCH_CHECK();
}

void pad_scatter(sfem_ctx* ctx, const sfem_system* S, const double* padded, double* compact) {
    const int64_t n = S->nfb * S->bs;
    if (n == 0) return;
    hipLaunchKernelGGL(pad_scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, S->pad_of.get(), n,
                       padded, compact);
    SFEM_LAUNCH_CHECK();
}

void eliminate(sfem_ctx* ctx, const sfem_csr* K, const sfem_bc* bc, sfem_system* S) {
    if (bc->n != K->n || K->m != K->n) fail(SFEM_ERR_DIMENSION, "eliminate: BC length differs from size(K)");
    hipEvent_t e0, e1;
    SFEM_HIP(hipEventCreate(&e0));
    SFEM_HIP(hipEventCreate(&e1));
    SFEM_HIP(hipEventRecord(e0, ctx->stream));
    const int64_t n = K->n;
    const int bs = K->bs;
    S->ctx = ctx;
    S->K = K;
    S->n = n;
    S->bs = bs;
    S->free_of.reset(ctx, (size_t)n);
    S->qd.reset(ctx, (size_t)n);
    S->f.reset(ctx, (size_t)n);
    SFEM_HIP(hipMemcpyAsync(S->qd.get(), bc->qd.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    SFEM_HIP(hipMemcpyAsync(S->f.get(), bc->f.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));

    DevBuf<uint32_t> pos(ctx, (size_t)n), tot(ctx, 1);
    const dim3 gn(grid_for(n, 256));
    hipLaunchKernelGGL(free_flags_kernel, gn, dim3(256), 0, ctx->stream, bc->fixed.get(), n, pos.get());
    SFEM_LAUNCH_CHECK();
    exclusive_scan_u32_dev(ctx, pos.get(), pos.get(), n, tot.get());
    hipLaunchKernelGGL(free_map_kernel, gn, dim3(256), 0, ctx->stream, bc->fixed.get(), pos.get(), n,
                       S->free_of.get());
    SFEM_LAUNCH_CHECK();
    uint32_t nf = 0;
    d2h(ctx, &nf, tot.get(), sizeof(nf));
    S->nf = nf;

    // scalar CSR: row pointers, columns and values on first use (ensure_rowptr, ensure_csr)
    const int64_t nunits = K->nunits;
    const dim3 gu(grid_for(nunits, 256));
    S->nnz = -1;
    S->rowptr.release();
    S->col.release();
    S->val64.release();
    S->val32.release();
    S->rhs.reset(ctx, nf);
    S->x.reset(ctx, nf);
    S->dinv64.release();  // point Jacobi's 1/diag: on first use (ensure_dinv)
    S->dinv32.release();
    memset_d(ctx, S->x.get(), 0, sizeof(double) * nf);
    if (bc->qd_zero)
        hipLaunchKernelGGL(rhs_gather_kernel, gn, dim3(256), 0, ctx->stream, S->free_of.get(), S->f.get(), n,
                           S->rhs.get());
    else
        hipLaunchKernelGGL(rhs_kernel, gn, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), S->free_of.get(), S->qd.get(), S->f.get(), n, bs, S->rhs.get());
    SFEM_LAUNCH_CHECK();

    // node-block sliced format: every node with a free dof is a block row;
    // here the slices' shape, the row map and the inverse diagonal blocks,
    // the values on first use (ensure_sell)
    S->has_block = false;
    S->partial = false;
    S->sell_ready = false;
    S->scol.release();
    S->sval64.release();
    S->sval32.release();
    S->sval16.release();
    if (bs > 1 && nf > 0) {
        DevBuf<uint32_t> ufree(ctx, (size_t)nunits), upos(ctx, (size_t)nunits), utot(ctx, 1);
        DevBuf<int> dpartial(ctx, 1);
        memset_d(ctx, dpartial.get(), 0, sizeof(int));
        hipLaunchKernelGGL(unit_status_kernel, gu, dim3(256), 0, ctx->stream, bc->fixed.get(), nunits, bs,
                           ufree.get(), dpartial.get());
        SFEM_LAUNCH_CHECK();
        int part = 0;
        d2h(ctx, &part, dpartial.get(), sizeof(int));
        S->partial = part != 0;
        {
            exclusive_scan_u32_dev(ctx, ufree.get(), upos.get(), nunits, utot.get());
            uint32_t nfb = 0;
            d2h(ctx, &nfb, utot.get(), sizeof(nfb));
            S->ufree_of.reset(ctx, (size_t)nunits);
            int32_t* ufree_of = S->ufree_of.get();
            // unit map: fixed unit <=> ufree == 0
            DevBuf<uint8_t> ufixed(ctx, (size_t)nunits);
            hipLaunchKernelGGL(unit_fixed_kernel, gu, dim3(256), 0, ctx->stream, ufree.get(), ufixed.get(), nunits);
            SFEM_LAUNCH_CHECK();
            hipLaunchKernelGGL(free_map_kernel, gu, dim3(256), 0, ctx->stream, ufixed.get(), upos.get(), nunits,
                               ufree_of);
            SFEM_LAUNCH_CHECK();
            DevBuf<int32_t> blen(ctx, nfb);
            hipLaunchKernelGGL(blk_count_kernel, dim3(grid_for(nunits * 32, 256)), dim3(256), 0, ctx->stream,
                               K->bcolptr.get(), K->browidx.get(), (const int32_t*)ufree_of, nunits, blen.get());
            SFEM_LAUNCH_CHECK();
            const int64_t nslices = (nfb + 63) / 64;
            S->nfb = nfb;
            S->nslices = nslices;
            {  // node blocks of K_ff (with the masked blocks of partially fixed nodes)
                DevBuf<uint32_t> btot(ctx, 1);
                DevBuf<uint32_t> bscan(ctx, nfb);
                exclusive_scan_u32_dev(ctx, (const uint32_t*)blen.get(), bscan.get(), nfb, btot.get());
                uint32_t nb = 0;
                d2h(ctx, &nb, btot.get(), sizeof(nb));
                S->nnzb = nb;
            }
            if (S->partial) {
                S->pad_of.reset(ctx, (size_t)nfb * bs);
                S->prhs.reset(ctx, (size_t)nfb * bs);
                S->px.reset(ctx, (size_t)nfb * bs);
                S->pdinv64.reset(ctx, (size_t)nfb * bs);
                S->pdinv32.reset(ctx, (size_t)nfb * bs);
                hipLaunchKernelGGL(pad_map_kernel, gu, dim3(256), 0, ctx->stream, (const int32_t*)ufree_of,
                                   S->free_of.get(), nunits, bs, S->pad_of.get());
                SFEM_LAUNCH_CHECK();
            }
            double* pdinv = S->partial ? S->pdinv64.get() : nullptr;
            S->sl_ptr.reset(ctx, (size_t)nslices + 1);
            S->sl_w.reset(ctx, (size_t)nslices);
            DevBuf<int64_t> cnt(ctx, (size_t)nslices + 1);
            memset_d(ctx, cnt.get(), 0, sizeof(int64_t) * (nslices + 1));
            hipLaunchKernelGGL(slice_width_kernel, dim3(grid_for(nslices, 256)), dim3(256), 0, ctx->stream,
                               blen.get(), (int64_t)nfb, nslices, S->sl_w.get(), cnt.get());
            SFEM_LAUNCH_CHECK();
            const int64_t total = exclusive_scan_i64(ctx, cnt.get(), S->sl_ptr.get(), nslices + 1);
            if (total >= (int64_t(1) << 31)) fail(SFEM_ERR_UNSUPPORTED, "block format too large");
            S->sell_blocks = total;
            S->binv64.reset(ctx, (size_t)nfb * bs * bs);
            S->binv32.reset(ctx, (size_t)nfb * bs * bs);
            if (bs == 3)
                hipLaunchKernelGGL(diag_blocks_kernel<3>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(),
                                   K->browidx.get(), K->bval.get(), (const int32_t*)ufree_of, S->free_of.get(),
                                   nunits, (int64_t)nfb, S->binv64.get(), S->binv32.get(), pdinv);
            else
                hipLaunchKernelGGL(diag_blocks_kernel<2>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(),
                                   K->browidx.get(), K->bval.get(), (const int32_t*)ufree_of, S->free_of.get(),
                                   nunits, (int64_t)nfb, S->binv64.get(), S->binv32.get(), pdinv);
            SFEM_LAUNCH_CHECK();
            if (S->partial) {
                hipLaunchKernelGGL(to_f32_kernel, dim3(grid_for((int64_t)nfb * bs, 256)), dim3(256), 0, ctx->stream,
                                   S->pdinv64.get(), S->pdinv32.get(), (int64_t)nfb * bs);
                SFEM_LAUNCH_CHECK();
                pad_gather(ctx, S, S->rhs.get(), S->prhs.get());
            }
            S->has_block = true;
        }
    }
    SFEM_HIP(hipEventRecord(e1, ctx->stream));
    sync(ctx);
    float ms = 0;
    SFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    S->t_eliminate_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void ensure_sell(sfem_ctx* ctx, sfem_system* S) {
    if (!S->has_block || S->sell_ready) return;
    const sfem_csr* K = S->K;
    const int bs = S->bs;
    const int64_t nunits = K->nunits, total = S->sell_blocks;
    const dim3 gu(grid_for(nunits, 256));
    S->scol.reset(ctx, (size_t)total);
    S->sval64.reset(ctx, (size_t)total * bs * bs);
    S->sval32.reset(ctx, (size_t)total * bs * bs);
    S->sval16.reset(ctx, (size_t)total * bs * bs);
    // the inverse blocks (and padded 1/diag) are written again, the same values
    double* pdinv = S->partial ? S->pdinv64.get() : nullptr;
    if (bs == 3) {
        hipLaunchKernelGGL(sell_fill_kernel<3>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), S->ufree_of.get(), nunits, S->sl_ptr.get(), S->sl_w.get(), S->scol.get(),
                           S->sval64.get(), S->sval32.get(), S->sval16.get(), S->nfb, S->binv64.get(),
                           S->binv32.get(), S->free_of.get(), pdinv);
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(sell_pad_tail_kernel<3>, dim3(1), dim3(64), 0, ctx->stream, S->nfb, S->nslices,
                           S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval64.get(), S->sval32.get(),
                           S->sval16.get());
    } else {
        hipLaunchKernelGGL(sell_fill_kernel<2>, gu, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), S->ufree_of.get(), nunits, S->sl_ptr.get(), S->sl_w.get(), S->scol.get(),
                           S->sval64.get(), S->sval32.get(), S->sval16.get(), S->nfb, S->binv64.get(),
                           S->binv32.get(), S->free_of.get(), pdinv);
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(sell_pad_tail_kernel<2>, dim3(1), dim3(64), 0, ctx->stream, S->nfb, S->nslices,
                           S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval64.get(), S->sval32.get(),
                           S->sval16.get());
    }
    SFEM_LAUNCH_CHECK();
    S->sell_ready = true;
}

void ensure_dinv(sfem_ctx* ctx, sfem_system* S) {
    if (S->dinv64.size() || S->nf == 0) return;
    const sfem_csr* K = S->K;
    S->dinv64.reset(ctx, (size_t)S->nf);
    S->dinv32.reset(ctx, (size_t)S->nf);
    hipLaunchKernelGGL(dinv_kernel, dim3(grid_for(S->n, 256)), dim3(256), 0, ctx->stream, K->bcolptr.get(),
                       K->browidx.get(), K->bval.get(), S->free_of.get(), S->n, S->bs, S->dinv64.get());
    SFEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(to_f32_kernel, dim3(grid_for(S->nf, 256)), dim3(256), 0, ctx->stream, S->dinv64.get(),
                       S->dinv32.get(), S->nf);
    SFEM_LAUNCH_CHECK();
}

void ensure_rowptr(sfem_ctx* ctx, sfem_system* S) {
    if (S->nnz >= 0) return;
    const sfem_csr* K = S->K;
    const int64_t nf = S->nf;
    DevBuf<uint32_t> rowptr(ctx, (size_t)nf + 1);
    memset_d(ctx, rowptr.get(), 0, sizeof(uint32_t) * (nf + 1));
    hipLaunchKernelGGL(csr_count_kernel, dim3(grid_for(K->nunits, 256)), dim3(256), 0, ctx->stream,
                       K->bcolptr.get(), K->browidx.get(), S->free_of.get(), K->nunits, S->bs, rowptr.get());
    SFEM_LAUNCH_CHECK();
    {
        DevBuf<uint32_t> t2(ctx, 1);
        exclusive_scan_u32_dev(ctx, rowptr.get(), rowptr.get(), nf + 1, t2.get());
    }
    uint32_t nnz = 0;
    if (nf) d2h(ctx, &nnz, rowptr.get() + nf, sizeof(nnz));
    if ((uint64_t)nnz >= (1ull << 31)) fail(SFEM_ERR_UNSUPPORTED, "K_ff has 2^31 or more entries");
    S->rowptr.reset(ctx, (size_t)nf + 1);
    SFEM_HIP(hipMemcpyAsync(S->rowptr.get(), rowptr.get(), sizeof(int32_t) * (nf + 1), hipMemcpyDeviceToDevice,
                            ctx->stream));
    sync(ctx);
    S->nnz = nnz;
}

void ensure_csr(sfem_ctx* ctx, sfem_system* S) {
    ensure_rowptr(ctx, S);
    if (S->nnz == 0 || S->col.size() == (size_t)S->nnz) return;
    const sfem_csr* K = S->K;
    S->col.reset(ctx, (size_t)S->nnz);
    S->val64.reset(ctx, (size_t)S->nnz);
    // the fill also forms rhs and 1/diag; those go to scratch (S keeps its own)
    DevBuf<double> rhs(ctx, (size_t)S->nf), dinv(ctx, (size_t)S->nf);
    hipLaunchKernelGGL(csr_fill_kernel, dim3(grid_for(S->n, 256)), dim3(256), 0, ctx->stream, K->bcolptr.get(),
                       K->browidx.get(), K->bval.get(), S->free_of.get(), S->qd.get(), S->f.get(), S->n, S->bs,
                       (const uint32_t*)S->rowptr.get(), S->col.get(), S->val64.get(), rhs.get(), dinv.get());
    SFEM_LAUNCH_CHECK();
}

void system_set_bc(sfem_ctx* ctx, sfem_system* S, const sfem_bc* bc) {
    if (bc->n != S->n) fail(SFEM_ERR_DIMENSION, "system_set_bc: BC length differs from the system");
    const int64_t n = S->n;
    const dim3 gn(grid_for(n, 256));
    DevBuf<int> diff(ctx, 1);
    memset_d(ctx, diff.get(), 0, sizeof(int));
    hipLaunchKernelGGL(mask_diff_kernel, gn, dim3(256), 0, ctx->stream, bc->fixed.get(), S->free_of.get(), n,
                       diff.get());
    SFEM_LAUNCH_CHECK();
    int d = 0;
    d2h(ctx, &d, diff.get(), sizeof(int));
    if (d) fail(SFEM_ERR_ARGUMENT, "system_set_bc: the fixed dofs differ from the eliminated system's");
    SFEM_HIP(hipMemcpyAsync(S->qd.get(), bc->qd.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    SFEM_HIP(hipMemcpyAsync(S->f.get(), bc->f.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    const sfem_csr* K = S->K;
    if (bc->qd_zero)
        hipLaunchKernelGGL(rhs_gather_kernel, gn, dim3(256), 0, ctx->stream, S->free_of.get(), S->f.get(), n,
                           S->rhs.get());
    else
        hipLaunchKernelGGL(rhs_kernel, gn, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), S->free_of.get(), S->qd.get(), S->f.get(), n, S->bs, S->rhs.get());
    SFEM_LAUNCH_CHECK();
    if (S->partial && S->has_block) pad_gather(ctx, S, S->rhs.get(), S->prhs.get());
    sync(ctx);
}

void launch_rhs_many(sfem_ctx* ctx, const sfem_system* S, int k, int kp, const double* qd, const double* f,
                     double* rhs) {
    const sfem_csr* K = S->K;
    hipLaunchKernelGGL(rhs_many_kernel, dim3(grid_for(S->n, 256)), dim3(256), 0, ctx->stream, K->bcolptr.get(),
                       K->browidx.get(), K->bval.get(), S->free_of.get(), qd, f, S->n, S->bs, k, kp, rhs);
    SFEM_LAUNCH_CHECK();
}

void system_solution(sfem_ctx* ctx, const sfem_system* S, double* q, double* reactions) {
    const int64_t n = S->n;
    DevBuf<double> dq(ctx, (size_t)n);
    const dim3 gn(grid_for(n, 256));
    hipLaunchKernelGGL(recombine_kernel, gn, dim3(256), 0, ctx->stream, S->free_of.get(), S->x.get(), S->qd.get(), n,
                       dq.get());
    SFEM_LAUNCH_CHECK();
    if (reactions) {
        DevBuf<double> dr(ctx, (size_t)n);
        const sfem_csr* K = S->K;
        hipLaunchKernelGGL(reactions_kernel, gn, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), S->free_of.get(), dq.get(), S->f.get(), n, S->bs, dr.get());
        SFEM_LAUNCH_CHECK();
        d2h(ctx, reactions, dr.get(), sizeof(double) * n);
    }
    if (q) d2h(ctx, q, dq.get(), sizeof(double) * n);
}

}  // namespace sfem
