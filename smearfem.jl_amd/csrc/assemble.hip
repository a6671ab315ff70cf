// assemble.hip -- K = sparse(E, J, V) on the device by sort-by-key and an
// in-order segmented reduction.
//
// Reference: smearFEM src/fem.jl:135-256; the combine semantics are those of
// Julia's SparseArrays.sparse(I,J,V) at src/fem.jl:253: entries with equal
// (row, col) are summed left to right in COO index order, explicit zeros are
// kept (the pattern is the union of element couplings), rows ascend within
// each column, and the dimensions are the largest row / column index.
//
// Pipeline (one stream, stage times recorded with HIP events):
//   validate -> element Ke (COO values) -> keys -> stable radix sort ->
//   segment heads -> scan -> in-order segment sums -> block CSC pointers.
// On a mesh declared a structured grid (mesh_set_grid) the pattern is the
// grid's: block counts -> scan -> one gather of every node pair's element
// blocks in ascending element order (the same sums, bit for bit).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sfem_objects.hpp"

namespace sfem {

namespace {

__global__ __launch_bounds__(256) void seg_flags_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                        uint32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// pos = exclusive scan of the flags; heads of segments record where they start.
__global__ __launch_bounds__(256) void seg_heads_kernel(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ pos, int64_t n,
                                                        uint32_t nseg, uint32_t* __restrict__ head,
                                                        uint64_t* __restrict__ ukey) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    if (i == 0 || k != keys[i - 1]) {
        head[pos[i]] = (uint32_t)i;
        ukey[pos[i]] = k;
    }
    if (i == n - 1) head[nseg] = (uint32_t)n;
}

// One thread per unique key: values of the segment summed in sorted order,
// which is ascending COO index (stable sort), starting from the first value
// as Julia's combine does.  Block values for payload p sit at
// V[p + i*ndof + j] (contiguous when bs == ndof).
template <int BS>
__global__ __launch_bounds__(256) void seg_sum_kernel(const double* __restrict__ V,
                                                      const uint32_t* __restrict__ payload,
                                                      const uint32_t* __restrict__ head,
                                                      const uint64_t* __restrict__ ukey, int64_t nseg,
                                                      int ndof, int row_bits,
                                                      double* __restrict__ bval,
                                                      int32_t* __restrict__ browidx,
                                                      int32_t* __restrict__ bcolidx) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nseg) return;
    const uint32_t s = head[u], e = head[u + 1];
    double acc[BS * BS];
    {
        const uint32_t p = payload[s];
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < BS; ++j) acc[i * BS + j] = V[(int64_t)p + i * ndof + j];
    }
    for (uint32_t q = s + 1; q < e; ++q) {
        const uint32_t p = payload[q];
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < BS; ++j) acc[i * BS + j] += V[(int64_t)p + i * ndof + j];
    }
#pragma unroll
    for (int v = 0; v < BS * BS; ++v) bval[u * BS * BS + v] = acc[v];
    const uint64_t k = ukey[u];
    browidx[u] = (int32_t)(k & ((1ull << row_bits) - 1));
    bcolidx[u] = (int32_t)(k >> row_bits);
}

// bcolptr[c] = first block whose column unit is >= c.
__global__ __launch_bounds__(256) void colptr_kernel(const int32_t* __restrict__ bcolidx, int64_t nnzb,
                                                     int64_t nunits, int64_t* __restrict__ bcolptr) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nnzb) return;
    const int64_t c = bcolidx[u];
    const int64_t cp = u == 0 ? -1 : bcolidx[u - 1];
    for (int64_t cc = cp + 1; cc <= c; ++cc) bcolptr[cc] = u;
    if (u == nnzb - 1)
        for (int64_t cc = c + 1; cc <= nunits; ++cc) bcolptr[cc] = nnzb;
}

__global__ void fill_i64_kernel(int64_t* p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Scalar CSC (1-based, Int64 indices) from the block CSC.
__global__ __launch_bounds__(256) void export_colptr_kernel(const int64_t* __restrict__ bcolptr, int bs,
                                                            int64_t n, int64_t nnz,
                                                            int64_t* __restrict__ colptr) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > n) return;
    if (s == n) { colptr[n] = nnz + 1; return; }
    const int64_t cu = s / bs, j = s % bs;
    const int64_t b0 = bcolptr[cu], nbc = bcolptr[cu + 1] - b0;
    colptr[s] = b0 * bs * bs + j * nbc * bs + 1;
}

__global__ __launch_bounds__(256) void export_entries_kernel(
    const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
    const int32_t* __restrict__ bcolidx, const double* __restrict__ bval, int bs, int64_t nnz,
    int64_t* __restrict__ rowval, double* __restrict__ nzval) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    const int bs2 = bs * bs;
    const int64_t u = t / bs2;
    const int v = (int)(t - u * bs2), i = v / bs, j = v % bs;
    const int64_t cu = bcolidx[u];
    const int64_t b0 = bcolptr[cu], nbc = bcolptr[cu + 1] - b0, k = u - b0;
    const int64_t pos = b0 * bs2 + (int64_t)j * nbc * bs + k * bs + i;
    rowval[pos] = (int64_t)browidx[u] * bs + i + 1;
    nzval[pos] = bval[t];
}

int bits_for(int64_t nunits) {
    int b = 0;
    while ((int64_t(1) << b) < nunits) ++b;
    return b;
}

struct Timer {
    hipEvent_t ev[8];
    int n = 0;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        for (auto& e : ev) SFEM_HIP(hipEventCreate(&e));
    }
    ~Timer() {
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    void mark() { SFEM_HIP(hipEventRecord(ev[n++], s)); }
    double ms(int a, int b) const {
        float t = 0;
        SFEM_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
        return t;
    }
};

int code_of_step(int step) {
    switch (step) {
        case 0: case 4: return SFEM_ERR_BOUNDS;
        case 2: case 5: return SFEM_ERR_SINGULAR;
        default: return SFEM_ERR_DIMENSION;
    }
}

const char* name_of_code(int code) {
    switch (code) {
        case SFEM_ERR_BOUNDS: return "BoundsError";
        case SFEM_ERR_SINGULAR: return "SingularException";
        case SFEM_ERR_ARGUMENT: return "ArgumentError";
        case SFEM_ERR_DIMENSION: return "DimensionMismatch";
        case SFEM_ERR_UNDEFVAR: return "UndefVarError";
        default: return "error";
    }
}

}  // namespace

namespace {

// ------------------------------------------------ structured assembly --
// On a structured grid every node pair (r, c) that shares an element shares
// the cells between them: per direction the two cells around the node when
// r and c have the same coordinate, else the one cell between them.  Their
// contributions are summed in ascending element order starting from the
// first, as the stable sort and segmented sum do, so K is bitwise the same.
// K summed from the element scales when every Ke(e) / s_e is within this
// much of Ke(0) (relative to Ke(0)'s largest entry): K's values then stay
// within it of the element sums, 10x inside the 1e-12 the assembly is held to
constexpr double kGridScaleDev = 1e-13;

struct GridAsm {
    int64_t n[3];   // nodes per direction (1 beyond ndim)
    int64_t e[3];   // elements per direction (1 beyond ndim)
    int corner[8];  // local index of the cell corner with code ox + 2 oy + 4 oz
};

template <int ND>
__device__ __forceinline__ void grid_coords(int64_t node, const GridAsm& g, int64_t c[3]) {
    c[0] = node % g.n[0];
    c[1] = (node / g.n[0]) % g.n[1];
    c[2] = ND == 3 ? node / (g.n[0] * g.n[1]) : 0;
}

// blocks per node column: product over directions of its valid offsets
template <int ND>
__global__ __launch_bounds__(256) void grid_counts_kernel(GridAsm g, int64_t nn, int64_t* __restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nn) return;
    if (c == nn) {
        cnt[nn] = 0;
        return;
    }
    int64_t a[3];
    grid_coords<ND>(c, g, a);
    int64_t k = 1;
    for (int d = 0; d < ND; ++d) k *= 1 + (a[d] > 0) + (a[d] < g.n[d] - 1);
    cnt[c] = k;
}

// out[block] = max of the block's grid-stride share of in[0, n) (NaN
// propagates); a second launch with one block reduces the block results
__global__ __launch_bounds__(1024) void max_reduce_kernel(const double* __restrict__ in, int64_t n,
                                                          double* __restrict__ out) {
    __shared__ double wmax[16];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max_nan(m, in[i]);
    for (int off = 32; off > 0; off >>= 1) m = max_nan(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = max_nan(b, wmax[w]);
        out[blockIdx.x] = b;
    }
}

// Half a wavefront per node column c (4 columns per block of 128).
//   phase 1: lane (cell, part) copies its part of the column node's section
//            in one of the cells around c (the element's NEN blocks whose
//            column corner is c, NEN * ndof^2 contiguous values) from V to LDS
//            with 16-byte loads;
//   phase 2: lane s (stencil slot, ascending row order: z, then y, then x
//            offset) sums the shared cells' blocks for row node c + offset(s)
//            in ascending element order and stores K block k, the slot's rank
//            among the column's valid slots (each direction's valid offsets
//            as a mixed radix, x fastest).
// Element matrices are bitwise symmetric, so block (row ar, column c) is read
// as the transpose of block (c, ar), which lies in c's section.
template <int ND>
__global__ __launch_bounds__(128) void grid_gather_kernel(const double* __restrict__ V, GridAsm g, int64_t nn,
                                                          const int64_t* __restrict__ bcolptr,
                                                          double* __restrict__ bval, int32_t* __restrict__ browidx,
                                                          int32_t* __restrict__ bcolidx) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND, SEC = NEN * B2, KE = NEN * SEC;
    constexpr int NS = ND == 3 ? 27 : 9, CPB = 128 / 32;
    constexpr int LPC = 32 / NEN, PER = SEC / LPC;  // lanes per cell, values per lane
    static_assert(PER % 2 == 0 && (SEC * 8) % 16 == 0, "16-byte copies");
    __shared__ __align__(16) double sv[CPB][NEN * SEC];
    const int hl = threadIdx.x & 31, w = threadIdx.x >> 5;
    const int64_t c = (int64_t)blockIdx.x * CPB + w;
    const bool live = c < nn;
    int64_t a[3] = {0, 0, 0};
    if (live) grid_coords<ND>(c, g, a);
    {
        const int kc = hl / LPC, part = hl % LPC;
        bool ex = live;
        int code = 0;
        int64_t cell[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int k = d < ND ? (kc >> d) & 1 : 1;
            cell[d] = a[d] - 1 + k;
            ex = ex && cell[d] >= 0 && cell[d] < g.e[d];
            code += (1 - k) << d;
        }
        if (ex) {
            const int64_t e = cell[0] + g.e[0] * (cell[1] + g.e[1] * cell[2]);
            const double2* src = reinterpret_cast<const double2*>(V + e * KE + g.corner[code] * SEC + part * PER);
            double2* dst = reinterpret_cast<double2*>(&sv[w][kc * SEC + part * PER]);
            double2 t[PER / 2];
#pragma unroll
            for (int r = 0; r < PER / 2; ++r) t[r] = src[r];
#pragma unroll
            for (int r = 0; r < PER / 2; ++r) dst[r] = t[r];
        }
    }
    __syncthreads();
    if (!live || hl >= NS) return;
    const int s = hl;
    const int o[3] = {s % 3, (s / 3) % 3, ND == 3 ? s / 9 : 1};  // offset + 1 per direction
    int64_t b[3];
    int k = 0, radix = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b[d] = a[d] + o[d] - 1;
        if (b[d] < 0 || b[d] >= g.n[d]) return;
        const int lo = a[d] > 0 ? 0 : 1, hi = a[d] < g.n[d] - 1 ? 2 : 1;
        k += (o[d] - lo) * radix;
        radix *= hi - lo + 1;
    }
    const int64_t u = bcolptr[c] + k;
    double acc[B2];
    bool first = true;
#pragma unroll
    for (int kc = 0; kc < NEN; ++kc) {  // cell (a - 1 + k) per direction: ascending element index
        bool ok = true;
        int rc = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int kd = d < ND ? (kc >> d) & 1 : 1;
            const int64_t cd = a[d] - 1 + kd;
            const int rb = o[d] - kd;  // row node's corner bit in the cell
            ok = ok && cd >= 0 && cd < g.e[d] && (rb == 0 || rb == 1);
            rc += (rb & 1) << d;
        }
        if (!ok) continue;
        const double* p = &sv[w][kc * SEC + g.corner[rc] * B2];
#pragma unroll
        for (int m = 0; m < B2; ++m) {
            const double pv = p[(m % BS) * BS + m / BS];
            acc[m] = first ? pv : acc[m] + pv;
        }
        first = false;
    }
#pragma unroll
    for (int m = 0; m < B2; ++m) bval[u * B2 + m] = acc[m];
    browidx[u] = (int32_t)(b[0] + g.n[0] * (b[1] + g.n[1] * b[2]));
    bcolidx[u] = (int32_t)c;
}

// K from the element scales, when every element matrix is s_e Ke(0) to
// within the measured deviation: half a wavefront per node column c as in
// grid_gather_kernel (the same blocks in the same places), lane s summing
// over the shared cells in ascending element order s_e Ke(0)[row corner]
// [column corner], Ke(0) (local corner numbering) staged in LDS.  The
// column's blocks are put in LDS by rank and written out as one contiguous
// run (consecutive lanes, consecutive doubles).  Ke(0) is bitwise symmetric
// and the blocks (r, c) and (c, r) meet the same cells in the same order, so
// K is bitwise symmetric.
constexpr int kScaleThreads = 128;

template <int ND>
__global__ __launch_bounds__(kScaleThreads) void grid_scale_kernel(const double* __restrict__ ke0,
                                                                   const double* __restrict__ es, GridAsm g,
                                                                   int64_t nn, const int64_t* __restrict__ bcolptr,
                                                                   double* __restrict__ bval,
                                                                   int32_t* __restrict__ browidx,
                                                                   int32_t* __restrict__ bcolidx) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND, KE = NEN * NEN * B2;
    constexpr int NS = ND == 3 ? 27 : 9, CPB = kScaleThreads / 32;
    __shared__ double kl[KE];
    __shared__ double sv[CPB][NS * B2];
    for (int t = (int)threadIdx.x; t < KE; t += kScaleThreads) kl[t] = ke0[t];
    __syncthreads();
    const int hl = threadIdx.x & 31, w = threadIdx.x >> 5;
    const int64_t c = (int64_t)blockIdx.x * CPB + w;
    bool live = c < nn && hl < NS;
    int64_t a[3] = {0, 0, 0};
    if (c < nn) grid_coords<ND>(c, g, a);
    const int s = hl;
    const int o[3] = {s % 3, (s / 3) % 3, ND == 3 ? s / 9 : 1};  // offset + 1 per direction
    int64_t b[3];
    int k = 0, radix = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b[d] = a[d] + o[d] - 1;
        live = live && b[d] >= 0 && b[d] < g.n[d];
        const int lo = a[d] > 0 ? 0 : 1, hi = a[d] < g.n[d] - 1 ? 2 : 1;
        k += (o[d] - lo) * radix;
        radix *= hi - lo + 1;
    }
    if (live) {
        double acc[B2];
#pragma unroll
        for (int m = 0; m < B2; ++m) acc[m] = 0.0;
#pragma unroll
        for (int kc = 0; kc < NEN; ++kc) {  // cell (a - 1 + k) per direction: ascending element index
            bool ok = true;
            int rc = 0, cc = 0;
            int64_t cd[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int kd = d < ND ? (kc >> d) & 1 : 1;
                cd[d] = a[d] - 1 + kd;
                const int rb = o[d] - kd;  // row node's corner bit in the cell
                ok = ok && cd[d] >= 0 && cd[d] < g.e[d] && (rb == 0 || rb == 1);
                rc += (rb & 1) << d;
                cc += (1 - kd) << d;
            }
            if (!ok) continue;
            const double se = es[cd[0] + g.e[0] * (cd[1] + g.e[1] * cd[2])];
            const double* p = &kl[(g.corner[rc] * NEN + g.corner[cc]) * B2];
#pragma unroll
            for (int m = 0; m < B2; ++m) acc[m] = fma(se, p[m], acc[m]);
        }
#pragma unroll
        for (int m = 0; m < B2; ++m) sv[w][k * B2 + m] = acc[m];
        const int64_t u = bcolptr[c] + k;
        browidx[u] = (int32_t)(b[0] + g.n[0] * (b[1] + g.n[1] * b[2]));
        bcolidx[u] = (int32_t)c;
    }
    __syncthreads();
    if (c < nn) {
        const int64_t u0 = bcolptr[c];
        const int nv = (int)(bcolptr[c + 1] - u0) * B2;
        double* dst = bval + u0 * B2;
        for (int t = hl; t < nv; t += 32) dst[t] = sv[w][t];
    }
}

// Is IEN the meshgrid connectivity (corner order of element 0) and the node
// numbering the identity interleave (dof of node m, component 0, = ndof m)?
template <int ND>
__global__ __launch_bounds__(256) void grid_check_kernel(const int32_t* __restrict__ ien,
                                                         const int32_t* __restrict__ dof, GridAsm g, int64_t n_elem,
                                                         int64_t nn, int ndof, int* __restrict__ bad) {
    constexpr int NEN = 1 << ND;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_elem) {
        const int64_t cx = t % g.e[0], cy = (t / g.e[0]) % g.e[1], cz = ND == 3 ? t / (g.e[0] * g.e[1]) : 0;
        for (int code = 0; code < NEN; ++code) {
            const int64_t node = (cx + (code & 1)) + g.n[0] * ((cy + ((code >> 1) & 1)) + g.n[1] * (cz + (code >> 2)));
            if (ien[t * NEN + g.corner[code]] != node) atomicOr(bad, 1);
        }
    }
    if (t < nn && dof[t * ndof] != ndof * t) atomicOr(bad, 2);
}

}  // namespace

void mesh_set_grid(sfem_ctx* ctx, sfem_mesh* m, int ndim, const int64_t* ne) {
    if (!ne) fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: NULL element counts");
    if (m->face || m->cls != SFEM_Q1 || m->ndim != ndim || (ndim != 2 && ndim != 3) || m->nen != (1 << ndim) ||
        m->ndof != ndim)
        fail(SFEM_ERR_UNSUPPORTED, "mesh_set_grid: Q4 / Hex8 vector meshes (n_dof = ndim) only");
    GridAsm g{};
    int64_t ncount = 1, ecount = 1;
    for (int d = 0; d < 3; ++d) {
        g.e[d] = d < ndim ? ne[d] : 1;
        g.n[d] = d < ndim ? ne[d] + 1 : 1;
        if (g.e[d] < 1) fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: element counts must be >= 1");
        ncount *= g.n[d];
        ecount *= g.e[d];
    }
    if (ecount != m->n_elem || ncount != m->n_nodes || m->id_rows < m->n_nodes)
        fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: element or node count differs from the mesh");
    // corner order from element 0 (cell (0, 0, 0))
    const int nen = m->nen;
    std::vector<int32_t> e0((size_t)nen);
    d2h(ctx, e0.data(), m->ien.get(), sizeof(int32_t) * nen);
    int seen = 0;
    for (int a = 0; a < nen; ++a) {
        const int64_t v = e0[(size_t)a];
        const int64_t x = v >= 0 ? v % g.n[0] : -1, y = v >= 0 ? (v / g.n[0]) % g.n[1] : -1;
        const int64_t z = v >= 0 ? v / (g.n[0] * g.n[1]) : -1;
        if (x < 0 || x > 1 || y < 0 || y > 1 || z < 0 || z > 1)
            fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: element 1 is not the first cell of the grid");
        const int code = (int)(x + 2 * y + 4 * z);
        if (seen & (1 << code)) fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: element 1 repeats a corner");
        seen |= 1 << code;
        g.corner[code] = a;
    }
    DevBuf<int> bad(ctx, 1);
    memset_d(ctx, bad.get(), 0, sizeof(int));
    const int64_t nt = std::max(m->n_elem, m->n_nodes);
    if (ndim == 3)
        hipLaunchKernelGGL(grid_check_kernel<3>, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, m->ien.get(),
                           m->dof.get(), g, m->n_elem, m->n_nodes, m->ndof, bad.get());
    else
        hipLaunchKernelGGL(grid_check_kernel<2>, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, m->ien.get(),
                           m->dof.get(), g, m->n_elem, m->n_nodes, m->ndof, bad.get());
    SFEM_LAUNCH_CHECK();
    int hb = 0;
    d2h(ctx, &hb, bad.get(), sizeof(int));
    if (hb & 1) fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: IEN is not the meshgrid connectivity of these element counts");
    if (hb & 2) fail(SFEM_ERR_ARGUMENT, "mesh_set_grid: ID is not the node-interleaved numbering ndof*(node-1)+l");
    m->grid_nd = ndim;
    for (int d = 0; d < 3; ++d) m->grid_ne[d] = g.e[d];
    for (int c = 0; c < 8; ++c) m->grid_corner[c] = g.corner[c];
}

// The element counts of the meshgrid numbering that element 1's corners
// imply: they are the nodes {0, 1, n0, n0+1} (+ n0 n1 + the same in 3D), n_d
// nodes along direction d.  Then the full device check of mesh_set_grid.
// Returns false, and leaves the mesh as it was, when the mesh is not that
// grid.
bool mesh_detect_grid(sfem_ctx* ctx, sfem_mesh* m, int64_t* ne_out) {
    const int nd = m->ndim;
    if (m->grid_nd) {
        for (int d = 0; d < 3; ++d) ne_out[d] = m->grid_ne[d];
        return true;
    }
    if (m->face || m->cls != SFEM_Q1 || (nd != 2 && nd != 3) || m->nen != (1 << nd) || m->ndof != nd ||
        m->n_elem < 1 || m->pre_code || m->dof.size() == 0 || m->id_rows < m->n_nodes)
        return false;
    std::vector<int32_t> c((size_t)m->nen);
    d2h(ctx, c.data(), m->ien.get(), sizeof(int32_t) * m->nen);
    std::sort(c.begin(), c.end());
    if (c[0] != 0 || c[1] != 1) return false;
    const int64_t n0 = c[2];
    if (n0 < 2 || c[3] != n0 + 1) return false;
    int64_t n[3] = {n0, 1, 1};
    if (nd == 2) {
        if (m->n_nodes % n0) return false;
        n[1] = m->n_nodes / n0;
    } else {
        const int64_t n01 = c[4];
        if (n01 % n0 || c[5] != n01 + 1 || c[6] != n01 + n0 || c[7] != n01 + n0 + 1 || m->n_nodes % n01)
            return false;
        n[1] = n01 / n0;
        n[2] = m->n_nodes / n01;
    }
    int64_t ne[3], ecount = 1;
    for (int d = 0; d < nd; ++d) {
        if (n[d] < 2) return false;
        ne[d] = n[d] - 1;
        ecount *= ne[d];
    }
    if (ecount != m->n_elem) return false;
    try {
        mesh_set_grid(ctx, m, nd, ne);
    } catch (const Error& e) {
        if (e.code != SFEM_ERR_ARGUMENT) throw;
        return false;
    }
    for (int d = 0; d < 3; ++d) ne_out[d] = d < nd ? ne[d] : 0;
    return true;
}

void assemble(sfem_ctx* ctx, const sfem_mesh* m, int flags, sfem_csr* K) {
    K->ctx = ctx;
    Timer tm(ctx->stream);
    tm.mark();
    DevBuf<ValidateOut> vout(ctx, kValidateCopies);
    ValidateOut vcopies[kValidateCopies];
    for (ValidateOut& c : vcopies) c = ValidateOut{~0ull, -1, INT32_MAX, 1, 0};
    h2d(ctx, vout.get(), vcopies, sizeof(vcopies));
    launch_validate(ctx, m, vout.get());
    tm.mark();
    const int nen = m->nen, ndof = m->ndof;
    const int64_t KE = (int64_t)nen * nen * ndof * ndof;
    const int64_t ncoo = m->n_elem * KE;
    if (ncoo >= (int64_t(1) << 32)) fail(SFEM_ERR_UNSUPPORTED, "assemble_system: 2^32 or more COO entries");
    DevBuf<double> V;  // the element matrices (not built when K is summed from the scales)
    // a mesh declared a grid gets the element-matrix check first: Ke(0), the
    // scales s_e and the largest deviation of Ke(e) / s_e from Ke(0), without
    // storing the element matrices
    const bool grid_try = m->grid_nd && ndof == m->ndim && m->n_elem > 0 && !(flags & SFEM_ASM_FORCE_SCALAR);
    DevBuf<double> ke0, escale, bmax, dev;
    if (grid_try) {
        ke0.reset(ctx, (size_t)KE);
        escale.reset(ctx, (size_t)m->n_elem);
        bmax.reset(ctx, element_launch_blocks(m));
        dev.reset(ctx, 1);
        launch_element_zero(ctx, m, ke0.get(), vout.get());
        launch_element_checked(ctx, m, nullptr, vout.get(), ElemCheck{ke0.get(), escale.get(), bmax.get()});
        const unsigned rblocks = std::min(256u, grid_for((int64_t)bmax.size(), 1024));
        DevBuf<double> part(ctx, rblocks);
        hipLaunchKernelGGL(max_reduce_kernel, dim3(rblocks), dim3(1024), 0, ctx->stream, (const double*)bmax.get(),
                           (int64_t)bmax.size(), part.get());
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const double*)part.get(),
                           (int64_t)rblocks, dev.get());
        SFEM_LAUNCH_CHECK();
    } else {
        V.reset(ctx, (size_t)ncoo);
        launch_element_stiffness(ctx, m, V.get(), vout.get());
    }
    d2h(ctx, vcopies, vout.get(), sizeof(vcopies));
    const ValidateOut vr = combine_validate(vcopies);

    // first error in the reference's order: element-level errors, then the
    // bounds check inside sparse() on the dof ids (ArgumentError)
    unsigned long long key = vr.err_key;
    int code = key == ~0ull ? 0 : code_of_step((int)(key & 7));
    if (m->pre_code && (unsigned long long)m->pre_key < key) {
        key = (unsigned long long)m->pre_key;
        code = m->pre_code;
    }
    if (code) {
        char buf[160];
        snprintf(buf, sizeof buf, m->face ? "apply_boundary_conditions: %s at face %lld"
                                          : "assemble_system: %s at element %lld",
                 name_of_code(code), (long long)(key >> 3) + 1);
        fail(code, buf);
    }
    if (ndof > 1 && vr.min_dof == -1) fail(SFEM_ERR_ARGUMENT, "sparse: row/column index < 1 (ID entry < 1)");
    if (ndof > 1 && vr.min_dof < -1) fail(SFEM_ERR_UNSUPPORTED, "ID entry >= 2^31");

    int bs = 1;
    int64_t nunits = vr.max_unit + 1;  // nodes (ndof == 1) or dofs
    if (ndof > 1 && vr.blocked && !(flags & SFEM_ASM_FORCE_SCALAR)) {
        bs = ndof;
        nunits = vr.max_unit / ndof + 1;
    }
    K->bs = bs;
    K->nunits = m->n_elem ? nunits : 0;
    K->m = K->n = K->nunits * bs;
    const int row_bits = bits_for(K->nunits);
    if (2 * row_bits > 64) fail(SFEM_ERR_UNSUPPORTED, "too many dofs for 64-bit keys");

    // the grid's pattern when the dof numbering is the nodes'; its values from
    // the scales when every Ke(e) is s_e Ke(0) to within kGridScaleDev of
    // Ke(0)'s largest entry (SFEM_ASM_EXACT_SUMS: always the element sums)
    const bool grid_asm = grid_try && bs == ndof && K->nunits == m->n_nodes;
    double ke_dev = HUGE_VAL;
    if (grid_try) {
        K->ke0.resize((size_t)KE);
        d2h(ctx, K->ke0.data(), ke0.get(), sizeof(double) * KE);
        double dabs = 0.0;
        d2h(ctx, &dabs, dev.get(), sizeof(dabs));
        double kmax = 0.0;
        for (double v : K->ke0) kmax = std::max(kmax, std::fabs(v));
        ke_dev = (K->ke0[0] > 0.0 && kmax > 0.0 && std::isfinite(dabs)) ? dabs / kmax : HUGE_VAL;
    }
    const bool by_scale = grid_asm && !(flags & SFEM_ASM_EXACT_SUMS) && ke_dev <= kGridScaleDev;
    if (grid_try && !by_scale) {
        V.reset(ctx, (size_t)ncoo);
        launch_element_checked(ctx, m, V.get(), vout.get(), ElemCheck{ke0.get(), escale.get(), bmax.get()});
    }
    tm.mark();

    if (grid_asm) {
        // structured grid: the pattern is the grid's, no keys or sort
        GridAsm g{};
        for (int d = 0; d < 3; ++d) {
            g.e[d] = m->grid_ne[d];
            g.n[d] = d < m->grid_nd ? m->grid_ne[d] + 1 : 1;
        }
        for (int c = 0; c < 8; ++c) g.corner[c] = m->grid_corner[c];
        const int64_t nn = K->nunits;
        DevBuf<int64_t> cnt(ctx, (size_t)nn + 1);
        K->bcolptr.reset(ctx, (size_t)nn + 1);
        if (m->grid_nd == 3)
            hipLaunchKernelGGL(grid_counts_kernel<3>, dim3(grid_for(nn + 1, 256)), dim3(256), 0, ctx->stream, g, nn,
                               cnt.get());
        else
            hipLaunchKernelGGL(grid_counts_kernel<2>, dim3(grid_for(nn + 1, 256)), dim3(256), 0, ctx->stream, g, nn,
                               cnt.get());
        SFEM_LAUNCH_CHECK();
        const int64_t nnzb = exclusive_scan_i64(ctx, cnt.get(), K->bcolptr.get(), nn + 1);
        tm.mark();
        K->nnzb = nnzb;
        K->nnz = nnzb * bs * bs;
        K->browidx.reset(ctx, (size_t)nnzb);
        K->bcolidx.reset(ctx, (size_t)nnzb);
        K->bval.reset(ctx, (size_t)nnzb * bs * bs);
        K->escale = std::move(escale);
        const unsigned gblocks = grid_for(nn, 128 / 32);
        if (by_scale) {
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(grid_for(nn, kScaleThreads / 32)), dim3(kScaleThreads), 0,
                                   ctx->stream, (const double*)ke0.get(),
                                   (const double*)K->escale.get(), g, nn, (const int64_t*)K->bcolptr.get(),
                                   K->bval.get(), K->browidx.get(), K->bcolidx.get());
            };
            if (m->grid_nd == 3) go(grid_scale_kernel<3>);
            else go(grid_scale_kernel<2>);
        } else if (m->grid_nd == 3) {
            hipLaunchKernelGGL(grid_gather_kernel<3>, dim3(gblocks), dim3(128), 0, ctx->stream, V.get(), g, nn,
                               K->bcolptr.get(), K->bval.get(), K->browidx.get(), K->bcolidx.get());
        } else {
            hipLaunchKernelGGL(grid_gather_kernel<2>, dim3(gblocks), dim3(128), 0, ctx->stream, V.get(), g, nn,
                               K->bcolptr.get(), K->bval.get(), K->browidx.get(), K->bcolidx.get());
        }
        SFEM_LAUNCH_CHECK();
        tm.mark();
        K->ke_dev = ke_dev;
        K->from_scales = by_scale;
        K->grid_nd = m->grid_nd;
        for (int d = 0; d < 3; ++d) K->grid_ne[d] = m->grid_ne[d];
        for (int cc = 0; cc < 8; ++cc) K->grid_corner[cc] = m->grid_corner[cc];
        sync(ctx);
        K->t_validate = tm.ms(0, 1);
        K->t_element = tm.ms(1, 2);
        K->t_keys = tm.ms(2, 3);  // block counts and their scan
        K->t_sort = 0.0;
        K->t_reduce = tm.ms(3, 4);
        K->t_total = tm.ms(0, 4);
        return;
    }

    const int64_t npairs = (bs == 1 && ndof > 1) ? ncoo : m->n_elem * nen * nen;
    DevBuf<uint64_t> keys(ctx, (size_t)npairs), keys_alt(ctx, (size_t)npairs);
    DevBuf<uint32_t> pay(ctx, (size_t)npairs), pay_alt(ctx, (size_t)npairs);
    launch_keys(ctx, m, bs, row_bits, keys.get(), pay.get());
    tm.mark();
    uint64_t* kp = keys.get();
    uint32_t* vp = pay.get();
    radix_sort_pairs(ctx, &kp, keys_alt.get(), &vp, pay_alt.get(), npairs, 2 * row_bits);
    tm.mark();

    // segments
    K->bcolptr.reset(ctx, (size_t)K->nunits + 1);
    if (npairs == 0) {
        K->nnzb = K->nnz = 0;
        hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(K->nunits + 1, 256)), dim3(256), 0, ctx->stream,
                           K->bcolptr.get(), K->nunits + 1, (int64_t)0);
        SFEM_LAUNCH_CHECK();
    } else {
        DevBuf<uint32_t> pos(ctx, (size_t)npairs);
        DevBuf<uint32_t> dtot(ctx, 1);
        hipLaunchKernelGGL(seg_flags_kernel, dim3(grid_for(npairs, 256)), dim3(256), 0, ctx->stream, kp, npairs,
                           pos.get());
        SFEM_LAUNCH_CHECK();
        exclusive_scan_u32_dev(ctx, pos.get(), pos.get(), npairs, dtot.get());
        uint32_t nseg = 0;
        d2h(ctx, &nseg, dtot.get(), sizeof(nseg));
        K->nnzb = nseg;
        K->nnz = K->nnzb * bs * bs;
        DevBuf<uint32_t> head(ctx, (size_t)nseg + 1);
        DevBuf<uint64_t> ukey(ctx, (size_t)nseg);
        hipLaunchKernelGGL(seg_heads_kernel, dim3(grid_for(npairs, 256)), dim3(256), 0, ctx->stream, kp,
                           pos.get(), npairs, nseg, head.get(), ukey.get());
        SFEM_LAUNCH_CHECK();
        K->browidx.reset(ctx, nseg);
        K->bcolidx.reset(ctx, nseg);
        K->bval.reset(ctx, (size_t)nseg * bs * bs);
        const dim3 g(grid_for(nseg, 256));
        switch (bs) {
            case 1:
                hipLaunchKernelGGL(seg_sum_kernel<1>, g, dim3(256), 0, ctx->stream, V.get(), vp, head.get(),
                                   ukey.get(), (int64_t)nseg, ndof, row_bits, K->bval.get(), K->browidx.get(),
                                   K->bcolidx.get());
                break;
            case 2:
                hipLaunchKernelGGL(seg_sum_kernel<2>, g, dim3(256), 0, ctx->stream, V.get(), vp, head.get(),
                                   ukey.get(), (int64_t)nseg, ndof, row_bits, K->bval.get(), K->browidx.get(),
                                   K->bcolidx.get());
                break;
            default:
                hipLaunchKernelGGL(seg_sum_kernel<3>, g, dim3(256), 0, ctx->stream, V.get(), vp, head.get(),
                                   ukey.get(), (int64_t)nseg, ndof, row_bits, K->bval.get(), K->browidx.get(),
                                   K->bcolidx.get());
                break;
        }
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(colptr_kernel, g, dim3(256), 0, ctx->stream, K->bcolidx.get(), (int64_t)nseg,
                           K->nunits, K->bcolptr.get());
        SFEM_LAUNCH_CHECK();
    }
    tm.mark();
    sync(ctx);
    K->t_validate = tm.ms(0, 1);
    K->t_element = tm.ms(1, 2);
    K->t_keys = tm.ms(2, 3);
    K->t_sort = tm.ms(3, 4);
    K->t_reduce = tm.ms(4, 5);
    K->t_total = tm.ms(0, 5);
}

namespace {
__global__ void csc_keys_kernel(const int64_t* __restrict__ colptr, const int64_t* __restrict__ rowval, int64_t n,
                                int64_t m, int row_bits, uint64_t* __restrict__ keys, uint32_t* __restrict__ payload,
                                int* __restrict__ bad) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
        const int64_t r = rowval[k] - 1;
        if (r < 0 || r >= m) { atomicOr(bad, 1); continue; }
        // transposed key: (row of K) is the column unit of the stored CSC of K'
        keys[k] = ((uint64_t)r << row_bits) | (uint64_t)j;
        payload[k] = (uint32_t)k;
    }
}
}  // namespace

// Stores a host CSC of K as the block CSC (bs = 1) of K' -- i.e. the CSR of K
// -- so the solver's "columns as rows" products apply K itself even when K is
// not exactly symmetric.
void upload_csc1(sfem_ctx* ctx, int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const double* nzval, sfem_csr* K) {
    if (m != n) fail(SFEM_ERR_DIMENSION, "solve needs a square matrix");
    const int64_t nnz = colptr[n] - 1;
    if (nnz < 0 || nnz >= (int64_t(1) << 31)) fail(SFEM_ERR_UNSUPPORTED, "matrix has 2^31 or more entries");
    K->ctx = ctx;
    K->bs = 1;
    K->m = m;
    K->n = n;
    K->nunits = n;
    K->transposed = true;
    K->bcolptr.reset(ctx, (size_t)n + 1);
    if (nnz == 0) {
        K->nnz = K->nnzb = 0;
        hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(n + 1, 256)), dim3(256), 0, ctx->stream, K->bcolptr.get(),
                           n + 1, (int64_t)0);
        SFEM_LAUNCH_CHECK();
        sync(ctx);
        return;
    }
    DevBuf<int64_t> dcp(ctx, (size_t)n + 1), drv(ctx, (size_t)nnz);
    DevBuf<double> dval(ctx, (size_t)nnz);
    h2d(ctx, dcp.get(), colptr, sizeof(int64_t) * (n + 1));
    h2d(ctx, drv.get(), rowval, sizeof(int64_t) * nnz);
    h2d(ctx, dval.get(), nzval, sizeof(double) * nnz);
    const int row_bits = bits_for(n);
    DevBuf<uint64_t> keys(ctx, (size_t)nnz), keys_alt(ctx, (size_t)nnz);
    DevBuf<uint32_t> pay(ctx, (size_t)nnz), pay_alt(ctx, (size_t)nnz);
    DevBuf<int> bad(ctx, 1);
    memset_d(ctx, bad.get(), 0, sizeof(int));
    hipLaunchKernelGGL(csc_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dcp.get(), drv.get(), n, m,
                       row_bits, keys.get(), pay.get(), bad.get());
    SFEM_LAUNCH_CHECK();
    int hbad = 0;
    d2h(ctx, &hbad, bad.get(), sizeof(int));
    if (hbad) fail(SFEM_ERR_ARGUMENT, "CSC row index out of range");
    uint64_t* kp = keys.get();
    uint32_t* vp = pay.get();
    radix_sort_pairs(ctx, &kp, keys_alt.get(), &vp, pay_alt.get(), nnz, 2 * row_bits);
    DevBuf<uint32_t> pos(ctx, (size_t)nnz), dtot(ctx, 1);
    hipLaunchKernelGGL(seg_flags_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, ctx->stream, kp, nnz, pos.get());
    SFEM_LAUNCH_CHECK();
    exclusive_scan_u32_dev(ctx, pos.get(), pos.get(), nnz, dtot.get());
    uint32_t nseg = 0;
    d2h(ctx, &nseg, dtot.get(), sizeof(nseg));
    K->nnzb = K->nnz = nseg;  // repeated (row, col) pairs are summed in stored order
    DevBuf<uint32_t> head(ctx, (size_t)nseg + 1);
    DevBuf<uint64_t> ukey(ctx, (size_t)nseg);
    hipLaunchKernelGGL(seg_heads_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, ctx->stream, kp, pos.get(), nnz, nseg,
                       head.get(), ukey.get());
    SFEM_LAUNCH_CHECK();
    K->browidx.reset(ctx, nseg);
    K->bcolidx.reset(ctx, nseg);
    K->bval.reset(ctx, nseg);
    hipLaunchKernelGGL(seg_sum_kernel<1>, dim3(grid_for(nseg, 256)), dim3(256), 0, ctx->stream, dval.get(), vp,
                       head.get(), ukey.get(), (int64_t)nseg, 1, row_bits, K->bval.get(), K->browidx.get(),
                       K->bcolidx.get());
    SFEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(colptr_kernel, dim3(grid_for(nseg, 256)), dim3(256), 0, ctx->stream, K->bcolidx.get(),
                       (int64_t)nseg, n, K->bcolptr.get());
    SFEM_LAUNCH_CHECK();
    sync(ctx);
}

void export_csc1(sfem_ctx* ctx, const sfem_csr* K, int64_t* colptr, int64_t* rowval, double* nzval) {
    if (K->transposed) fail(SFEM_ERR_STATE, "export: matrix was uploaded from host CSC (the caller holds it)");
    const int64_t n = K->n, nnz = K->nnz;
    DevBuf<int64_t> dcp(ctx, (size_t)n + 1);
    hipLaunchKernelGGL(export_colptr_kernel, dim3(grid_for(n + 1, 256)), dim3(256), 0, ctx->stream,
                       K->bcolptr.get(), K->bs, n, nnz, dcp.get());
    SFEM_LAUNCH_CHECK();
    if (nnz) {
        DevBuf<int64_t> drv(ctx, (size_t)nnz);
        DevBuf<double> dnz(ctx, (size_t)nnz);
        hipLaunchKernelGGL(export_entries_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, ctx->stream,
                           K->bcolptr.get(), K->browidx.get(), K->bcolidx.get(), K->bval.get(), K->bs, nnz,
                           drv.get(), dnz.get());
        SFEM_LAUNCH_CHECK();
        if (rowval) d2h(ctx, rowval, drv.get(), sizeof(int64_t) * nnz);
        if (nzval) d2h(ctx, nzval, dnz.get(), sizeof(double) * nnz);
    }
    if (colptr) d2h(ctx, colptr, dcp.get(), sizeof(int64_t) * (n + 1));
}

}  // namespace sfem
