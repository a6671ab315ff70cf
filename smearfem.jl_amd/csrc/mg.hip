// mg.hip -- geometric multigrid preconditioned CG for node-block systems whose
// nodes form a structured grid in the reference's meshgrid numbering
// (examples/vector3D.jl:67-101: x fastest, then y, then z; 2D: x, then y).
//
// The reference solves K_free q_f = C'(f - K q_d) with a dense inverse
// (examples/vector3D.jl:316-320).  This solver reaches the same solution (the
// stopping test is on the FP64 true residual) in a number of iterations that
// does not grow with the mesh: the preconditioner is one V-cycle over a
// hierarchy of nested grids.
//
//   level 0    the system's node-block sliced operator with FP32 values and
//              its inverse diagonal blocks; vectors in block-row order.
//   level l+1  the grid coarsened by 2 along every direction whose element
//              count is even (others kept).  Operator A_{l+1} = P' A_l P, P
//              the multilinear interpolation between the nested grids,
//              restricted to free dofs on both sides (a coarse dof is fixed
//              when the fine dof at the same node is fixed).  Computed in
//              FP64 on 3^d-point node-block stencils (the grid gives every
//              column, so none is stored), symmetrised, and kept in FP32 for
//              smoothing.
//   coarsest   dense inverse in FP64, computed on the host.
//
// Smoother: Chebyshev polynomial in B^-1 A (B = the diagonal node blocks) on
// [lo, hi] x lambda_max, lambda_max from power iteration on every level;
// `degree` steps before and after the coarse correction.  The V-cycle is a
// fixed symmetric linear operator up to FP32 rounding and preconditions an
// FP64-vector CG whose operator is the FP64, split or FP32 node-block format.
//
// Every reduction is deterministic (fixed grids, per-block partials summed
// in a fixed order by the last block), so solves are bitwise reproducible.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "block_ops.hpp"
#include "sfem_objects.hpp"
#include "solver_common.hpp"

namespace sfem {

constexpr int kMgMaxDegree = 8;
constexpr int64_t kMgMaxCoarseNodes = 400;  // dense coarsest level (host inverse, O(dofs^3))

struct MgLevel {
    int64_t n[3] = {1, 1, 1};  // nodes per grid direction
    int f[3] = {1, 1, 1};      // coarsening factor from the finer level
    int64_t nn = 0;            // grid nodes
    int64_t nrows = 0;         // rows of the level's vectors (level 0: block rows)
    DevBuf<double> st64;       // masked symmetric FP64 stencil (levels >= 1)
    DevBuf<float> st32;        // smoothing operator: st64 with a unit diagonal on fixed dofs
    DevBuf<_Float16> st16;     // or (FP16 coarse levels) its FP16 values times `scale`
    bool node_major = false;   // st32 / st16 layout (st_index): lane-group kernels
    double scale = 1.0;        // 2^-e: the level's smoother works on residuals times scale
    DevBuf<float> binv;        // inverse diagonal blocks, [(j*BS+i)*nrows + row]
    DevBuf<uint8_t> fixed;     // nrows*bs: the dof is fixed
    DevBuf<float> r, x, da, db, b;
    double lmax = 0;
    int degree = 2;  // Chebyshev steps before and after the coarse correction
    float c1[kMgMaxDegree] = {}, c2[kMgMaxDegree] = {};
};

struct MgHierarchy {
    int nd = 0, bs = 0;
    int degree = 2;
    std::vector<MgLevel> L;
    DevBuf<int32_t> row_of_node;  // fine grid node -> block row (-1: every dof fixed)
    DevBuf<int32_t> node_of_row;  // block row -> fine grid node
    DevBuf<double> cinv;          // coarsest level: dense inverse, nc x nc row-major
    int64_t nc = 0;
    double setup_ms = 0;
    // FP16 finest-level smoothing operator: s * K_ff (s = 2^-e, max |s a| in
    // (1/2, 1]) in the system's sliced layout, and the inverses of its
    // diagonal blocks; the finest-level smoother then works in units of s
    bool fp16 = false;
    double scale0 = 1.0;
    DevBuf<_Float16> h0;
    DevBuf<float> binv0;
    // Matrix-free finest level on a grid whose element matrices are s_e Ke(0)
    // (K's ke_dev): 1 the smoother, 2 also the outer FP64 operator.  kc:
    // Ke(0) by corner codes; es32: the scales in FP32 (FP64: K's escale).
    int mf = 0;
    double ke_dev = HUGE_VAL;
    DevBuf<float> kc32, es32;     // es: the scales with a ghost layer
    DevBuf<double> kc64, es64;
    DevBuf<int32_t> mf_rows;      // row_of_node with a ghost layer
    DevBuf<uint32_t> mf_offs;     // its rows as byte offsets in an FP32 vector
    // the smoothing steps on the matrix cores (3D): Ke(0) 2^-ek as FP16 hi / lo
    // fragments of the 32x32x8 instruction, [part][k step][lane][4]
    bool mfma = false;
    DevBuf<_Float16> kfrag;
    float kscale = 1.0f;          // 2^ek
};

void MgDelete::operator()(MgHierarchy* h) const { delete h; }

namespace {

// ------------------------------------------------------------ grid --
struct Grid {
    int64_t n0, n1, n2;  // nodes per direction (n2 = 1 in 2D)
};

template <int ND>
struct Sten {
    static constexpr int NS = ND == 3 ? 27 : 9;  // stencil slots
    static constexpr int C = NS / 2;             // slot of offset 0
};

template <int ND>
__host__ __device__ __forceinline__ void coords_of(int64_t node, const Grid& g, int64_t c[3]) {
    c[0] = node % g.n0;
    const int64_t t = node / g.n0;
    c[1] = t % g.n1;
    c[2] = ND == 3 ? t / g.n1 : 0;
}

__host__ __device__ __forceinline__ int64_t node_at(const Grid& g, int64_t x, int64_t y, int64_t z) {
    return x + g.n0 * (y + g.n1 * z);
}

template <int ND>
__host__ __device__ __forceinline__ bool inside(const Grid& g, int64_t x, int64_t y, int64_t z) {
    return x >= 0 && x < g.n0 && y >= 0 && y < g.n1 && (ND == 2 ? z == 0 : (z >= 0 && z < g.n2));
}

// slot of the offset (dx, dy, dz), each in {-1, 0, 1} (dz = 0 in 2D); the
// slot of the opposite offset is NS - 1 - slot
template <int ND>
__host__ __device__ __forceinline__ int slot_of(int dx, int dy, int dz) {
    return (dx + 1) + 3 * (dy + 1) + (ND == 3 ? 9 * (dz + 1) : 0);
}

template <int ND>
__host__ __device__ __forceinline__ void offset_of(int s, int& dx, int& dy, int& dz) {
    dx = s % 3 - 1;
    dy = (s / 3) % 3 - 1;
    dz = ND == 3 ? s / 9 - 1 : 0;
}

// Nested coarsening along one direction with e fine elements: coarse node c
// sits on fine node min(2c, e), so e elements become ceil(e/2) (for odd e the
// last coarse element spans one fine element).  child() gives the k-th
// (k = 0, 1, 2) fine node of coarse node c's interpolation support and its
// weight: 2c-1, 2c, 2c+1 with 1/2, 1, 1/2, except that for odd e the last fine
// node belongs to the last coarse node alone (weight 1).
__host__ __device__ __forceinline__ bool child(int64_t c, int k, int64_t e, int64_t& i, double& w) {
    if ((e & 1) && c == (e + 1) / 2) {
        i = e;
        w = 1.0;
        return k == 1;
    }
    const int a = k - 1;
    i = 2 * c + a;
    if (i < 0 || i > e || ((e & 1) && i == e)) return false;
    w = a ? 0.5 : 1.0;
    return true;
}

// ------------------------------------------------------- hierarchy --
// The fine stencil from the element scales (matrix-free finest level): slot
// s of node I is the sum, over the cells I shares with its neighbour, of
// s_e Ke(0)'s block (I's corner, the neighbour's corner); fixed rows and
// columns zero, as stencil_from_k.  es: the scales with a ghost layer of
// zero cells.  c: the node's grid coordinates.
template <int ND>
__device__ __forceinline__ void mf_stencil_block(const double* __restrict__ kc, const double* __restrict__ es,
                                                 const int32_t* __restrict__ free_of, const Grid& g, int s,
                                                 int64_t node, const int64_t (&c)[3], double (&v)[ND * ND]) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND;
    int d[3];
    offset_of<ND>(s, d[0], d[1], d[2]);
    const bool in = c[0] + d[0] >= 0 && c[0] + d[0] < g.n0 && c[1] + d[1] >= 0 && c[1] + d[1] < g.n1 &&
                    (ND == 2 || (c[2] + d[2] >= 0 && c[2] + d[2] < g.n2));
    double acc[B2];
#pragma unroll
    for (int m = 0; m < B2; ++m) acc[m] = 0.0;
    if (in) {
        const int64_t q0 = g.n0 + 1, q1 = g.n1 + 1;  // padded cells per direction
#pragma unroll
        for (int q = 0; q < NEN; ++q) {  // cell c - 1 + (qx, qy, qz)
            const int qx = q & 1, qy = (q >> 1) & 1, qz = ND == 3 ? (q >> 2) & 1 : 0;
            const int bx = d[0] + 1 - qx, by = d[1] + 1 - qy, bz = ND == 3 ? d[2] + 1 - qz : 0;
            if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;  // not in cell q
            const int a = (1 - qx) + 2 * (1 - qy) + (ND == 3 ? 4 * (1 - qz) : 0);
            const int b = bx + 2 * by + 4 * bz;
            const double se = es[(c[0] + qx) + q0 * ((c[1] + qy) + (ND == 3 ? q1 * (c[2] + qz) : 0))];
#pragma unroll
            for (int m = 0; m < B2; ++m) acc[m] = fma(se, kc[(a * NEN + b) * B2 + m], acc[m]);
        }
    }
    const int64_t nb = in ? node + d[0] + g.n0 * (d[1] + g.n1 * d[2]) : node;
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            const bool masked = !in || free_of[node * BS + i] < 0 || free_of[nb * BS + j] < 0;
            v[i * BS + j] = masked ? 0.0 : acc[i * BS + j];
        }
}

// The input of a Galerkin line product: a stored slot-major stencil, or the
// matrix-free finest level's blocks (mf_stencil_block), formed where they are
// read instead of stored.
template <int ND>
struct StencilMem {
    const double* st;
    int64_t nn;
    __device__ __forceinline__ void load(int s, int64_t node, const int64_t (&)[3], double (&v)[ND * ND]) const {
        const double* src = st + (int64_t)s * ND * ND * nn + node;
#pragma unroll
        for (int m = 0; m < ND * ND; ++m) v[m] = src[(int64_t)m * nn];
    }
};

template <int ND>
struct StencilMf {
    const double* kc;
    const double* es;
    const int32_t* free_of;
    Grid g;
    __device__ __forceinline__ void load(int s, int64_t node, const int64_t (&c)[3], double (&v)[ND * ND]) const {
        mf_stencil_block<ND>(kc, es, free_of, g, s, node, c, v);
    }
};

// Level-0 maps: a node with any free dof is a block row (rank in node order,
// as the elimination numbers them).
__global__ void node_free_kernel(const int32_t* __restrict__ free_of, int64_t nn, int bs,
                                 uint32_t* __restrict__ flag) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nn) return;
    int c = 0;
    for (int i = 0; i < bs; ++i) c += free_of[u * bs + i] >= 0 ? 1 : 0;
    flag[u] = c > 0 ? 1u : 0u;
}

__global__ void row_maps_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                const int32_t* __restrict__ free_of, int64_t nn, int bs,
                                int32_t* __restrict__ row_of_node, int32_t* __restrict__ node_of_row,
                                uint8_t* __restrict__ fixed0, uint8_t* __restrict__ fixnode) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nn) return;
    const int32_t row = flag[u] ? (int32_t)pos[u] : -1;
    row_of_node[u] = row;
    for (int i = 0; i < bs; ++i) fixnode[u * bs + i] = free_of[u * bs + i] < 0 ? 1 : 0;
    if (row < 0) return;
    node_of_row[row] = (int32_t)u;
    for (int i = 0; i < bs; ++i) fixed0[(int64_t)row * bs + i] = free_of[u * bs + i] < 0 ? 1 : 0;
}

// Fine stencil from K's node blocks, one thread per stored block (u, in
// column c = bcolidx[u], row r = browidx[u]): K is symmetric, so the block
// gives row c's slot of offset r - c,
// st[(s*B2 + i*BS + j)*nn + c] = K[(c,i),(r,j)] = bval[u*B2 + j*BS + i]
// (exact for K stored transposed as well), 0 where either dof is fixed.  The
// caller zeroes st first (slots without a block).  A block between nodes
// that are not grid neighbours sets *bad.
template <int ND>
__global__ void stencil_from_k_kernel(const int32_t* __restrict__ bcolidx, const int32_t* __restrict__ browidx,
                                      const double* __restrict__ bval, int64_t nnzb, int64_t nn,
                                      const int32_t* __restrict__ free_of, Grid g, double* __restrict__ st,
                                      int* __restrict__ bad) {
    constexpr int BS = ND, B2 = BS * BS;
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nnzb) return;
    const int64_t c = bcolidx[u], r = browidx[u];
    int64_t a[3], b[3];
    coords_of<ND>(c, g, a);
    coords_of<ND>(r, g, b);
    const int64_t dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) {
        atomicOr(bad, 1);
        return;
    }
    const int s = slot_of<ND>((int)dx, (int)dy, (int)dz);
    double v[B2];
#pragma unroll
    for (int m = 0; m < B2; ++m) v[m] = bval[u * B2 + m];
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            const bool masked = free_of[c * BS + i] < 0 || free_of[r * BS + j] < 0;
            st[((int64_t)s * B2 + i * BS + j) * nn + c] = masked ? 0.0 : v[j * BS + i];
        }
}

// Coarse dof fixed <=> the fine dof at the same node is fixed.
template <int ND>
__global__ void coarse_fixed_kernel(const uint8_t* __restrict__ ffix, Grid gf, Grid gc, int f0, int f1, int f2,
                                    uint8_t* __restrict__ cfix) {
    constexpr int BS = ND;
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= gc.n0 * gc.n1 * gc.n2) return;
    int64_t c[3];
    coords_of<ND>(I, gc, c);
    // coarse node c sits on fine node min(2c, e) along a coarsened direction
    const int64_t x = f0 == 2 ? min(2 * c[0], gf.n0 - 1) : c[0];
    const int64_t y = f1 == 2 ? min(2 * c[1], gf.n1 - 1) : c[1];
    const int64_t z = f2 == 2 ? min(2 * c[2], gf.n2 - 1) : c[2];
    const int64_t fn = node_at(gf, x, y, z);
    for (int a = 0; a < BS; ++a) cfix[I * BS + a] = ffix[fn * BS + a];
}

// Galerkin product along one direction: P = P_x (x) P_y (x) P_z, so P'AP
// is three one-direction products, each coarsening direction d (child())
// and keeping the others:
// A'(I, I+D)[a][b] = sum over the support pairs (i_d of I_d, j_d of I_d+D_d)
// of w_i w_j A(i, i+o)[a][b], with i = I except along d, o = D except
// o_d = j_d - i_d (pairs more than one apart add nothing).  One thread per
// (coarse node, line of the 3 slots that differ only along d): the line's
// outputs read the same <= 9 input blocks (3 fine children x 3 offsets along
// d), each once.  Each output sums its terms in ascending order of the child
// index, then of the neighbour's fine index.
template <int ND, typename SRC>
__global__ __launch_bounds__(kThreads) void galerkin_line_kernel(SRC in, Grid gi, int d, Grid go,
                                                              double* __restrict__ out) {
    constexpr int BS = ND, B2 = BS * BS;
    const int64_t nno = go.n0 * go.n1 * go.n2;
    // the line (the offsets off direction d) is the block's y index: uniform,
    // so the slots and Ke(0) blocks it addresses are scalar loads
    const int l = (int)blockIdx.y;
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nno) return;
    int64_t c[3];
    coords_of<ND>(I, go, c);
    int D[3] = {0, 0, 0};
    {
        int q = l;
        for (int dd = 0; dd < ND; ++dd) {
            if (dd == d) continue;
            D[dd] = q % 3 - 1;
            q /= 3;
        }
    }
    bool ok[3];
    for (int k = 0; k < 3; ++k) {
        int64_t cc[3] = {c[0] + D[0], c[1] + D[1], c[2] + D[2]};
        cc[d] = c[d] + k - 1;
        ok[k] = inside<ND>(go, cc[0], cc[1], cc[2]);
    }
    double acc[3][B2];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < B2; ++m) acc[k][m] = 0.0;
    const int64_t e = (d == 0 ? gi.n0 : d == 1 ? gi.n1 : gi.n2) - 1;
    if (ok[0] || ok[1] || ok[2]) {
        for (int ka = 0; ka < 3; ++ka) {
            int64_t ia;
            double wa;
            if (!child(c[d], ka, e, ia, wa)) continue;
            for (int o = -1; o <= 1; ++o) {
                const int64_t jb = ia + o;
                if (jb < 0 || jb > e) continue;
                int64_t ic[3] = {c[0], c[1], c[2]};
                ic[d] = ia;
                int od[3] = {D[0], D[1], D[2]};
                od[d] = o;
                const int src_slot = slot_of<ND>(od[0], od[1], od[2]);
                const int64_t src_node = node_at(gi, ic[0], ic[1], ic[2]);
                double v[B2];
                bool loaded = false;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (!ok[k]) continue;
                    // jb a child of the coarse neighbour c[d] + k - 1, with weight wb
                    for (int kb = 0; kb < 3; ++kb) {
                        int64_t jj;
                        double wb;
                        if (!child(c[d] + k - 1, kb, e, jj, wb) || jj != jb) continue;
                        if (!loaded) {
                            in.load(src_slot, src_node, ic, v);
                            loaded = true;
                        }
                        const double w = wa * wb;
#pragma unroll
                        for (int m = 0; m < B2; ++m) acc[k][m] = fma(w, v[m], acc[k][m]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int od[3] = {D[0], D[1], D[2]};
        od[d] = k - 1;
        const int s = slot_of<ND>(od[0], od[1], od[2]);
#pragma unroll
        for (int m = 0; m < B2; ++m) out[((int64_t)s * B2 + m) * nno + I] = acc[k][m];
    }
}

// ------------------------- first coarse level from the element scales --
// With every direction coarsened by 2, a coarse cell E (lower corner Ec)
// holds the fine elements 2 Ec + u, u in {0, 1}^ND, and P restricted to a
// fine element interpolates its corners from E's corners alone: along a
// direction, fine corner k of sub-element u sits at t = (u + k) / 2 of the
// cell, weight 1 - t for E's lower corner and t for the upper one.  So
// P'AP = sum over cells E, over their fine elements e, of s_e W_e' Ke(0)_m W_e
// (Ke(0)_m: Ke(0) with the rows and columns of e's fixed dofs zero).  A cell
// is regular when it has two fine elements along every direction and no
// fixed dof among its fine nodes; its elements' W_e' Ke(0) W_e depend on u
// alone: gt[u][a][b] (a, b: the corners of E, 2^ND x 2^ND blocks; computed
// from Ke(0) by galerkin_table_kernel).  The last cell of an odd element
// count (one fine element, the identity weights) and cells with a fixed dof
// are irregular: their cell matrices are summed first (cell_matrix_kernel),
// the fine elements with a fixed dof (all, in a one-element cell) from Ke(0)
// with that element's masked weights.  This replaces the one-direction
// products and the symmetrising pass for that level.

// gt[((u NEN + a) NEN + b) B2 + m] = sum over e's corners i, then j, of
// w_a(i) w_b(j) Ke(0)[i][j][m], for m = (r, t) with a's dof r at most b's
// dof t; the other entries are their transposes' sums (bit for bit
// symmetric).  One thread per entry.
template <int ND>
__global__ void galerkin_table_kernel(const double* __restrict__ kc, double* __restrict__ gt) {
    constexpr int B2 = ND * ND, NEN = 1 << ND;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= NEN * NEN * NEN * B2) return;
    int m = t % B2, b = (t / B2) % NEN, a = (t / (B2 * NEN)) % NEN;
    const int u = t / (B2 * NEN * NEN);
    if (a * ND + m / ND > b * ND + m % ND) {  // summed as the transposed entry (b, a)
        const int r = m / ND;
        m = (m % ND) * ND + r;
        const int x = a;
        a = b;
        b = x;
    }
    auto w = [u](int c, int k) {  // weight of corner c of E at fine corner k of sub-element u
        double p = 1.0;
        for (int d = 0; d < ND; ++d) {
            const double x = 0.5 * (((u >> d) & 1) + ((k >> d) & 1));
            p *= ((c >> d) & 1) ? x : 1.0 - x;
        }
        return p;
    };
    double acc = 0.0;
    for (int i = 0; i < NEN; ++i)
        for (int j = 0; j < NEN; ++j) acc = fma(w(a, i) * w(b, j), kc[(i * NEN + j) * B2 + m], acc);
    gt[t] = acc;
}

// The irregular coarse cells (see above) numbered in any order: slot[E] =
// k >= 0 and cells[k] = E for an irregular cell E, slot[E] = -1 otherwise
template <int ND>
__global__ void cell_irregular_kernel(const uint8_t* __restrict__ ffix, Grid gf, Grid gc, int32_t* __restrict__ slot,
                                      int32_t* __restrict__ cells, unsigned* __restrict__ count) {
    constexpr int BS = ND;
    const int64_t c0 = gc.n0 - 1, c1 = gc.n1 - 1, c2 = ND == 3 ? gc.n2 - 1 : 1;
    const int64_t E = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (E >= c0 * c1 * c2) return;
    const int64_t ec[3] = {E % c0, (E / c0) % c1, ND == 3 ? E / (c0 * c1) : 0};
    const int64_t ef[3] = {gf.n0 - 1, gf.n1 - 1, gf.n2 - 1};
    bool bad = false;
    int64_t hi[3] = {0, 0, 0};
    for (int d = 0; d < ND; ++d) {
        hi[d] = 2 * ec[d] + 2;
        if (hi[d] > ef[d]) {  // one fine element
            bad = true;
            hi[d] = ef[d];
        }
    }
    for (int64_t z = 2 * ec[2]; z <= hi[2]; ++z)
        for (int64_t y = 2 * ec[1]; y <= hi[1]; ++y)
            for (int64_t x = 2 * ec[0]; x <= hi[0]; ++x) {
                const int64_t fn = node_at(gf, x, y, z);
                for (int a = 0; a < BS; ++a) bad = bad || ffix[fn * BS + a];
            }
    int32_t k = -1;
    if (bad) {
        k = (int32_t)atomicAdd(count, 1u);
        cells[k] = (int32_t)E;
    }
    slot[E] = k;
}

// The cell matrix of each irregular cell, kcell[k][a][b] = sum over its fine
// elements (u ascending) of s_e T_u(a, b): T_u = gt[u] for an element with
// no fixed dof in a cell of two fine elements per direction, else
// sum over i, then j, of (w_a(i) [i's dof free]) (w_b(j) [j's dof free]) Ke(0)[i][j]
// with the element's own weights (1 - t, t; the identity along a
// one-element direction).  Each entry (row dof of corner a, column dof of
// corner b) is summed in the order of the entry whose row dof is the lower
// one, so kcell[k][a][b] and kcell[k][b][a] are transposes bit for bit.
// One wavefront per cell, lane (a, b); fine nodes outside a one-element
// cell are never read.  Ke(0) is staged in LDS; the inner corner's
// weights are formed once per element.
template <int ND>
__global__ __launch_bounds__(64) void cell_matrix_kernel(const double* __restrict__ gt, const double* __restrict__ kc,
                                                         const double* __restrict__ es,
                                                         const uint8_t* __restrict__ ffix, Grid gf, Grid gc,
                                                         const int32_t* __restrict__ cells,
                                                         double* __restrict__ kcell) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND;
    __shared__ double kl[NEN * NEN * B2];
    for (int t = (int)threadIdx.x; t < NEN * NEN * B2; t += 64) kl[t] = kc[t];
    __syncthreads();
    const int k = (int)blockIdx.x;
    const int lane = (int)threadIdx.x;
    if (lane >= NEN * NEN) return;
    const int a = lane / NEN, b = lane % NEN;
    const int64_t E = cells[k];
    const int64_t c0 = gc.n0 - 1, c1 = gc.n1 - 1;
    const int64_t ec[3] = {E % c0, (E / c0) % c1, ND == 3 ? E / (c0 * c1) : 0};
    const int64_t p0 = gf.n0 + 1, p1 = gf.n1 + 1;  // padded fine cells per direction
    const int64_t ef[3] = {gf.n0 - 1, gf.n1 - 1, gf.n2 - 1};
    int ns[3] = {1, 1, 1};  // fine elements along each direction
    bool two = true;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        ns[d] = 2 * ec[d] + 2 <= ef[d] ? 2 : 1;
        two = two && ns[d] == 2;
    }
    double acc[B2];
#pragma unroll
    for (int m = 0; m < B2; ++m) acc[m] = 0.0;
#pragma unroll 1
    for (int u = 0; u < NEN; ++u) {
        const int ud[3] = {u & 1, (u >> 1) & 1, ND == 3 ? (u >> 2) & 1 : 0};
        if (ud[0] >= ns[0] || ud[1] >= ns[1] || ud[2] >= ns[2]) continue;
        int64_t fc[3];  // the fine element's lower corner
#pragma unroll
        for (int d = 0; d < 3; ++d) fc[d] = d < ND ? 2 * ec[d] + ud[d] : 0;
        const double se = es[(fc[0] + 1) + p0 * ((fc[1] + 1) + (ND == 3 ? p1 * (fc[2] + 1) : 0))];
        uint32_t fx = 0;  // bit n BS + j: corner n's dof j fixed
#pragma unroll
        for (int n = 0; n < NEN; ++n) {
            const int64_t fn =
                node_at(gf, fc[0] + (n & 1), fc[1] + ((n >> 1) & 1), ND == 3 ? fc[2] + ((n >> 2) & 1) : 0);
#pragma unroll
            for (int j = 0; j < BS; ++j) fx |= (ffix[fn * BS + j] ? 1u : 0u) << (n * BS + j);
        }
        double T[B2];
        if (two && !fx) {
            const double* g = gt + (((int64_t)u * NEN + a) * NEN + b) * B2;
#pragma unroll
            for (int m = 0; m < B2; ++m) T[m] = g[m];
        } else {
            // corner c's weight at the element's corner n
            auto wt = [&](int c, int n) {
                double p = 1.0;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const int nd = (n >> d) & 1, cd = (c >> d) & 1;
                    const double t = 0.5 * (ud[d] + nd);
                    p *= ns[d] == 2 ? (cd ? t : 1.0 - t) : (cd == nd ? 1.0 : 0.0);
                }
                return p;
            };
            double gaj[NEN], gbj[NEN];  // the inner corner's weights, formed once
#pragma unroll
            for (int j = 0; j < NEN; ++j) {
                gaj[j] = wt(a, j);
                gbj[j] = wt(b, j);
            }
#pragma unroll
            for (int m = 0; m < B2; ++m) T[m] = 0.0;
#pragma unroll 1
            for (int i = 0; i < NEN; ++i) {
                const double gai = wt(a, i), gbi = wt(b, i);
                const double* ki = kl + i * NEN * B2;
#pragma unroll
                for (int j = 0; j < NEN; ++j) {
                    const double* kb = ki + j * B2;
#pragma unroll
                    for (int r = 0; r < BS; ++r)
#pragma unroll
                        for (int t = 0; t < BS; ++t) {
                            // (a, r) first: w_a(i) w_b(j) Ke[i][j][r][t]; else the
                            // transposed entry's sum, w_b(i) w_a(j) Ke[i][j][t][r]
                            const bool lo = a * BS + r <= b * BS + t;
                            const bool fir = (fx >> (i * BS + r)) & 1u, fit = (fx >> (i * BS + t)) & 1u;
                            const bool fjr = (fx >> (j * BS + r)) & 1u, fjt = (fx >> (j * BS + t)) & 1u;
                            const double x = lo ? (fir ? 0.0 : gai) : (fit ? 0.0 : gbi);
                            const double y = lo ? (fjt ? 0.0 : gbj[j]) : (fjr ? 0.0 : gaj[j]);
                            T[r * BS + t] = fma(x * y, lo ? kb[r * BS + t] : kb[t * BS + r], T[r * BS + t]);
                        }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < B2; ++m) acc[m] = fma(se, T[m], acc[m]);
    }
    double* o = kcell + ((int64_t)k * NEN * NEN + lane) * B2;
#pragma unroll
    for (int m = 0; m < B2; ++m) o[m] = acc[m];
}

// The cells holding coarse node c and its neighbour c + D: q in {0,1}^ND
// names the cell with lower corner c - 1 + q; it holds the neighbour when
// its corner b = 1 - q + D (per direction) is 0 or 1, and exists when its
// lower corner is inside the cell grid.  a = 1 - q: c's corner.
template <int ND>
__device__ __forceinline__ bool shared_cell(const Grid& gc, const int64_t (&c)[3], const int (&D)[3], int q,
                                            int64_t (&ec)[3], int& a, int& b) {
    a = b = 0;
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
        ec[d] = 0;
        if (d >= ND) continue;
        const int qd = (q >> d) & 1;
        const int bd = 1 - qd + D[d];
        const int64_t n = d == 0 ? gc.n0 : d == 1 ? gc.n1 : gc.n2;
        ec[d] = c[d] - 1 + qd;
        ok = ok && bd >= 0 && bd <= 1 && ec[d] >= 0 && ec[d] <= n - 2;
        a |= (1 - qd) << d;
        b |= bd << d;
    }
    return ok;
}

// A coarse smoothing stencil's value (slot s, entry m) of a node: slot-major
// [(s*B2 + m)*nn + node] on the large levels (one thread per node, loads
// coalesced along nodes), node-major [(node*NS + s)*B2 + m] on the small ones
// (a lane group per node reads one contiguous run).
template <int ND>
__host__ __device__ __forceinline__ int64_t st_index(bool node_major, int64_t nn, int64_t node, int s, int m) {
    constexpr int B2 = ND * ND, NS = Sten<ND>::NS;
    return node_major ? (node * NS + s) * B2 + m : ((int64_t)s * B2 + m) * nn + node;
}

// The first coarse level's stencil, FP64 (slot-major) and the smoother's
// FP32 copy, as sym_mask_kernel leaves them.  Thread (coarse node, slot =
// blockIdx.y): the block sums, over the cells shared with the slot's
// neighbour (q ascending), a regular cell's fine elements' s_e gt[u][a][b]
// (u ascending) or an irregular cell's kcell block.  gt and kcell are
// symmetric bit for bit and both blocks (I, J) and (J, I) meet the same
// cells in the same order, so the stencil needs no symmetrising.
template <int ND>
__global__ __launch_bounds__(kThreads) void galerkin_cells_kernel(
    const double* __restrict__ gt, const double* __restrict__ kcell, const double* __restrict__ es,
    const int32_t* __restrict__ slot, Grid gf, Grid gc, const uint8_t* __restrict__ cfix,
    double* __restrict__ st64, float* __restrict__ st32, bool node_major) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND;
    const int s = (int)blockIdx.y;
    const int64_t nno = gc.n0 * gc.n1 * gc.n2;
    const int64_t I = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (I >= nno) return;
    int D[3];
    offset_of<ND>(s, D[0], D[1], D[2]);
    const int64_t c0 = gc.n0 - 1, c1 = gc.n1 - 1;
    const int64_t p0 = gf.n0 + 1, p1 = gf.n1 + 1;  // padded fine cells per direction
    int64_t c[3];
    coords_of<ND>(I, gc, c);
    double acc[B2];
#pragma unroll
    for (int m = 0; m < B2; ++m) acc[m] = 0.0;
    // (cells and elements not unrolled: the table blocks are scalar loads,
    // a few at a time)
#pragma unroll 1
    for (int q = 0; q < NEN; ++q) {
        int64_t ec[3];
        int a, b;
        if (!shared_cell<ND>(gc, c, D, q, ec, a, b)) continue;
        const int32_t k = slot[ec[0] + c0 * (ec[1] + c1 * ec[2])];
        if (k >= 0) {
            const double* o = kcell + ((int64_t)k * NEN * NEN + a * NEN + b) * B2;
#pragma unroll
            for (int m = 0; m < B2; ++m) acc[m] += o[m];
            continue;
        }
#pragma unroll 2
        for (int u = 0; u < NEN; ++u) {
            const int64_t fe = (2 * ec[0] + (u & 1) + 1) +
                               p0 * ((2 * ec[1] + ((u >> 1) & 1) + 1) +
                                     (ND == 3 ? p1 * (2 * ec[2] + ((u >> 2) & 1) + 1) : 0));
            const double se = es[fe];
            const double* g = gt + (((int64_t)u * NEN + a) * NEN + b) * B2;
#pragma unroll
            for (int m = 0; m < B2; ++m) acc[m] = fma(se, g[m], acc[m]);
        }
    }
    // fixed coarse dofs: zero rows and columns, unit diagonal in the FP32 copy
    const bool valid = inside<ND>(gc, c[0] + D[0], c[1] + D[1], c[2] + D[2]);
    const int64_t J = valid ? node_at(gc, c[0] + D[0], c[1] + D[1], c[2] + D[2]) : I;
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int t = 0; t < BS; ++t) {
            const bool masked = cfix[I * BS + r] || cfix[J * BS + t];
            const double v = valid && !masked ? acc[r * BS + t] : 0.0;
            st64[((int64_t)s * B2 + r * BS + t) * nno + I] = v;
            st32[st_index<ND>(node_major, nno, I, s, r * BS + t)] =
                valid && masked && s == Sten<ND>::C && r == t ? 1.0f : (float)v;
        }
}

// Symmetrise (the mean of A(I,J) and A(J,I)', the same bits from either
// side), zero the rows and columns of fixed dofs (st64, slot-major, the next
// Galerkin product's input) and put a unit diagonal on them (st32, the
// smoother's operator, in the level's layout).  Thread (node, slot s =
// blockIdx.y): a wavefront reads and writes 64 consecutive nodes of a slot.
template <int ND>
__global__ __launch_bounds__(kThreads) void sym_mask_kernel(const double* __restrict__ raw, Grid g,
                                                            const uint8_t* __restrict__ fixed,
                                                            double* __restrict__ st64, float* __restrict__ st32,
                                                            bool node_major) {
    constexpr int BS = ND, B2 = BS * BS, NS = Sten<ND>::NS;
    const int64_t nn = g.n0 * g.n1 * g.n2;
    const int s = (int)blockIdx.y;
    const int64_t I = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (I >= nn) return;
    int64_t c[3];
    coords_of<ND>(I, g, c);
    int dx, dy, dz;
    offset_of<ND>(s, dx, dy, dz);
    const bool valid = inside<ND>(g, c[0] + dx, c[1] + dy, c[2] + dz);
    const int64_t J = valid ? node_at(g, c[0] + dx, c[1] + dy, c[2] + dz) : I;
    const int sp = NS - 1 - s;
    for (int a = 0; a < BS; ++a)
        for (int b = 0; b < BS; ++b) {
            double v = 0.0;
            bool unit = false;
            if (valid) {
                const bool masked = fixed[I * BS + a] || fixed[J * BS + b];
                v = masked ? 0.0
                           : 0.5 * (raw[((int64_t)s * B2 + a * BS + b) * nn + I] +
                                    raw[((int64_t)sp * B2 + b * BS + a) * nn + J]);
                unit = masked && s == Sten<ND>::C && a == b;
            }
            st64[((int64_t)s * B2 + a * BS + b) * nn + I] = v;
            st32[st_index<ND>(node_major, nn, I, s, a * BS + b)] = unit ? 1.0f : (float)v;
        }
}

template <int ND>
__global__ void stencil_binv_kernel(const double* __restrict__ st64, const uint8_t* __restrict__ fixed, int64_t nn,
                                    float* __restrict__ binv) {
    constexpr int BS = ND, B2 = BS * BS, C = Sten<ND>::C;
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    double m[B2], inv[B2];
    for (int a = 0; a < BS; ++a)
        for (int b = 0; b < BS; ++b) {
            const double v = st64[((int64_t)C * B2 + a * BS + b) * nn + node];
            m[a * BS + b] = (fixed[node * BS + a] && a == b) ? 1.0 : v;
        }
    invert_block<BS>(m, inv);
    for (int k = 0; k < B2; ++k) binv[(int64_t)k * nn + node] = (float)inv[k];
}

// max |a| of K_ff: K_ff is symmetric positive definite, so |a_ij| <=
// sqrt(a_ii a_jj) and the largest entry is a diagonal one, max 1 / dinv
// (exact: the bit patterns of non-negative doubles order as unsigned integers)
__global__ void maxdiag_kernel(const double* __restrict__ dinv, int64_t n, unsigned long long* __restrict__ out) {
    unsigned long long m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (unsigned long long)__double_as_longlong(fabs(1.0 / dinv[i])));
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned long long)__shfl_xor((long long)m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// FP16 of the FP32 values times a power of two (the scaling is exact)
__global__ void to_half_kernel(const float* __restrict__ v, int64_t n, float scale, _Float16* __restrict__ h) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        h[i] = (_Float16)(v[i] * scale);
}

// The scale 2^-e of a level's FP16 stencil, amax = f 2^e with f in [1/2, 1)
// (amax as stencil_maxdiag_kernel's bit pattern); 0 when amax is zero or not
// finite
__global__ void half_scale_kernel(const unsigned* __restrict__ amax_bits, double* __restrict__ scale) {
    const float amax = __uint_as_float(*amax_bits);
    double sc = 0.0;
    if (amax > 0.0f && isfinite(amax)) {
        int e = 0;
        (void)frexp((double)amax, &e);
        sc = ldexp(1.0, -e);
    }
    *scale = sc;
}

// to_half_kernel with the scale read from the device
__global__ void to_half_dev_kernel(const float* __restrict__ v, int64_t n, const double* __restrict__ scale,
                                   _Float16* __restrict__ h) {
    const float sc = (float)*scale;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        h[i] = (_Float16)(v[i] * sc);
}

// Inverse of each block row's diagonal block, read from the FP16 values.
template <int BS>
__global__ void binv_half_kernel(const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                 const int32_t* __restrict__ scol, const _Float16* __restrict__ h, int64_t nrows,
                                 float* __restrict__ binv) {
    constexpr int B2 = BS * BS;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const int64_t sl = row / 64;
    const int lane = (int)(row % 64);
    double m[B2], inv[B2];
    for (int k = 0; k < B2; ++k) m[k] = (k % (BS + 1) == 0) ? 1.0 : 0.0;
    const int w = sl_w[sl];
    for (int k = 0; k < w; ++k) {
        const int64_t slot = sl_ptr[sl] + (int64_t)k * 64;
        if (scol[slot + lane] != row) continue;
        for (int v = 0; v < B2; ++v) m[v] = (double)h[slot * B2 + (int64_t)v * 64 + lane];
        break;
    }
    invert_block<BS>(m, inv);
    for (int k = 0; k < B2; ++k) binv[(int64_t)k * nrows + row] = (float)inv[k];
}

// max over the diagonal of an FP32 smoothing stencil (non-negative: SPD
// operator with unit diagonals on fixed dofs), as ordered bit patterns
template <int ND>
__global__ void stencil_maxdiag_kernel(const float* __restrict__ st32, int64_t nn, bool node_major,
                                       unsigned* __restrict__ out) {
    constexpr int BS = ND, B2 = BS * BS, NS = Sten<ND>::NS, C = Sten<ND>::C;
    unsigned m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn * BS;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t node = i / BS;
        const int a = (int)(i % BS);
        m = max(m, (unsigned)__float_as_int(fabsf(st32[st_index<ND>(node_major, nn, node, C, a * BS + a)])));
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// inverse diagonal blocks of an FP16 smoothing stencil (the values as
// stored, i.e. times the level's scale), planar [(j*BS+i)*nn + node]
template <int ND>
__global__ void stencil_binv_half_kernel(const _Float16* __restrict__ st16, int64_t nn, bool node_major,
                                         float* __restrict__ binv) {
    constexpr int BS = ND, B2 = BS * BS, NS = Sten<ND>::NS, C = Sten<ND>::C;
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    double m[B2], inv[B2];
    for (int k = 0; k < B2; ++k) m[k] = (double)st16[st_index<ND>(node_major, nn, node, C, k)];
    invert_block<BS>(m, inv);
    for (int k = 0; k < B2; ++k) binv[(int64_t)k * nn + node] = (float)inv[k];
}

// ------------------------------------------------------- V-cycle ----
// finest-level smoothing operator: the system's sliced format with FP32
// values, or FP16 values of the operator scaled by 1 / max |a|
template <typename TV>
struct SellOp {
    const int64_t* sl_ptr;
    const int32_t* sl_w;
    const int32_t* scol;
    const TV* val;
    int64_t nslices, nrows, per_block;
};

// coarse level operator: stencil values (st_index: slot-major for one thread
// per node, node-major for lane groups), FP32 or FP16 (scaled)
template <typename TV>
struct StOp {
    const TV* val;
    Grid g;
    int64_t nn;
};

// Row `node` of a coarse level's operator times v, one thread per node: slot
// s's block (read coalesced along nodes) times the neighbour's values, and
// the G = 32 (3D) / 16 (2D) slot products (slots past the stencil's NS and
// outside the grid are zero) summed by a fixed pairwise tree, p[i] += p[i+h]
// for h = G/2, ..., 1.  The tree is walked depth first over its leaves in
// bit-reversed order (leaf k is slot bitrev(k), and leaves 2m, 2m+1 are
// slots i, i + G/2), four leaves per loop trip, so that few products are
// live at a time; a stack holds the pending subtree sums.  Leaves of slots
// >= NS are all right children of present ones and are left out: no product
// or sum is -0 (each is a rounded fma chain from +0), so x + 0 = x and
// leaving them out changes no bit.  A slot outside the grid loads the
// node's own values (in bounds) and its product is dropped by a select.
// Offsets are 32-bit byte counts from uniform bases (the stencil and the
// vector under 4 GB each, checked when the level is built).
template <int ND>
struct Lanes {
    static constexpr int G = ND == 3 ? 32 : 16;
    static constexpr int BITS = ND == 3 ? 5 : 4;
};

template <int ND>
__device__ __forceinline__ int stencil_leaf_slot(int k) {
    return (int)(__builtin_bitreverse32((uint32_t)k) >> (32 - Lanes<ND>::BITS));
}

template <int ND, typename TV>
__device__ __forceinline__ void stencil_leaf(const char* __restrict__ val, const char* __restrict__ v, const Grid& g,
                                             uint32_t nb, uint32_t own, int node, const int (&c)[3], int sl,
                                             float (&t)[ND]) {
    constexpr int BS = ND, B2 = BS * BS;
    int dx, dy, dz;
    offset_of<ND>(sl, dx, dy, dz);
    const bool in = inside<ND>(g, c[0] + dx, c[1] + dy, c[2] + dz);
    const int m = in ? node + dx + (int)g.n0 * (dy + (int)g.n1 * dz) : node;
    const uint32_t mo = (uint32_t)m * (uint32_t)(BS * sizeof(float));
    const uint32_t so = own + (uint32_t)(sl * B2) * nb;
    float av[B2], xv[BS];
#pragma unroll
    for (int k = 0; k < B2; ++k) av[k] = (float)*(const TV*)(val + (so + (uint32_t)k * nb));
    if constexpr (BS == 3) {  // the neighbour's 3 values in one load
        const float3 w = *(const float3*)(v + mo);
        xv[0] = w.x;
        xv[1] = w.y;
        xv[2] = w.z;
    } else {
#pragma unroll
        for (int i = 0; i < BS; ++i) xv[i] = *(const float*)(v + (mo + (uint32_t)(i * sizeof(float))));
    }
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < BS; ++i) a = fmaf(av[j * BS + i], xv[i], a);
        t[j] = in ? a : 0.0f;
    }
}

template <int ND, typename TV>
__device__ __forceinline__ void stencil_row(const StOp<TV>& op, int64_t node64, const float* __restrict__ v,
                                            float (&out)[ND]) {
    constexpr int BS = ND, NS = Sten<ND>::NS, NG = Lanes<ND>::G / 4;
    const char* val = (const char*)op.val;
    const char* vb = (const char*)v;
    const uint32_t nb = (uint32_t)(op.nn * (int64_t)sizeof(TV));
    const uint32_t own = (uint32_t)(node64 * (int64_t)sizeof(TV));
    const int node = (int)node64;
    int c[3];
    c[0] = node % (int)op.g.n0;
    c[1] = (node / (int)op.g.n0) % (int)op.g.n1;
    c[2] = ND == 3 ? node / ((int)op.g.n0 * (int)op.g.n1) : 0;
    float st[3][BS];  // pending sums of 4, 8 and 16 leaves
#pragma unroll 1
    for (int q = 0; q < NG; ++q) {
        float l[4][BS];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int sl = stencil_leaf_slot<ND>(4 * q + h);
            if (sl < NS) stencil_leaf<ND, TV>(val, vb, op.g, nb, own, node, c, sl, l[h]);  // uniform
        }
        const bool r1 = stencil_leaf_slot<ND>(4 * q + 1) < NS, r3 = stencil_leaf_slot<ND>(4 * q + 3) < NS;
        float w[BS];
#pragma unroll
        for (int j = 0; j < BS; ++j) w[j] = (r1 ? l[0][j] + l[1][j] : l[0][j]) + (r3 ? l[2][j] + l[3][j] : l[2][j]);
        // combine with the pending subtrees of this one's ancestors (q's trailing ones)
        if (!(q & 1)) {
#pragma unroll
            for (int j = 0; j < BS; ++j) st[0][j] = w[j];
            continue;
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) w[j] = st[0][j] + w[j];
        if (!(q & 2)) {
#pragma unroll
            for (int j = 0; j < BS; ++j) st[1][j] = w[j];
            continue;
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) w[j] = st[1][j] + w[j];
        if (NG == 4 || !(q & 4)) {  // 2D: 16 leaves, the root
#pragma unroll
            for (int j = 0; j < BS; ++j) st[2][j] = w[j];
            continue;
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) st[2][j] = st[2][j] + w[j];  // 3D: the root
    }
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = st[2][j];
}

// The same row by a group of G lanes on a node-major stencil (small levels,
// where one thread per node leaves the device idle and its 8 rounds of loads
// exposed): lane s < NS forms slot s's product from the node's contiguous
// run, then a butterfly over the group, p[i] += p[i ^ h] for h = G/2, ...,
// 1 -- the tree above, so both forms give the same bits.  The group's first
// lane holds the sum; every lane of the group must call it.
template <int ND, typename TV>
__device__ __forceinline__ void stencil_row_group(const StOp<TV>& op, int64_t node, int sl, const float* __restrict__ v,
                                                  float (&out)[ND]) {
    constexpr int BS = ND, B2 = BS * BS, NS = Sten<ND>::NS, G = Lanes<ND>::G;
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = 0.0f;
    if (sl < NS && node < op.nn) {
        const int nd = (int)node;
        int c[3];
        c[0] = nd % (int)op.g.n0;
        c[1] = (nd / (int)op.g.n0) % (int)op.g.n1;
        c[2] = ND == 3 ? nd / ((int)op.g.n0 * (int)op.g.n1) : 0;
        int dx, dy, dz;
        offset_of<ND>(sl, dx, dy, dz);
        if (inside<ND>(op.g, c[0] + dx, c[1] + dy, c[2] + dz)) {
            const int m = nd + dx + (int)op.g.n0 * (dy + (int)op.g.n1 * dz);
            const TV* a = op.val + ((int64_t)nd * NS + sl) * B2;
            float av[B2], xv[BS];
#pragma unroll
            for (int k = 0; k < B2; ++k) av[k] = (float)a[k];
#pragma unroll
            for (int i = 0; i < BS; ++i) xv[i] = v[(int64_t)m * BS + i];
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                float t = 0.0f;
#pragma unroll
                for (int i = 0; i < BS; ++i) t = fmaf(av[j * BS + i], xv[i], t);
                out[j] = t;
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < BS; ++j) out[j] += __shfl_xor(out[j], off, G);
}

// Levels with fewer nodes than this store their stencil node-major and run
// the lane-group form (SFEM_MG_GROUP_BELOW overrides it, read when the
// hierarchy is built: both forms give the same bits).
constexpr int64_t kStencilGroupBelow = 65536;

// SFEM_MG_GALERKIN_LINES=1: the first coarse level of a matrix-free finest
// level by the one-direction products too (read when the hierarchy is
// built; the two forms agree to FP64 rounding)
bool galerkin_lines_forced() {
    const char* e = std::getenv("SFEM_MG_GALERKIN_LINES");
    return e && e[0] == '1';
}

int64_t stencil_group_below() {
    const char* e = std::getenv("SFEM_MG_GROUP_BELOW");
    return e && *e ? std::atoll(e) : kStencilGroupBelow;
}

// MODE 0: r -= A v; d_out = c1 d_in + c2 B^-1 r; x += d_out (a Chebyshev step)
// MODE 1: r -= A v (the residual after the last step)
template <int BS, int MODE>
__device__ __forceinline__ void cheb_epilogue(int64_t row, int64_t nrows, const float (&av)[BS], float* __restrict__ r,
                                              float* __restrict__ x, const float* d_in, float* __restrict__ d_out,
                                              float c1, float c2, const float* __restrict__ binv) {
    float rv[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        const int64_t d = row * BS + j;
        rv[j] = r[d] - av[j];
        r[d] = rv[j];
    }
    if (MODE == 1) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        float z = 0.0f;
#pragma unroll
        for (int i = 0; i < BS; ++i) z = fmaf(binv[(int64_t)(j * BS + i) * nrows + row], rv[i], z);
        const int64_t d = row * BS + j;
        const float dn = fmaf(c2, z, d_in ? c1 * d_in[d] : 0.0f);
        d_out[d] = dn;
        x[d] += dn;
    }
}

// The same step with its operands loaded first (cheb_load, beside the row's
// own loads, so the kernel's chain of dependent loads is one shorter), then
// cheb_finish: the arithmetic of cheb_epilogue.
template <int BS>
struct ChebIn {
    float r[BS], x[BS], din[BS], binv[BS * BS];
};

template <int BS, int MODE>
__device__ __forceinline__ void cheb_load(int64_t row, int64_t nrows, const float* __restrict__ r,
                                          const float* __restrict__ x, const float* d_in,
                                          const float* __restrict__ binv, ChebIn<BS>& o) {
#pragma unroll
    for (int j = 0; j < BS; ++j) o.r[j] = r[row * BS + j];
    if (MODE == 1) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        o.x[j] = x[row * BS + j];
        o.din[j] = d_in ? d_in[row * BS + j] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < BS * BS; ++k) o.binv[k] = binv[(int64_t)k * nrows + row];
}

template <int BS, int MODE>
__device__ __forceinline__ void cheb_finish(int64_t row, const ChebIn<BS>& o, const float (&av)[BS],
                                            float* __restrict__ r, float* __restrict__ x, bool has_din,
                                            float* __restrict__ d_out, float c1, float c2) {
    float rv[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        rv[j] = o.r[j] - av[j];
        r[row * BS + j] = rv[j];
    }
    if (MODE == 1) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        float z = 0.0f;
#pragma unroll
        for (int i = 0; i < BS; ++i) z = fmaf(o.binv[j * BS + i], rv[i], z);
        const float dn = fmaf(c2, z, has_din ? c1 * o.din[j] : 0.0f);
        d_out[row * BS + j] = dn;
        x[row * BS + j] = o.x[j] + dn;
    }
}

template <int BS, int MODE, bool NT, typename TV>
__global__ __launch_bounds__(kThreads) void cheb_sell_kernel(SellOp<TV> op, const float* v, float* __restrict__ r,
                                                             float* __restrict__ x, const float* d_in,
                                                             float* __restrict__ d_out, float c1, float c2,
                                                             const float* __restrict__ binv, const int* done) {
    if (*done) return;
    const int64_t s0 = logical_block() * op.per_block;
    const int64_t s1 = min(op.nslices, s0 + op.per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto gx = [=](int64_t c) { return v[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        float out[BS];
        sell_row<BS, TV, float, NT>(op.sl_ptr, op.sl_w, op.scol, Vals<TV>{op.val}, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < op.nrows) cheb_epilogue<BS, MODE>(row, op.nrows, out, r, x, d_in, d_out, c1, c2, binv);
    }
}

template <int ND, int MODE, bool GROUP, typename TV>
__global__ __launch_bounds__(kThreads, 4) void cheb_stencil_kernel(StOp<TV> op, const float* v, float* __restrict__ r,
                                                                   float* __restrict__ x, const float* d_in,
                                                                   float* __restrict__ d_out, float c1, float c2,
                                                                   const float* __restrict__ binv, const int* done) {
    constexpr int G = Lanes<ND>::G;
    // the flag is read beside the row's loads and checked before any store
    const bool stop = *done != 0;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    float out[ND];
    ChebIn<ND> in;
    if (GROUP) {
        const int64_t node = t / G;
        const bool lead = t % G == 0 && node < op.nn;
        if (lead) cheb_load<ND, MODE>(node, op.nn, r, x, d_in, binv, in);
        stencil_row_group<ND, TV>(op, node, (int)(t % G), v, out);
        if (lead && !stop) cheb_finish<ND, MODE>(node, in, out, r, x, d_in != nullptr, d_out, c1, c2);
    } else {
        if (t >= op.nn) return;
        cheb_load<ND, MODE>(t, op.nn, r, x, d_in, binv, in);
        stencil_row<ND, TV>(op, t, v, out);
        if (stop) return;
        cheb_finish<ND, MODE>(t, in, out, r, x, d_in != nullptr, d_out, c1, c2);
    }
}

// ------------------------------------------- matrix-free finest level --
// K_ff v = sum over cells e of s_e Ke(0) v_e, per block row (a node): the
// node's 3^ND neighbours are gathered once, each contributes to the cells it
// shares with the node, and the 2^ND cell sums are scaled by their s_e.
// kc: Ke(0) blocks (a*NEN + b) for corner codes a, b (ox + 2 oy + 4 oz).
// Inputs are zero on fixed dofs (the V-cycle and CG keep them so); a fixed
// dof of the row returns its input (the unit diagonal of the masked K_ff).
// The node rows and cell scales are read from copies with a ghost layer
// (one node / cell more on every side of the grid: row -1, scale 0), so
// every neighbour and cell is the node's padded index plus a fixed offset;
// rows of fixed nodes read the vectors' zero pad row `nrows`.
template <typename T>
struct MfOp {
    const T* kc;
    const T* es;                 // cells (e0+2) x (e1+2) [x (e2+2)], meshgrid order
    const int32_t* node_of_row;
    const int32_t* rows;         // nodes (n0+2) x (n1+2) [x (n2+2)]: block row or -1
    const uint32_t* offs;        // the same nodes: byte offset of the row (the pad row for -1) in
                                 // an FP32 vector, 3 x 4 B per row (FP64: twice that)
    const uint8_t* fixed;        // per row dof
    uint32_t n0, n1;             // nodes per direction x, y
    int64_t nrows;
};

// kc: Ke(0), read uniformly (scalar loads).  Node indices fit 31 bits (the
// row maps are int32) and the vectors 4 GB (offsets are 32-bit byte counts).
// Every load is unconditional and in bounds.  The neighbours are gathered
// and applied one plane (z offset) at a time: a third of them in registers.
// NR rows per thread share Ke(0)'s scalar loads and give the FMA chains of
// independent rows to interleave; each row's arithmetic is the same in any
// NR (the same bits).  The smoothing kernels use NR = 1: at 2 the FP32 step
// needs 223 VGPRs (2 waves per SIMD) and ran slower (solve 19.5 against
// 15.1 ms at 257 x 65 x 65 nodes).
template <int ND, int NR, typename T, typename VT>
__device__ __forceinline__ void mf_block_rows(const MfOp<T>& op, const T* kc, const int64_t (&row)[NR],
                                              const VT* __restrict__ v, T (&out)[NR][ND]) {
    constexpr int BS = ND, B2 = BS * BS, NEN = 1 << ND, NP = ND == 3 ? 3 : 1;
    const uint32_t n0 = op.n0, n1 = op.n1;
    const int P0 = (int)n0 + 2, P1 = (int)n1 + 2;  // padded nodes per direction
    const int Q0 = (int)n0 + 1, Q1 = (int)n1 + 1;  // padded cells per direction
    const char* vb = (const char*)v;
    int pn[NR];
    T se[NR][NEN];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uint32_t node = (uint32_t)op.node_of_row[row[r]];
        const uint32_t cx = node % n0, yz = node / n0;
        const uint32_t cy = ND == 3 ? yz % n1 : yz, cz = ND == 3 ? yz / n1 : 0;
        pn[r] = (int)(cx + 1) + P0 * ((int)(cy + 1) + (ND == 3 ? P1 * (int)(cz + 1) : 0));
        const int pc = (int)cx + Q0 * ((int)cy + (ND == 3 ? Q1 * (int)cz : 0));  // cell c - 1, padded
#pragma unroll
        for (int q = 0; q < NEN; ++q)  // cell c - 1 + (qx, qy, qz)
            se[r][q] = op.es[(uint32_t)(pc + (q & 1) + Q0 * (((q >> 1) & 1) + (ND == 3 ? Q1 * (q >> 2) : 0)))];
    }
    T t[NR][NEN][BS];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int q = 0; q < NEN; ++q)
#pragma unroll
            for (int i = 0; i < BS; ++i) t[r][q][i] = T(0);
    T last = T(0);  // the previous block's last sum
    // FP64: every neighbour's row first (one round trip, not one per plane);
    // FP32 per plane (hoisted, the scalar registers spill)
    constexpr bool HOIST = sizeof(T) == 8;
    // FP32: a neighbour's byte offset (the offset map); FP64: its row (the
    // row map; fewer registers across the hoisted loads)
    auto row_at = [&](int r, int pz, int s) {
        const int dx = s % 3 - 1, dy = s / 3 - 1, dz = ND == 3 ? pz - 1 : 0;
        const uint32_t p = (uint32_t)(pn[r] + dx + P0 * (dy + P1 * dz));
        if constexpr (sizeof(VT) == 8) {
            const int32_t rr = op.rows[p];
            return rr >= 0 ? (uint32_t)rr : (uint32_t)op.nrows;
        } else {
            return op.offs[p];
        }
    };
    uint32_t rn[HOIST ? NR : 1][NP][9];
    if constexpr (HOIST)
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int pz = 0; pz < NP; ++pz)
#pragma unroll
                for (int s = 0; s < 9; ++s) rn[r][pz][s] = row_at(r, pz, s);
#pragma unroll
    for (int pz = 0; pz < NP; ++pz) {
        const int sz = ND == 3 ? pz : 1;  // z offset + 1
        T vs[NR][9][BS];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int s = 0; s < 9; ++s) {
                const uint32_t q = HOIST ? rn[HOIST ? r : 0][pz][s] : row_at(r, pz, s);
                if constexpr (sizeof(VT) == 4 && BS == 3) {  // the row's 3 values in one load
                    const float3 w = *(const float3*)(vb + q);
                    vs[r][s][0] = (T)w.x;
                    vs[r][s][1] = (T)w.y;
                    vs[r][s][2] = (T)w.z;
                } else if constexpr (sizeof(VT) == 4 && BS == 2) {
                    const float2 w = *(const float2*)(vb + q);
                    vs[r][s][0] = (T)w.x;
                    vs[r][s][1] = (T)w.y;
                } else {
#pragma unroll
                    for (int j = 0; j < BS; ++j) {
                        const uint32_t o = sizeof(VT) == 8 ? (q * BS + j) * (uint32_t)sizeof(VT)
                                                           : q + j * (uint32_t)sizeof(VT);
                        vs[r][s][j] = (T)*(const VT*)(vb + o);
                    }
                }
            }
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const int sx = s % 3, sy = s / 3;
#pragma unroll
            for (int q = 0; q < NEN; ++q) {
                const int qx = q & 1, qy = (q >> 1) & 1, qz = ND == 3 ? (q >> 2) & 1 : 0;
                const int bx = sx - qx, by = sy - qy, bz = ND == 3 ? sz - qz : 0;
                if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;  // not in cell q
                const int a = (1 - qx) + 2 * (1 - qy) + (ND == 3 ? 4 * (1 - qz) : 0);
                const int b = bx + 2 * by + 4 * bz;
                // Ke(0)'s block (a, b): scalar loads (constant address space:
                // unchanged during the kernel) next to their use.  Hoisted all
                // together they exceed the scalar registers and spill; the
                // address takes a zero the compiler cannot see through, made
                // after the previous block's sums, so they stay in order.
                int zero;
                asm volatile("s_mov_b32 %0, 0" : "=s"(zero) : "v"(last));
                const __attribute__((address_space(4))) T* k =
                    (const __attribute__((address_space(4))) T*)kc + zero + (a * NEN + b) * B2;
#pragma unroll
                for (int r = 0; r < NR; ++r)
#pragma unroll
                    for (int i = 0; i < BS; ++i)
#pragma unroll
                        for (int j = 0; j < BS; ++j) t[r][q][i] = fma(k[i * BS + j], vs[r][s][j], t[r][q][i]);
                last = t[NR - 1][q][BS - 1];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < NEN; ++q) acc = fma(se[r][q], t[r][q][i], acc);
            const T own = (T)v[row[r] * BS + i];
            out[r][i] = op.fixed[row[r] * BS + i] ? own : acc;
        }
}

template <int ND, typename T, typename VT>
__device__ __forceinline__ void mf_row(const MfOp<T>& op, const T* kc, int64_t row, const VT* __restrict__ v,
                                       T (&out)[ND]) {
    const int64_t rows[1] = {row};
    T o[1][ND];
    mf_block_rows<ND, 1, T, VT>(op, kc, rows, v, o);
#pragma unroll
    for (int i = 0; i < ND; ++i) out[i] = o[0][i];
}

// ghost-layer copies for the matrix-free operator: node rows (-1 outside)
// and cell scales (0 outside); n: nodes or cells per direction, the copy
// has n + 2 (directions beyond ND: 1, no layer)
template <int ND>
__global__ void ghost_rows_kernel(const int32_t* __restrict__ in, int64_t n0, int64_t n1, int64_t n2,
                                  int32_t* __restrict__ out) {
    const int64_t p0 = n0 + 2, p1 = n1 + 2, p2 = ND == 3 ? n2 + 2 : 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p0 * p1 * p2) return;
    const int64_t x = i % p0 - 1, y = (i / p0) % p1 - 1, z = ND == 3 ? i / (p0 * p1) - 1 : 0;
    const bool in_ = x >= 0 && x < n0 && y >= 0 && y < n1 && z >= 0 && z < (ND == 3 ? n2 : 1);
    out[i] = in_ ? in[x + n0 * (y + n1 * z)] : -1;
}

// the ghost-layer row map as FP32 byte offsets (3 x 4 B per row; -1: the
// pad row nrows)
__global__ void ghost_offs_kernel(const int32_t* __restrict__ rows, int64_t n, int64_t nrows, int bs,
                                  uint32_t* __restrict__ offs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = rows[i] >= 0 ? rows[i] : nrows;
    offs[i] = (uint32_t)(r * bs * (int64_t)sizeof(float));
}

template <int ND, typename T>
__global__ void ghost_cells_kernel(const double* __restrict__ in, int64_t e0, int64_t e1, int64_t e2,
                                   T* __restrict__ out) {
    const int64_t p0 = e0 + 2, p1 = e1 + 2, p2 = ND == 3 ? e2 + 2 : 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p0 * p1 * p2) return;
    const int64_t x = i % p0 - 1, y = (i / p0) % p1 - 1, z = ND == 3 ? i / (p0 * p1) - 1 : 0;
    const bool in_ = x >= 0 && x < e0 && y >= 0 && y < e1 && z >= 0 && z < (ND == 3 ? e2 : 1);
    out[i] = in_ ? (T)in[x + e0 * (y + e1 * z)] : T(0);
}

template <int ND, int MODE>
__global__ __launch_bounds__(kThreads) void cheb_mf_kernel(MfOp<float> op, const float* v, float* __restrict__ r,
                                                           float* __restrict__ x, const float* d_in,
                                                           float* __restrict__ d_out, float c1, float c2,
                                                           const float* __restrict__ binv, const int* done) {
    if (*done) return;
    const int64_t row = logical_block() * kThreads + threadIdx.x;
    if (row >= op.nrows) return;
    float out[ND];
    mf_row<ND, float, float>(op, op.kc, row, v, out);
    cheb_epilogue<ND, MODE>(row, op.nrows, out, r, x, d_in, d_out, c1, c2, binv);
}

// First step from x = 0: r = b, d = c2 B^-1 b, x = d.  bscale: the units of
// the level's operator (the FP16 finest level works on b / max |a|).
template <int BS, typename TB>
__global__ __launch_bounds__(kThreads) void cheb_first_kernel(const TB* __restrict__ b, TB bscale, int64_t nrows,
                                                              const float* __restrict__ binv, float c2,
                                                              float* __restrict__ r, float* __restrict__ x,
                                                              float* __restrict__ d, const int* done) {
    const bool stop = *done != 0;  // checked before the stores
    const int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (row >= nrows) return;
    float rv[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) rv[j] = (float)(b[row * BS + j] * bscale);
    float bi[BS * BS];
#pragma unroll
    for (int k = 0; k < BS * BS; ++k) bi[k] = binv[(int64_t)k * nrows + row];
    if (stop) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) r[row * BS + j] = rv[j];
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        float z = 0.0f;
#pragma unroll
        for (int i = 0; i < BS; ++i) z = fmaf(bi[j * BS + i], rv[i], z);
        const float dn = c2 * z;
        d[row * BS + j] = dn;
        x[row * BS + j] = dn;
    }
}

// b_c = P' r_f on the coarse grid (0 on fixed coarse dofs); row_of_node maps
// fine grid nodes to rows of r_f (NULL: the identity); r_f is in units of
// 1 / rscale.  FIRST: also the coarse level's first smoothing step
// (cheb_first_kernel: r_c = b_c, d_c = c2 B^-1 b_c, x_c = d_c).
template <int ND, bool FIRST>
__device__ __forceinline__ void restrict_row(int64_t I, const float* __restrict__ rf, float rscale, const Grid& gf,
                                             const int32_t* __restrict__ row_of_node, const Grid& gc, int f0, int f1,
                                             int f2, const uint8_t* __restrict__ cfix, float* __restrict__ bc,
                                             const float* __restrict__ binv, float c2, float cscale,
                                             float* __restrict__ rc, float* __restrict__ xc, float* __restrict__ dc,
                                             bool skip) {
    constexpr int BS = ND;
    int64_t ci[3];
    coords_of<ND>(I, gc, ci);
    // the interpolation support of I per direction (child(); the identity
    // along directions that are not coarsened)
    const int f[3] = {f0, f1, f2};
    const int64_t ef[3] = {gf.n0 - 1, gf.n1 - 1, gf.n2 - 1};
    int64_t fi[3][3] = {};  // unused entries: node 0 (their terms are not added)
    float fw[3][3] = {};
    int nk[3];
    for (int d = 0; d < 3; ++d) {
        nk[d] = 0;
        if (f[d] != 2) {
            fi[d][0] = ci[d];
            fw[d][0] = 1.0f;
            nk[d] = 1;
            continue;
        }
        for (int k = 0; k < 3; ++k) {
            int64_t i;
            double w;
            if (child(ci[d], k, ef[d], i, w)) {
                fi[d][nk[d]] = i;
                fw[d][nk[d]] = (float)w;
                ++nk[d];
            }
        }
    }
    float acc[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) acc[j] = 0.0f;
    // a fixed 3x3x3 loop, so all loads can be in flight at once: the missing
    // support points (and fully fixed fine nodes) load the first row and
    // their term is then dropped by a select, so the sums are exactly those
    // of the support points alone, whatever that row holds (Inf / NaN too)
#pragma unroll
    for (int kz = 0; kz < 3; ++kz)
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const bool in = kx < nk[0] && ky < nk[1] && kz < nk[2];
                const int64_t fn = node_at(gf, fi[0][kx], fi[1][ky], fi[2][kz]);
                const int64_t mrow = row_of_node ? (int64_t)row_of_node[fn] : fn;
                const bool use = in && mrow >= 0;
                const int64_t row = use ? mrow : 0;
                const float w = fw[0][kx] * fw[1][ky] * fw[2][kz];
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const float t = fmaf(w, rf[row * BS + j] * rscale, acc[j]);
                    acc[j] = use ? t : acc[j];
                }
            }
    const int64_t nnc = gc.n0 * gc.n1 * gc.n2;
#pragma unroll
    for (int j = 0; j < BS; ++j) acc[j] = cfix[I * BS + j] ? 0.0f : acc[j];
    if (skip) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        bc[I * BS + j] = acc[j];
        if (FIRST) {
            acc[j] *= cscale;  // the coarse level's units
            rc[I * BS + j] = acc[j];
        }
    }
    if (FIRST) {
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            float z = 0.0f;
#pragma unroll
            for (int i = 0; i < BS; ++i) z = fmaf(binv[(int64_t)(j * BS + i) * nnc + I], acc[i], z);
            const float dn = c2 * z;
            dc[I * BS + j] = dn;
            xc[I * BS + j] = dn;
        }
    }
}

template <int ND, bool FIRST>
__global__ __launch_bounds__(kThreads) void restrict_kernel(const float* __restrict__ rf, float rscale, Grid gf,
                                                            const int32_t* __restrict__ row_of_node, Grid gc, int f0,
                                                            int f1, int f2, const uint8_t* __restrict__ cfix,
                                                            float* __restrict__ bc, const float* __restrict__ binv,
                                                            float c2, float cscale, float* __restrict__ rc,
                                                            float* __restrict__ xc, float* __restrict__ dc,
                                                            const int* done) {
    const bool stop = *done != 0;  // checked before the stores
    const int64_t I = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (I >= gc.n0 * gc.n1 * gc.n2) return;
    restrict_row<ND, FIRST>(I, rf, rscale, gf, row_of_node, gc, f0, f1, f2, cfix, bc, binv, c2, cscale, rc, xc, dc,
                            stop);
}

// e = P x_c on the fine rows (0 on fixed fine dofs), x_f += e; node_of_row
// maps rows to fine grid nodes (NULL: the identity).
template <int ND>
__device__ __forceinline__ void prolong_row(int64_t row, const float* __restrict__ xc, const Grid& gc, const Grid& gf,
                                            int f0, int f1, int f2, const int32_t* __restrict__ node_of_row,
                                            const uint8_t* __restrict__ ffix, float* __restrict__ xf,
                                            float* __restrict__ e, bool skip) {
    constexpr int BS = ND;
    const int64_t node = node_of_row ? (int64_t)node_of_row[row] : row;
    int64_t cf[3];
    coords_of<ND>(node, gf, cf);
    const int f[3] = {f0, f1, f2};
    int64_t p[3][2];
    float w[3][2];
    int np[3];
    const int64_t ef[3] = {gf.n0 - 1, gf.n1 - 1, gf.n2 - 1};
    for (int d = 0; d < 3; ++d) {
        if (f[d] == 2 && (ef[d] & 1) && cf[d] == ef[d]) {  // odd e: the last node has one parent
            p[d][0] = p[d][1] = (ef[d] + 1) / 2;
            w[d][0] = w[d][1] = 1.0f;
            np[d] = 1;
        } else if (f[d] == 2 && (cf[d] & 1)) {
            p[d][0] = (cf[d] - 1) / 2;
            p[d][1] = p[d][0] + 1;
            w[d][0] = w[d][1] = 0.5f;
            np[d] = 2;
        } else {
            p[d][0] = p[d][1] = f[d] == 2 ? cf[d] / 2 : cf[d];
            w[d][0] = w[d][1] = 1.0f;
            np[d] = 1;
        }
    }
    float acc[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) acc[j] = 0.0f;
    // a fixed 2x2x2 loop, all loads in flight: a second parent that does not
    // exist loads the first one again and its term is dropped by a select
    // (the sums are those of the real parents, whatever the loaded value)
#pragma unroll
    for (int kz = 0; kz < 2; ++kz)
#pragma unroll
        for (int ky = 0; ky < 2; ++ky)
#pragma unroll
            for (int kx = 0; kx < 2; ++kx) {
                const int64_t I = node_at(gc, p[0][kx], p[1][ky], p[2][kz]);
                const bool in = kx < np[0] && ky < np[1] && kz < np[2];
                const float ww = w[0][kx] * w[1][ky] * w[2][kz];
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const float t = fmaf(ww, xc[I * BS + j], acc[j]);
                    acc[j] = in ? t : acc[j];
                }
            }
    if (skip) return;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        const int64_t d = row * BS + j;
        const float ev = ffix[d] ? 0.0f : acc[j];
        e[d] = ev;
        xf[d] += ev;
    }
}

template <int ND>
__global__ __launch_bounds__(kThreads) void prolong_kernel(const float* __restrict__ xc, Grid gc, Grid gf, int f0,
                                                           int f1, int f2, const int32_t* __restrict__ node_of_row,
                                                           int64_t nrows, const uint8_t* __restrict__ ffix,
                                                           float* __restrict__ xf, float* __restrict__ e,
                                                           const int* done) {
    const bool stop = *done != 0;  // checked before the stores
    const int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (row >= nrows) return;
    prolong_row<ND>(row, xc, gc, gf, f0, f1, f2, node_of_row, ffix, xf, e, stop);
}

// coarsest level: x = A^-1 b with the dense FP64 inverse (one block)
__global__ __launch_bounds__(kThreads) void dense_apply_kernel(const double* __restrict__ inv, int64_t nc,
                                                               const float* __restrict__ b, float* __restrict__ x,
                                                               const int* done) {
    const bool stop = *done != 0;  // checked before the stores
    for (int64_t i = threadIdx.x; i < nc; i += kThreads) {
        double s = 0.0;
        for (int64_t j = 0; j < nc; ++j) s = fma(inv[i * nc + j], (double)b[j], s);
        if (!stop) x[i] = (float)s;
    }
}

// ------------------------------------------- power iteration (setup) --
struct PowState {
    double lam;  // |B^-1 A v| for the last unit vector v
    double inv;  // 1 / |w| of the last step (scale of the next input)
    Arrivals arrive;
};

__device__ __forceinline__ double sum_parts(const double* partials, int nparts, double* sh) {
    double s[1];
    thread_partials<1>(partials, nparts, s);
    return block_sum_256(s[0], sh);
}

__device__ __forceinline__ void pow_finish(PowState* ps, const double* partials, double* sh) {
    const double s = sum_parts(partials, gridDim.x, sh);
    if (threadIdx.x == 0) {
        const double nrm = sqrt(s);
        ps->lam = nrm;
        ps->inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    }
}

__device__ __forceinline__ float hash_unit(int64_t i) {
    uint32_t h = (uint32_t)i * 2654435761u ^ (uint32_t)(i >> 32);
    h ^= h >> 15;
    h *= 0x2c1b3c6dU;
    h ^= h >> 12;
    h *= 0x297a2d39U;
    h ^= h >> 15;
    return (float)(h & 0xFFFF) * (1.0f / 32768.0f) - 1.0f;
}

// start vector: fixed pseudo-random values, 0 on fixed dofs
__global__ __launch_bounds__(kThreads) void pow_init_kernel(int64_t n, const uint8_t* __restrict__ fixed,
                                                            float* __restrict__ v, PowState* ps,
                                                            double* __restrict__ partials) {
    __shared__ double sh[8];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float x = fixed[i] ? 0.0f : hash_unit(i);
        v[i] = x;
        s += (double)x * (double)x;
    }
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&ps->arrive)) pow_finish(ps, partials, sh);
}

template <int BS>
__device__ __forceinline__ double pow_epilogue(int64_t row, int64_t nrows, const float (&av)[BS], float scale,
                                               const float* __restrict__ binv, float* __restrict__ w) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        float z = 0.0f;
#pragma unroll
        for (int i = 0; i < BS; ++i) z = fmaf(binv[(int64_t)(j * BS + i) * nrows + row], av[i], z);
        z *= scale;
        w[row * BS + j] = z;
        s += (double)z * (double)z;
    }
    return s;
}

// w = B^-1 A (v / |v|), partial |w|^2
template <int BS, typename TV>
__global__ __launch_bounds__(kThreads) void pow_sell_kernel(SellOp<TV> op, const float* __restrict__ v,
                                                            const float* __restrict__ binv, float* __restrict__ w,
                                                            PowState* ps, double* __restrict__ partials) {
    __shared__ double sh[8];
    const float scale = (float)ps->inv;
    const int64_t s0 = logical_block() * op.per_block;
    const int64_t s1 = min(op.nslices, s0 + op.per_block);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto gx = [=](int64_t c) { return v[c]; };
    double s = 0.0;
    for (int64_t sl = s0 + wv; sl < s1; sl += kThreads / 64) {
        float out[BS];
        sell_row<BS, TV, float, false>(op.sl_ptr, op.sl_w, op.scol, Vals<TV>{op.val}, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < op.nrows) s += pow_epilogue<BS>(row, op.nrows, out, scale, binv, w);
    }
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&ps->arrive)) pow_finish(ps, partials, sh);
}

template <int ND, bool GROUP, typename TV>
__global__ __launch_bounds__(kThreads, 4) void pow_stencil_kernel(StOp<TV> op, const float* __restrict__ v,
                                                                  const float* __restrict__ binv,
                                                                  float* __restrict__ w, PowState* ps,
                                                                  double* __restrict__ partials) {
    constexpr int G = GROUP ? Lanes<ND>::G : 1;
    __shared__ double sh[8];
    const float scale = (float)ps->inv;
    double sum = 0.0;
    // every lane of a group runs the same trip count: the stride and the
    // bound are multiples of G
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < op.nn * G;
         t += (int64_t)gridDim.x * kThreads) {
        const int64_t node = t / G;
        float out[ND];
        if (GROUP) stencil_row_group<ND, TV>(op, node, (int)(t % G), v, out);
        else stencil_row<ND, TV>(op, node, v, out);
        if (t % G == 0) sum += pow_epilogue<ND>(node, op.nn, out, scale, binv, w);
    }
    const double b = block_sum_256(sum, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&ps->arrive)) pow_finish(ps, partials, sh);
}

template <int ND>
__global__ __launch_bounds__(kThreads) void pow_mf_kernel(MfOp<float> op, const float* __restrict__ v,
                                                          const float* __restrict__ binv, float* __restrict__ w,
                                                          PowState* ps, double* __restrict__ partials) {
    __shared__ double sh[8];
    const float scale = (float)ps->inv;
    double sum = 0.0;
    for (int64_t row = logical_block() * kThreads + threadIdx.x; row < op.nrows;
         row += (int64_t)gridDim.x * kThreads) {
        float out[ND];
        mf_row<ND, float, float>(op, op.kc, row, v, out);
        sum += pow_epilogue<ND>(row, op.nrows, out, scale, binv, w);
    }
    const double b = block_sum_256(sum, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&ps->arrive)) pow_finish(ps, partials, sh);
}

// ------------------------------------------------------- outer CG ----
struct MgState {
    double bb, tol2bb;
    double rz, rz_old, alpha, beta;
    double rr, rr_true, best_true, floor_rr;
    long long iters, trues;
    int done;       // 1 converged, 2 breakdown, 3 stagnation
    int restart;    // next direction is z (after a residual replacement)
    int need_true;  // the recurrence claims convergence: check the FP64 residual
    int stall;
    Arrivals arrive;
};

__global__ __launch_bounds__(kThreads) void mg_norm_kernel(const double* __restrict__ rhs, int64_t n,
                                                           int64_t per_block, MgState* st,
                                                           double* __restrict__ partials, double rtol) {
    __shared__ double sh[8];
    const int64_t i0 = (int64_t)blockIdx.x * per_block, i1 = min(n, i0 + per_block);
    double s = 0.0;
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s = fma(rhs[i], rhs[i], s);
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&st->arrive)) {
        const double bb = sum_parts(partials, gridDim.x, sh);
        if (threadIdx.x == 0) {
            st->bb = bb;
            st->tol2bb = rtol * rtol * bb;
            if (bb == 0.0) {
                st->done = 1;
                st->rr_true = 0.0;
            }
        }
    }
}

__device__ void true_finish(MgState* st, double* __restrict__ partials, double* sh, double rr, double ff);

// r = rhs - K_ff x in FP64 from K's block columns (K is exactly symmetric:
// column c's blocks are row c's, transposed), ||r||^2 and the scale of the
// rounding floor; fixed dofs of a row keep the unit diagonal.  32 lanes per
// block row, one block per lane (adjacent blocks: coalesced), summed by a
// fixed butterfly.
template <int BS>
__global__ __launch_bounds__(kThreads) void mg_true_k_kernel(const int64_t* __restrict__ bcolptr,
                                                             const int32_t* __restrict__ browidx,
                                                             const double* __restrict__ bval,
                                                             const int32_t* __restrict__ node_of_row,
                                                             const int32_t* __restrict__ row_of_node,
                                                             const int32_t* __restrict__ free_of, int64_t nrows,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ rhs, double* __restrict__ r,
                                                             MgState* st, double* __restrict__ partials) {
    constexpr int B2 = BS * BS, G = 32;
    __shared__ double sh[8];
    if (st->done || !st->need_true) return;
    double rr = 0.0, ff = 0.0;
    const int lane = threadIdx.x % G;
    // every lane of a group runs the same trip count: stride and bound are
    // multiples of G
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < nrows * G;
         t += (int64_t)gridDim.x * kThreads) {
        const int64_t row = t / G;
        const int64_t cu = node_of_row[row];
        double out[BS], aout[BS];
#pragma unroll
        for (int j = 0; j < BS; ++j) out[j] = aout[j] = 0.0;
        for (int64_t u = bcolptr[cu] + lane; u < bcolptr[cu + 1]; u += G) {
            const int64_t ru = browidx[u];
            const int32_t c = row_of_node[ru];
            if (c < 0) continue;
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                if (free_of[ru * BS + i] < 0) continue;  // fixed column
                const double xv = x[(int64_t)c * BS + i];
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const double p = bval[u * B2 + i * BS + j] * xv;
                    out[j] += p;
                    aout[j] += fabs(p);
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                out[j] += __shfl_xor(out[j], off, G);
                aout[j] += __shfl_xor(aout[j], off, G);
            }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                if (free_of[cu * BS + j] < 0) {  // unit diagonal
                    out[j] = x[d];
                    aout[j] = fabs(x[d]);
                }
                const double res = rhs[d] - out[j];
                r[d] = res;
                rr = fma(res, res, rr);
                const double sc = fabs(rhs[d]) + aout[j];
                ff = fma(sc, sc, ff);
            }
        }
    }
    true_finish(st, partials, sh, rr, ff);
}

// r = rhs - A x in FP64 when a check is due (else exits at once); decides
// convergence, or replaces the residual and restarts the directions.
template <int BS>
__global__ __launch_bounds__(kThreads) void mg_true_kernel(const int64_t* __restrict__ sl_ptr,
                                                           const int32_t* __restrict__ sl_w,
                                                           const int32_t* __restrict__ scol,
                                                           const double* __restrict__ sval, int64_t nslices,
                                                           int64_t nrows, int64_t per_block,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ rhs, double* __restrict__ r,
                                                           MgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done || !st->need_true) return;
    const int64_t s0 = logical_block() * per_block;
    const int64_t s1 = min(nslices, s0 + per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto gx = [=](int64_t c) { return x[c]; };
    double rr = 0.0, ff = 0.0;
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        double out[BS], aout[BS];
        sell_row_abs<BS>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out, aout);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                const double res = rhs[d] - out[j];
                r[d] = res;
                rr = fma(res, res, rr);
                const double sc = fabs(rhs[d]) + aout[j];
                ff = fma(sc, sc, ff);
            }
        }
    }
    true_finish(st, partials, sh, rr, ff);
}

// Block partials of ||r||^2 and the floor scale; the last block to arrive
// forms the totals and decides: converged, breakdown, stagnation or restart.
__device__ void true_finish(MgState* st, double* __restrict__ partials, double* sh, double rr, double ff) {
    const double b0 = block_sum_256(rr, sh);
    const double b1 = block_sum_256(ff, sh + 4);
    if (threadIdx.x == 0) {
        st_coherent(partials + 2 * blockIdx.x, b0);
        st_coherent(partials + 2 * blockIdx.x + 1, b1);
    }
    if (last_block_arrive(&st->arrive)) {
        double sv[2];
        thread_partials<2>(partials, (int)gridDim.x, sv);
        const double trr = block_sum_256(sv[0], sh);
        const double tff = block_sum_256(sv[1], sh + 4);
        if (threadIdx.x == 0) {
            st->rr_true = trr;
            st->rr = trr;
            st->floor_rr = 16.0 * kEps64 * kEps64 * tff;
            st->trues += 1;
            st->need_true = 0;
            if (trr <= st->tol2bb) {
                st->done = 1;
            } else if (!(trr == trr) || trr > 1e300) {
                st->done = 2;
            } else {
                st->restart = 1;
                if (trr <= 0.81 * st->best_true) {
                    st->best_true = trr;
                    st->stall = 0;
                } else {
                    // no 10% gain since the best check: at the rounding floor,
                    // or for the third time in a row, the tolerance is out of reach
                    st->stall += 1;
                    if (trr <= st->floor_rr || st->stall >= 3) st->done = 3;
                }
            }
        }
    }
}

// r.z -> beta (0 after a restart)
// block partials of r.z; the last block to arrive forms rz and beta
__device__ void rz_finish(MgState* st, double* __restrict__ partials, double* sh, double s) {
    const double b = block_sum_256(s, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&st->arrive)) {
        const double rz = sum_parts(partials, gridDim.x, sh);
        if (threadIdx.x == 0) {
            if (!(rz > 0.0) || !isfinite(rz)) {
                st->done = 2;
            } else {
                st->beta = st->restart ? 0.0 : rz / st->rz_old;
                st->rz_old = rz;
            }
            st->rz = rz;
            st->restart = 0;
        }
    }
}

__global__ __launch_bounds__(kThreads) void mg_rz_kernel(const double* __restrict__ r, const float* __restrict__ z,
                                                         int64_t n, int64_t per_block, MgState* st,
                                                         double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const int64_t i0 = (int64_t)blockIdx.x * per_block, i1 = min(n, i0 + per_block);
    double s = 0.0;
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s = fma(r[i], (double)z[i], s);
    rz_finish(st, partials, sh, s);
}

// ------------------------------- matrix-free finest level on MFMA (3D) --
// K_ff v as sum over cells of s_e Ke(0) v_e, formed cell by cell on the
// matrix cores.  A workgroup owns a tile of kTX x kTY x kTZ grid nodes:
//   A  the values of the tile's nodes and of a one-node halo go to LDS, scaled
//      by 2^-ex (ex from the tile's largest |v|) and split into FP16 hi + lo
//      (hi = fp16(y), lo = fp16(y - hi): 22 significant bits together);
//   B  every cell touching the tile, (kTX+1)(kTY+1)(kTZ+1) of them, gets
//      Z_e = Ke(0) v_e from v_mfma_f32_32x32x8f16 (32 cells per instruction
//      column block, Ke(0)'s 24 rows padded to 32, K = 24 in 3 steps), with
//      Ke(0) split the same way and the FP32 sum hi*hi + hi*lo + lo*hi; the
//      cell's column is scaled by s_e 2^ex 2^ek into LDS;
//   C  each tile node sums its 8 cells' rows in a fixed order, then the
//      Chebyshev epilogue of the VALU kernel (cheb_mf_kernel).
// Halo cells are formed by both tiles that touch them (1.66x the cell
// products); the matrix cores leave the kernel bound by its memory stream.
// The products carry about FP32 accuracy; the V-cycle stays a fixed linear
// operator.  Lane layout of the 32x32x8 instruction: A[i][k] in lane
// i + 32 (k / 4), element k % 4; B[k][j] in lane j + 32 (k / 4), element
// k % 4; D[i][j] in lane j + 32 ((i / 4) % 2), element 4 (i / 8) + i % 4.
constexpr int kTX = 16, kTY = 4, kTZ = 4;                      // tile nodes
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;     // with the halo
constexpr int kCX = kTX + 1, kCY = kTY + 1, kCZ = kTZ + 1;     // cells touching the tile
constexpr int kHN = kHX * kHY * kHZ;
constexpr int kCN = kCX * kCY * kCZ;
constexpr int kCG = (kCN + 31) / 32;                           // column blocks of 32 cells
constexpr int kZS = 24;                                        // floats per cell in LDS

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

struct MfTiles {
    int64_t n0, n1, n2;  // grid nodes per direction
    int64_t tx, ty, tz;  // tiles per direction
};

template <int MODE>
__global__ __launch_bounds__(kThreads) void cheb_mf_mfma_kernel(
    MfOp<float> op, MfTiles T, const half4_t* __restrict__ kfrag, float kscale, const float* v,
    float* __restrict__ r, float* __restrict__ x, const float* d_in, float* __restrict__ d_out, float c1, float c2,
    const float* __restrict__ binv, const int* done) {
    __shared__ _Float16 xh[kHN * 3], xl[kHN * 3];
    __shared__ float zc[kCN * kZS];
    __shared__ float scell[kCN];
    __shared__ float wmax[kThreads / 64];
    if (*done != 0) return;  // uniform over the grid
    const int tid = threadIdx.x;
    const int64_t t = logical_block();
    const bool live = t < T.tx * T.ty * T.tz;
    int64_t x0 = 0, y0 = 0, z0 = 0;
    if (live) {
        x0 = (t % T.tx) * kTX;
        y0 = (t / T.tx % T.ty) * kTY;
        z0 = t / (T.tx * T.ty) * kTZ;
    }
    const int64_t P0 = T.n0 + 2, P1 = T.n1 + 2, Q0 = T.n0 + 1, Q1 = T.n1 + 1;
    // this thread's tile node: its row and everything the epilogue reads,
    // loaded first so that their latency overlaps the cell products
    const int tx = tid % kTX, ty = tid / kTX % kTY, tz = tid / (kTX * kTY);
    const int64_t gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
    int32_t row = -1;
    if (live && gx < T.n0 && gy < T.n1 && gz < T.n2) row = op.rows[(gx + 1) + P0 * ((gy + 1) + P1 * (gz + 1))];
    float rv[3] = {0, 0, 0}, xv[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, own[3] = {0, 0, 0}, bv[9] = {};
    uint8_t fx[3] = {0, 0, 0};
    if (row >= 0) {
        const int64_t r3 = (int64_t)row * 3;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            rv[j] = r[r3 + j];
            own[j] = v[r3 + j];
            fx[j] = op.fixed[r3 + j];
            if (MODE == 0) {
                xv[j] = x[r3 + j];
                dv[j] = d_in ? d_in[r3 + j] : 0.0f;
            }
        }
        if (MODE == 0)
#pragma unroll
            for (int m = 0; m < 9; ++m) bv[m] = binv[(int64_t)m * op.nrows + row];
    }
    // A: halo values (zero outside the grid and on fixed nodes), staged as
    // FP32 in zc, and the tile's largest magnitude; the cell scales
    float amax = 0.0f;
    for (int i = tid; i < kHN; i += kThreads) {
        const int64_t hx = x0 - 1 + i % kHX, hy = y0 - 1 + i / kHX % kHY, hz = z0 - 1 + i / (kHX * kHY);
        int32_t hr = -1;
        if (live && hx <= T.n0 && hy <= T.n1 && hz <= T.n2) hr = op.rows[(hx + 1) + P0 * ((hy + 1) + P1 * (hz + 1))];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float val = hr >= 0 ? v[(int64_t)hr * 3 + j] : 0.0f;
            zc[i * 3 + j] = val;
            amax = fmaxf(amax, fabsf(val));
        }
    }
    float se[2] = {0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = tid + k * kThreads;
        const int64_t cx = x0 - 1 + c % kCX, cy = y0 - 1 + c / kCX % kCY, cz = z0 - 1 + c / (kCX * kCY);
        if (c < kCN && live && cx <= T.n0 - 1 && cy <= T.n1 - 1 && cz <= T.n2 - 1)
            se[k] = op.es[(cx + 1) + Q0 * ((cy + 1) + Q1 * (cz + 1))];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, 64));
    if ((tid & 63) == 0) wmax[tid >> 6] = amax;
    __syncthreads();
    amax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    int ex = 0;
    (void)frexpf(amax, &ex);  // amax 2^-ex in [1/2, 1); 0 for a zero tile
    const float xs = ldexpf(1.0f, -ex);
    for (int i = tid; i < kHN * 3; i += kThreads) {
        const float y = zc[i] * xs;
        const _Float16 h = (_Float16)y;
        xh[i] = h;
        xl[i] = (_Float16)(y - (float)h);
    }
    const float unscale = ldexpf(kscale, ex);
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (tid + k * kThreads < kCN) scell[tid + k * kThreads] = se[k] * unscale;
    __syncthreads();
    // B: Z_e = s_e Ke(0) v_e on the matrix cores, 32 cells per column block
    const int w = tid >> 6, lane = tid & 63, half = lane >> 5;
    half4_t ahi[3], alo[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        ahi[kk] = kfrag[kk * 64 + lane];
        alo[kk] = kfrag[(3 + kk) * 64 + lane];
    }
    for (int g = w; g < kCG; g += kThreads / 64) {
        const int cell = g * 32 + (lane & 31);
        const bool cv = cell < kCN;
        const int cx = cell % kCX, cy = cell / kCX % kCY, cz = cell / (kCX * kCY);
        floatx16_t acc = {};
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            half4_t bh, bl;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = 8 * kk + 4 * half + e;  // dof of the cell: corner k / 3, component k % 3
                const int b = k / 3, j = k % 3;
                const int hn = (cx + (b & 1)) + kHX * ((cy + ((b >> 1) & 1)) + kHY * (cz + (b >> 2)));
                bh[e] = cv ? xh[hn * 3 + j] : (_Float16)0.0f;
                bl[e] = cv ? xl[hn * 3 + j] : (_Float16)0.0f;
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(ahi[kk], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(ahi[kk], bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(alo[kk], bh, acc, 0, 0, 0);
        }
        if (cv) {
            const float sc = scell[cell];
#pragma unroll
            for (int e = 0; e < 12; ++e) {  // rows 24..31 are the padding
                const int i = 8 * (e >> 2) + 4 * half + (e & 3);
                zc[cell * kZS + i] = acc[e] * sc;
            }
        }
    }
    __syncthreads();
    // C: a tile node sums its 8 cells (ascending corner code), then the
    // Chebyshev epilogue on the values loaded at the start
    double s = 0.0;
    if (row >= 0) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
            const int cell = (tx + qx) + kCX * ((ty + qy) + kCY * (tz + qz));
            const int a = (1 - qx) + 2 * (1 - qy) + 4 * (1 - qz);
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[i] += zc[cell * kZS + 3 * a + i];
        }
        const int64_t r3 = (int64_t)row * 3;
        float res[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            res[j] = rv[j] - (fx[j] ? own[j] : acc[j]);
            r[r3 + j] = res[j];
        }
        if (MODE == 0)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float z = 0.0f;
#pragma unroll
                for (int i = 0; i < 3; ++i) z = fmaf(bv[j * 3 + i], res[i], z);
                const float dn = fmaf(c2, z, d_in ? c1 * dv[j] : 0.0f);
                d_out[r3 + j] = dn;
                x[r3 + j] = xv[j] + dn;
            }
    }
}

__global__ __launch_bounds__(kThreads) void mg_p_kernel(double* __restrict__ p, const float* __restrict__ z, int64_t n,
                                                        int64_t per_block, const MgState* st) {
    if (st->done) return;
    const double beta = st->beta;
    const int64_t i0 = (int64_t)blockIdx.x * per_block, i1 = min(n, i0 + per_block);
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) p[i] = fma(beta, p[i], (double)z[i]);
}

// q = A p (operator values TV, FP64 vectors), p.q -> alpha
template <int BS, typename TV, bool NT>
__global__ __launch_bounds__(kThreads) void mg_spmv_kernel(const int64_t* __restrict__ sl_ptr,
                                                           const int32_t* __restrict__ sl_w,
                                                           const int32_t* __restrict__ scol, Vals<TV> sval,
                                                           int64_t nslices, int64_t nrows, int64_t per_block,
                                                           const double* __restrict__ p, double* __restrict__ q,
                                                           MgState* st, double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const int64_t s0 = logical_block() * per_block;
    const int64_t s1 = min(nslices, s0 + per_block);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double dot = 0.0;
    auto gx = [=](int64_t c) { return p[c]; };
    for (int64_t sl = s0 + w; sl < s1; sl += kThreads / 64) {
        double out[BS];
        sell_row<BS, TV, double, NT>(sl_ptr, sl_w, scol, sval, sl, lane, gx, out);
        const int64_t row = sl * 64 + lane;
        if (row < nrows) {
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int64_t d = row * BS + j;
                q[d] = out[j];
                dot = fma(p[d], out[j], dot);
            }
        }
    }
    const double b = block_sum_256(dot, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&st->arrive)) {
        const double pq = sum_parts(partials, gridDim.x, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz_old / pq;
            if (!(pq > 0.0) || !isfinite(st->alpha)) st->done = 2;
        }
    }
}

// q = K_ff p matrix-free in FP64, p.q -> alpha (4 waves per SIMD: <= 128
// registers)
template <int ND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void mg_spmv_mf_kernel(MfOp<double> op, const double* __restrict__ p,
                                                              double* __restrict__ q, MgState* st,
                                                              double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    double dot = 0.0;
    for (int64_t row = logical_block() * kThreads + threadIdx.x; row < op.nrows;
         row += (int64_t)gridDim.x * kThreads) {
        double out[ND];
        mf_row<ND, double, double>(op, op.kc, row, p, out);
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            const int64_t d = row * ND + j;
            q[d] = out[j];
            dot = fma(p[d], out[j], dot);
        }
    }
    const double b = block_sum_256(dot, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&st->arrive)) {
        const double pq = sum_parts(partials, gridDim.x, sh);
        if (threadIdx.x == 0) {
            st->alpha = st->rz_old / pq;
            if (!(pq > 0.0) || !isfinite(st->alpha)) st->done = 2;
        }
    }
}

// x += alpha p, r -= alpha q, r.r; a recurrence at the tolerance asks for
// the FP64 check
__global__ __launch_bounds__(kThreads) void mg_update_kernel(double* __restrict__ x, double* __restrict__ r,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ q, int64_t n,
                                                             int64_t per_block, MgState* st,
                                                             double* __restrict__ partials) {
    __shared__ double sh[8];
    if (st->done) return;
    const double a = st->alpha;
    const int64_t i0 = (int64_t)blockIdx.x * per_block, i1 = min(n, i0 + per_block);
    double rr = 0.0;
#pragma unroll 4  // four strides of loads in flight; the same order of operations
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        x[i] = fma(a, p[i], x[i]);
        const double ri = fma(-a, q[i], r[i]);
        r[i] = ri;
        rr = fma(ri, ri, rr);
    }
    const double b = block_sum_256(rr, sh);
    if (threadIdx.x == 0) st_coherent(partials + blockIdx.x, b);
    if (last_block_arrive(&st->arrive)) {
        const double trr = sum_parts(partials, gridDim.x, sh);
        if (threadIdx.x == 0) {
            st->rr = trr;
            st->iters += 1;
            if (trr <= st->tol2bb) st->need_true = 1;
            if (!isfinite(trr)) st->done = 2;
        }
    }
}

// ---------------------------------------------------------- host ----
struct Launch {
    unsigned grid;
    int64_t per_block;
};

Launch plan_vec(int64_t items, unsigned cap) {
    int64_t need = (items + kThreads - 1) / kThreads;
    int64_t grid = std::min<int64_t>(need, cap);
    grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
    int64_t per = (items + grid - 1) / grid;
    per = (per + kThreads - 1) / kThreads * kThreads;
    if (per == 0) per = kThreads;
    return {(unsigned)grid, per};
}

// one slice per wavefront, four per block
Launch plan_sell(int64_t nslices) {
    const int64_t per = kThreads / 64;
    int64_t grid = (nslices + per - 1) / per;
    grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
    return {(unsigned)grid, per};
}

unsigned rows_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }
// The same rounded up to a multiple of the 8 XCDs, for kernels that take
// their rows from logical_block(): each XCD then works a contiguous range of
// rows (a z-slab of the grid), and the neighbour rows it gathers are mostly
// in its own L2.  Blocks past the last row return.
unsigned xcd_rows_grid(int64_t n) { return (rows_grid(n) + 7u) / 8u * 8u; }

Grid grid_of(const MgLevel& l) { return Grid{l.n[0], l.n[1], l.n[2]}; }

// Element-matrix deviations (K's ke_dev) up to which the finest level is
// applied from the element scales: the smoother computes in FP32; the outer
// FP64 operator then differs from the assembled K by the rounding of the
// elements' coordinates (the stop still tests the assembled K's residual).
constexpr double kMfSmootherDev = 1e-6;
constexpr double kMfOuterDev = 1e-12;

template <typename T>
MfOp<T> mf_op(const sfem_system* S, const MgHierarchy& H, const T* kc, const T* es) {
    return {kc, es, H.node_of_row.get(), H.mf_rows.get(), H.mf_offs.get(), H.L[0].fixed.get(),
            (uint32_t)H.L[0].n[0], (uint32_t)H.L[0].n[1], S->nfb};
}

MfOp<float> mf_op32(const sfem_system* S, const MgHierarchy& H) {
    return mf_op<float>(S, H, H.kc32.get(), H.es32.get());
}

MfTiles mf_tiles(const MgHierarchy& H) {
    const MgLevel& L0 = H.L[0];
    return {L0.n[0], L0.n[1], L0.n[2], (L0.n[0] + kTX - 1) / kTX, (L0.n[1] + kTY - 1) / kTY,
            (L0.n[2] + kTZ - 1) / kTZ};
}

// Gauss-Jordan inverse with partial pivoting (FP64); false if singular.
bool dense_inverse(std::vector<double>& a, int64_t n, std::vector<double>& inv) {
    inv.assign((size_t)(n * n), 0.0);
    for (int64_t i = 0; i < n; ++i) inv[(size_t)(i * n + i)] = 1.0;
    for (int64_t c = 0; c < n; ++c) {
        int64_t piv = c;
        for (int64_t i = c + 1; i < n; ++i)
            if (std::fabs(a[(size_t)(i * n + c)]) > std::fabs(a[(size_t)(piv * n + c)])) piv = i;
        const double pv = a[(size_t)(piv * n + c)];
        if (!(std::fabs(pv) > 0.0) || !std::isfinite(pv)) return false;
        if (piv != c)
            for (int64_t j = 0; j < n; ++j) {
                std::swap(a[(size_t)(piv * n + j)], a[(size_t)(c * n + j)]);
                std::swap(inv[(size_t)(piv * n + j)], inv[(size_t)(c * n + j)]);
            }
        const double rp = 1.0 / pv;
        for (int64_t j = 0; j < n; ++j) {
            a[(size_t)(c * n + j)] *= rp;
            inv[(size_t)(c * n + j)] *= rp;
        }
        for (int64_t i = 0; i < n; ++i) {
            if (i == c) continue;
            const double m = a[(size_t)(i * n + c)];
            if (m == 0.0) continue;
            for (int64_t j = 0; j < n; ++j) {
                a[(size_t)(i * n + j)] -= m * a[(size_t)(c * n + j)];
                inv[(size_t)(i * n + j)] -= m * inv[(size_t)(c * n + j)];
            }
        }
    }
    return true;
}

void cheb_coefficients(MgLevel& l, int degree, double lo, double hi) {
    const double a = lo * l.lmax, b = hi * l.lmax;
    const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
    double rho = 1.0 / sigma;
    l.c1[0] = 0.0f;
    l.c2[0] = (float)(1.0 / theta);
    for (int k = 1; k < degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        l.c1[k] = (float)(rn * rho);
        l.c2[k] = (float)(2.0 * rn / delta);
        rho = rn;
    }
}

template <int ND>
void build_levels(sfem_ctx* ctx, sfem_system* S, MgHierarchy& H, const sfem_mg_opts& o) {
    constexpr int BS = ND, B2 = BS * BS, NS = Sten<ND>::NS;
    hipStream_t s = ctx->stream;
    const sfem_csr* K = S->K;
    MgLevel& L0 = H.L[0];
    const int64_t nn0 = L0.nn;
    const Grid g0 = grid_of(L0);

    // level-0 maps and masks
    DevBuf<uint32_t> flag(ctx, (size_t)nn0), pos(ctx, (size_t)nn0), tot(ctx, 1);
    hipLaunchKernelGGL(node_free_kernel, dim3(rows_grid(nn0)), dim3(kThreads), 0, s, S->free_of.get(), nn0, BS,
                       flag.get());
    SFEM_LAUNCH_CHECK();
    exclusive_scan_u32_dev(ctx, flag.get(), pos.get(), nn0, tot.get());
    uint32_t nrows = 0;
    d2h(ctx, &nrows, tot.get(), sizeof(nrows));
    if ((int64_t)nrows != S->nfb) fail(SFEM_ERR_STATE, "multigrid: block rows do not match the system");
    L0.nrows = nrows;
    H.row_of_node.reset(ctx, (size_t)nn0);
    H.node_of_row.reset(ctx, (size_t)nrows);
    L0.fixed.reset(ctx, (size_t)nrows * BS);
    DevBuf<uint8_t> fixnode(ctx, (size_t)nn0 * BS);
    hipLaunchKernelGGL(row_maps_kernel, dim3(rows_grid(nn0)), dim3(kThreads), 0, s, flag.get(), pos.get(),
                       S->free_of.get(), nn0, BS, H.row_of_node.get(), H.node_of_row.get(), L0.fixed.get(),
                       fixnode.get());
    SFEM_LAUNCH_CHECK();

    // matrix-free finest level: K assembled on this grid with element matrices
    // s_e Ke(0) (within FP32 rounding for the smoother, within a few FP64
    // roundings for the outer operator)
    bool same_grid = K->grid_nd == ND;
    for (int d = 0; d < ND; ++d) same_grid = same_grid && K->grid_ne[d] == L0.n[d] - 1;
    H.ke_dev = same_grid ? K->ke_dev : HUGE_VAL;
    // the matrix-free kernels address vectors with 32-bit byte offsets
    const bool mf_fits = (uint64_t)(S->nfb + 1) * BS * sizeof(double) < (uint64_t(1) << 32);
    if (o.matrix_free > 0 && same_grid && mf_fits && K->ke_dev <= kMfSmootherDev) {
        constexpr int NEN = 1 << ND;
        H.mf = o.matrix_free >= 2 && K->ke_dev <= kMfOuterDev ? 2 : 1;
        std::vector<double> kc((size_t)NEN * NEN * B2);
        std::vector<float> kc32(kc.size());
        for (int a = 0; a < NEN; ++a)
            for (int b = 0; b < NEN; ++b)
                for (int m = 0; m < B2; ++m) {
                    const size_t src = ((size_t)K->grid_corner[a] * NEN + K->grid_corner[b]) * B2 + m;
                    kc[((size_t)a * NEN + b) * B2 + m] = K->ke0[src];
                    kc32[((size_t)a * NEN + b) * B2 + m] = (float)K->ke0[src];
                }
        H.kc64.reset(ctx, kc.size());
        h2d(ctx, H.kc64.get(), kc.data(), kc.size() * sizeof(double));
        H.kc32.reset(ctx, kc32.size());
        h2d(ctx, H.kc32.get(), kc32.data(), kc32.size() * sizeof(float));
        const int64_t n0 = L0.n[0], n1 = L0.n[1], n2 = L0.n[2];
        const int64_t pn = (n0 + 2) * (n1 + 2) * (ND == 3 ? n2 + 2 : 1);
        const int64_t pc = (n0 + 1) * (n1 + 1) * (ND == 3 ? n2 + 1 : 1);
        H.mf_rows.reset(ctx, (size_t)pn);
        H.es32.reset(ctx, (size_t)pc);
        H.es64.reset(ctx, (size_t)pc);
        hipLaunchKernelGGL(ghost_rows_kernel<ND>, dim3(rows_grid(pn)), dim3(kThreads), 0, s,
                           (const int32_t*)H.row_of_node.get(), n0, n1, n2, H.mf_rows.get());
        SFEM_LAUNCH_CHECK();
        H.mf_offs.reset(ctx, (size_t)pn);
        hipLaunchKernelGGL(ghost_offs_kernel, dim3(rows_grid(pn)), dim3(kThreads), 0, s,
                           (const int32_t*)H.mf_rows.get(), pn, S->nfb, BS, H.mf_offs.get());
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL((ghost_cells_kernel<ND, float>), dim3(rows_grid(pc)), dim3(kThreads), 0, s,
                           K->escale.get(), n0 - 1, n1 - 1, n2 - 1, H.es32.get());
        SFEM_LAUNCH_CHECK();
        hipLaunchKernelGGL((ghost_cells_kernel<ND, double>), dim3(rows_grid(pc)), dim3(kThreads), 0, s,
                           K->escale.get(), n0 - 1, n1 - 1, n2 - 1, H.es64.get());
        SFEM_LAUNCH_CHECK();
        // SFEM_MF_MFMA=1: the smoothing steps on the matrix cores.  Off by
        // default: on MI355X the VALU kernel is faster at 257 x 65 x 65 nodes
        // (59 us a step against 66-94 us; the tile's halo, 2.5 x its nodes,
        // and 3 workgroups per CU leave the loads exposed)
        const char* mm = std::getenv("SFEM_MF_MFMA");
        if (ND == 3 && mm && mm[0] == '1') {
            double kmax = 0.0;
            for (double vv : kc) kmax = std::max(kmax, std::fabs(vv));
            int ek = 0;
            (void)std::frexp(kmax, &ek);
            std::vector<_Float16> fr((size_t)2 * 3 * 64 * 4);
            for (int kk = 0; kk < 3; ++kk)
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = 0; e < 4; ++e) {
                        const int i = lane % 32, k = 8 * kk + 4 * (lane / 32) + e;  // A[i][k], i < 24 real
                        float y = 0.0f;
                        if (i < 24) {
                            const int a = i / 3, ii = i % 3, b = k / 3, jj = k % 3;
                            y = (float)std::ldexp(kc[((size_t)a * NEN + b) * B2 + ii * 3 + jj], -ek);
                        }
                        const _Float16 hi = (_Float16)y;
                        fr[((size_t)(0 * 3 + kk) * 64 + lane) * 4 + e] = hi;
                        fr[((size_t)(1 * 3 + kk) * 64 + lane) * 4 + e] = (_Float16)(y - (float)hi);
                    }
            H.kfrag.reset(ctx, fr.size());
            h2d(ctx, H.kfrag.get(), fr.data(), fr.size() * sizeof(_Float16));
            H.kscale = (float)std::ldexp(1.0, ek);
            H.mfma = true;
        }
    }

    // Power iterations (lambda_max(B^-1 A) per level, the coarsest level
    // excepted) run on side streams as soon as a level's smoothing operator
    // exists, while the next levels are built: the finest level's first when
    // it is matrix-free.  Every level has its own vectors, sums and PowState;
    // all of them live until the side streams are joined (join syncs them
    // when setup fails).  A side stream starts after an event recorded on the
    // context's stream once its buffers are allocated, so memory the cache
    // hands out again is past its earlier users.
    DevBuf<PowState> ps(ctx, SFEM_MG_MAX_LEVELS);
    memset_d(ctx, ps.get(), 0, sizeof(PowState) * SFEM_MG_MAX_LEVELS);
    DevBuf<double> scales(ctx, SFEM_MG_MAX_LEVELS);  // FP16 stencil scales, 0: zero operator
    std::vector<DevBuf<double>> parts(SFEM_MG_MAX_LEVELS);
    DevBuf<float> p0a, p0b;
    struct SideJoin {
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> done;
        ~SideJoin() {
            for (hipStream_t t : streams) (void)hipStreamSynchronize(t);
            for (hipEvent_t e : done) (void)hipEventDestroy(e);
        }
    } join;
    auto fork = [&](hipStream_t t) {
        hipEvent_t f;
        SFEM_HIP(hipEventCreateWithFlags(&f, hipEventDisableTiming));
        SFEM_HIP(hipEventRecord(f, s));
        SFEM_HIP(hipStreamWaitEvent(t, f, 0));
        (void)hipEventDestroy(f);
        if (std::find(join.streams.begin(), join.streams.end(), t) == join.streams.end()) join.streams.push_back(t);
    };
    auto finish = [&](hipStream_t t) {
        hipEvent_t d;
        SFEM_HIP(hipEventCreateWithFlags(&d, hipEventDisableTiming));
        join.done.push_back(d);
        SFEM_HIP(hipEventRecord(d, t));
    };
    auto init_grid = [&](int64_t n) { return std::min<unsigned>(rows_grid(n), 1024); };
    // v's vectors: da, db (n + zero) floats
    auto power_start = [&](int l, const MgLevel& v, float* va, float* vb, unsigned nparts, hipStream_t t) {
        const int64_t n = v.nrows * BS;
        parts[(size_t)l].reset(ctx, std::max({nparts, init_grid(n), 1024u}));
        fork(t);
        hipLaunchKernelGGL(pow_init_kernel, dim3(init_grid(n)), dim3(kThreads), 0, t, n, v.fixed.get(), va,
                           ps.get() + l, parts[(size_t)l].get());
        SFEM_LAUNCH_CHECK();
    };
    auto power_coarse = [&](int l, const MgLevel& v) {
        const hipStream_t t = side_stream(ctx, (l - 1) % (sfem_ctx::kSideStreams - 1));
        const bool grp = v.node_major;
        const dim3 gp(std::min(rows_grid(v.nn * (grp ? Lanes<ND>::G : 1)), 1024u));
        power_start(l, v, v.da.get(), v.db.get(), gp.x, t);
        float* vin = v.da.get();
        float* vout = v.db.get();
        double* partials = parts[(size_t)l].get();
        for (int k = 0; k < o.power_iters; ++k) {
            // grid-stride over the nodes: few blocks meet at the last-block count
            auto go = [&](auto kern, auto op) {
                hipLaunchKernelGGL(kern, gp, dim3(kThreads), 0, t, op, vin, v.binv.get(), vout, ps.get() + l,
                                   partials);
            };
            if (v.st16.size()) {
                const StOp<_Float16> op{v.st16.get(), grid_of(v), v.nn};
                grp ? go(pow_stencil_kernel<ND, true, _Float16>, op) : go(pow_stencil_kernel<ND, false, _Float16>, op);
            } else {
                const StOp<float> op{v.st32.get(), grid_of(v), v.nn};
                grp ? go(pow_stencil_kernel<ND, true, float>, op) : go(pow_stencil_kernel<ND, false, float>, op);
            }
            SFEM_LAUNCH_CHECK();
            std::swap(vin, vout);
        }
        finish(t);
    };
    // The finest level's, when it is matrix-free, starts once the first
    // coarse level is built: the first Galerkin product (the longest kernel
    // of setup) runs alone.
    auto power_fine_mf = [&]() {
        const int64_t n = L0.nrows * BS;
        p0a.reset(ctx, (size_t)(n + BS));
        p0b.reset(ctx, (size_t)(n + BS));
        memset_d(ctx, p0a.get(), 0, p0a.bytes());
        memset_d(ctx, p0b.get(), 0, p0b.bytes());
        // a row per thread, blocks mapped XCD-contiguously as in the smoothing step
        const unsigned gm = xcd_rows_grid(L0.nrows);
        const hipStream_t t = side_stream(ctx, sfem_ctx::kSideStreams - 1);
        power_start(0, L0, p0a.get(), p0b.get(), gm, t);
        float* vin = p0a.get();
        float* vout = p0b.get();
        for (int k = 0; k < o.power_iters; ++k) {
            hipLaunchKernelGGL(pow_mf_kernel<ND>, dim3(gm), dim3(kThreads), 0, t, mf_op32(S, H), vin,
                               S->binv32.get(), vout, ps.get(), parts[0].get());
            SFEM_LAUNCH_CHECK();
            std::swap(vin, vout);
        }
        finish(t);
    };

    // the sliced values of K_ff: read by the finest level unless it is matrix-free
    if (!H.mf) ensure_sell(ctx, S);

    // fine stencil (FP64, fixed rows and columns zero): from the element
    // scales when the finest level is matrix-free (K on this grid couples
    // grid neighbours only) -- formed inside the first Galerkin product, not
    // stored -- else from K
    DevBuf<double> stf;
    if (!H.mf) {
        stf.reset(ctx, (size_t)nn0 * NS * B2);
        memset_d(ctx, stf.get(), 0, stf.bytes());
        DevBuf<int> bad(ctx, 1);
        memset_d(ctx, bad.get(), 0, sizeof(int));
        hipLaunchKernelGGL(stencil_from_k_kernel<ND>, dim3(rows_grid(K->nnzb)), dim3(kThreads), 0, s,
                           K->bcolidx.get(), K->browidx.get(), K->bval.get(), K->nnzb, nn0, S->free_of.get(), g0,
                           stf.get(), bad.get());
        SFEM_LAUNCH_CHECK();
        int hbad = 0;
        d2h(ctx, &hbad, bad.get(), sizeof(int));
        if (hbad) fail(SFEM_ERR_ARGUMENT, "multigrid: K couples nodes that are not neighbours on the given grid");
    }

    // coarse levels
    const double* prev_st = H.mf ? nullptr : stf.get();
    const uint8_t* prev_fix = fixnode.get();
    for (;;) {
        const MgLevel& P = H.L.back();
        int f[3] = {1, 1, 1};
        bool any = false;
        for (int d = 0; d < ND; ++d) {
            if (P.n[d] - 1 >= 2) {  // any element count from 2 halves (rounded up)
                f[d] = 2;
                any = true;
            }
        }
        if (!any) break;
        if (H.L.size() > 1 && P.nn <= o.coarse_nodes) break;
        MgLevel C;
        for (int d = 0; d < 3; ++d) {
            C.f[d] = f[d];
            C.n[d] = d < ND ? (f[d] == 2 ? P.n[d] / 2 + 1 : P.n[d]) : 1;  // ceil(e / 2) elements
        }
        C.nn = C.n[0] * C.n[1] * C.n[2];
        C.nrows = C.nn;
        const Grid gf = grid_of(P), gc = grid_of(C);
        C.fixed.reset(ctx, (size_t)C.nn * BS);
        hipLaunchKernelGGL(coarse_fixed_kernel<ND>, dim3(rows_grid(C.nn)), dim3(kThreads), 0, s, prev_fix, gf, gc,
                           f[0], f[1], f[2], C.fixed.get());
        SFEM_LAUNCH_CHECK();
        // P'AP: cell by cell from the element scales (the first coarse level
        // under a matrix-free finest level, every direction coarsened), else
        // one direction at a time (x, then y, then z)
        DevBuf<double> raw, tmp;
        const double* cur = prev_st;
        Grid gcur = gf;
        int last = -1;
        for (int d = 0; d < ND; ++d)
            if (f[d] == 2) last = d;
        bool by_cells = !prev_st && !galerkin_lines_forced();
        for (int d = 0; d < ND; ++d) by_cells = by_cells && f[d] == 2;
        // the smoothing kernels address a coarse stencil with 32-bit byte offsets
        if ((uint64_t)C.nn * NS * B2 * sizeof(float) >= (1ull << 32))
            fail(SFEM_ERR_UNSUPPORTED, "multigrid: a coarse level's stencil exceeds 4 GiB (" + std::to_string(C.nn) +
                                           " nodes)");
        C.st64.reset(ctx, (size_t)C.nn * NS * B2);
        C.st32.reset(ctx, (size_t)C.nn * NS * B2);
        C.node_major = C.nn < stencil_group_below();
        if (by_cells) {  // symmetric bit for bit and masked: no symmetrising pass
            constexpr int NEN = 1 << ND;
            DevBuf<double> gt(ctx, (size_t)NEN * NEN * NEN * B2);
            hipLaunchKernelGGL(galerkin_table_kernel<ND>, dim3(rows_grid((int64_t)gt.size())), dim3(kThreads), 0, s,
                               (const double*)H.kc64.get(), gt.get());
            SFEM_LAUNCH_CHECK();
            const int64_t ncell = (C.n[0] - 1) * (C.n[1] - 1) * (ND == 3 ? C.n[2] - 1 : 1);
            DevBuf<int32_t> slot(ctx, (size_t)ncell), cells(ctx, (size_t)ncell);
            DevBuf<unsigned> cnt(ctx, 1);
            memset_d(ctx, cnt.get(), 0, sizeof(unsigned));
            hipLaunchKernelGGL(cell_irregular_kernel<ND>, dim3(rows_grid(ncell)), dim3(kThreads), 0, s, prev_fix, gf,
                               gc, slot.get(), cells.get(), cnt.get());
            SFEM_LAUNCH_CHECK();
            unsigned nirr = 0;
            d2h(ctx, &nirr, cnt.get(), sizeof(unsigned));
            DevBuf<double> kcell(ctx, (size_t)std::max(nirr, 1u) * NEN * NEN * B2);
            if (nirr) {
                hipLaunchKernelGGL(cell_matrix_kernel<ND>, dim3(nirr), dim3(64), 0, s, (const double*)gt.get(),
                                   (const double*)H.kc64.get(), (const double*)H.es64.get(), prev_fix, gf, gc,
                                   (const int32_t*)cells.get(), kcell.get());
                SFEM_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(galerkin_cells_kernel<ND>, dim3(rows_grid(C.nn), NS), dim3(kThreads), 0, s,
                               (const double*)gt.get(), (const double*)kcell.get(), (const double*)H.es64.get(),
                               (const int32_t*)slot.get(), gf, gc, (const uint8_t*)C.fixed.get(), C.st64.get(),
                               C.st32.get(), C.node_major);
            SFEM_LAUNCH_CHECK();
        }
        for (int d = 0; d < ND && !by_cells; ++d) {
            if (f[d] != 2) continue;
            Grid gn = gcur;
            int64_t* nd = d == 0 ? &gn.n0 : d == 1 ? &gn.n1 : &gn.n2;
            *nd = *nd / 2 + 1;  // e = n - 1 elements -> ceil(e / 2)
            const int64_t nno = gn.n0 * gn.n1 * gn.n2;
            DevBuf<double> o(ctx, (size_t)nno * NS * B2);
            const dim3 gl(rows_grid(nno), NS / 3);
            if (!cur)  // the matrix-free finest level
                hipLaunchKernelGGL((galerkin_line_kernel<ND, StencilMf<ND>>), gl, dim3(kThreads), 0, s,
                                   StencilMf<ND>{H.kc64.get(), H.es64.get(), S->free_of.get(), gcur}, gcur, d, gn,
                                   o.get());
            else
                hipLaunchKernelGGL((galerkin_line_kernel<ND, StencilMem<ND>>), gl, dim3(kThreads), 0, s,
                                   StencilMem<ND>{cur, gcur.n0 * gcur.n1 * gcur.n2}, gcur, d, gn, o.get());
            SFEM_LAUNCH_CHECK();
            if (d == last) {
                raw = std::move(o);
            } else {
                tmp = std::move(o);  // the previous intermediate goes back to the allocator in stream order
                cur = tmp.get();
            }
            gcur = gn;
        }
        if (!by_cells) {
            hipLaunchKernelGGL(sym_mask_kernel<ND>, dim3(rows_grid(C.nn), NS), dim3(kThreads), 0, s, raw.get(), gc,
                               C.fixed.get(), C.st64.get(), C.st32.get(), C.node_major);
            SFEM_LAUNCH_CHECK();
        }
        C.binv.reset(ctx, (size_t)C.nn * B2);
        hipLaunchKernelGGL(stencil_binv_kernel<ND>, dim3(rows_grid(C.nn)), dim3(kThreads), 0, s, C.st64.get(),
                           C.fixed.get(), C.nn, C.binv.get());
        SFEM_LAUNCH_CHECK();
        // unless C is the coarsest level (the loop's exit tests for the next
        // pass; the coarsest is solved directly): its smoothing operator, FP16
        // values with a power-of-two scale (the scale found on the device),
        // then its power iteration on a side stream
        bool more = false;
        for (int d = 0; d < ND; ++d) more = more || C.n[d] - 1 >= 2;
        const int lc = (int)H.L.size();
        if (more && C.nn > o.coarse_nodes && lc + 1 < SFEM_MG_MAX_LEVELS) {
            if (o.coarse_fp16) {
                const int64_t nv = C.nn * NS * B2;
                DevBuf<unsigned> mx(ctx, 1);
                memset_d(ctx, mx.get(), 0, sizeof(unsigned));
                hipLaunchKernelGGL(stencil_maxdiag_kernel<ND>, dim3(std::min(rows_grid(C.nn * BS), 256u)),
                                   dim3(kThreads), 0, s, (const float*)C.st32.get(), C.nn, C.node_major, mx.get());
                SFEM_LAUNCH_CHECK();
                hipLaunchKernelGGL(half_scale_kernel, dim3(1), dim3(1), 0, s, (const unsigned*)mx.get(),
                                   scales.get() + lc);
                SFEM_LAUNCH_CHECK();
                C.st16.reset(ctx, (size_t)nv);
                hipLaunchKernelGGL(to_half_dev_kernel, dim3(std::min(rows_grid(nv), 2048u)), dim3(kThreads), 0, s,
                                   (const float*)C.st32.get(), nv, (const double*)scales.get() + lc, C.st16.get());
                SFEM_LAUNCH_CHECK();
                hipLaunchKernelGGL(stencil_binv_half_kernel<ND>, dim3(rows_grid(C.nn)), dim3(kThreads), 0, s,
                                   (const _Float16*)C.st16.get(), C.nn, C.node_major, C.binv.get());
                SFEM_LAUNCH_CHECK();
                C.st32.release();
            }
            C.da.reset(ctx, (size_t)C.nn * BS);
            C.db.reset(ctx, (size_t)C.nn * BS);
            power_coarse(lc, C);
        }
        if (lc == 1 && H.mf) power_fine_mf();
        H.L.push_back(std::move(C));
        prev_st = H.L.back().st64.get();
        prev_fix = H.L.back().fixed.get();
        if ((int)H.L.size() >= SFEM_MG_MAX_LEVELS) break;
    }
    sync(ctx);
    stf.release();
    if (H.L.size() < 2)
        fail(SFEM_ERR_UNSUPPORTED, "multigrid: no direction has 2 or more elements to coarsen");
    MgLevel& Lc = H.L.back();
    if (Lc.nn > kMgMaxCoarseNodes)
        fail(SFEM_ERR_UNSUPPORTED, "multigrid: the grid coarsens only to " + std::to_string(Lc.nn) +
                                       " nodes (at most " + std::to_string(kMgMaxCoarseNodes) +
                                       ")");

    // coarsest: dense inverse on the host
    {
        const int64_t nn = Lc.nn, nc = nn * BS;
        std::vector<double> st((size_t)(nn * NS * B2));
        std::vector<uint8_t> fx((size_t)nc);
        d2h(ctx, st.data(), Lc.st64.get(), st.size() * sizeof(double));
        d2h(ctx, fx.data(), Lc.fixed.get(), fx.size());
        std::vector<double> a((size_t)(nc * nc), 0.0), inv;
        const Grid gc = grid_of(Lc);
        for (int64_t I = 0; I < nn; ++I) {
            int64_t c[3];
            coords_of<ND>(I, gc, c);
            for (int sl = 0; sl < NS; ++sl) {
                int dx, dy, dz;
                offset_of<ND>(sl, dx, dy, dz);
                if (!inside<ND>(gc, c[0] + dx, c[1] + dy, c[2] + dz)) continue;
                const int64_t J = node_at(gc, c[0] + dx, c[1] + dy, c[2] + dz);
                for (int p = 0; p < BS; ++p)
                    for (int q = 0; q < BS; ++q)
                        a[(size_t)((I * BS + p) * nc + J * BS + q)] = st[(size_t)(((int64_t)sl * B2 + p * BS + q) * nn + I)];
            }
        }
        for (int64_t d = 0; d < nc; ++d)
            if (fx[(size_t)d]) a[(size_t)(d * nc + d)] = 1.0;
        if (!dense_inverse(a, nc, inv)) fail(SFEM_ERR_SINGULAR, "multigrid: the coarsest operator is singular");
        H.nc = nc;
        H.cinv.reset(ctx, (size_t)(nc * nc));
        h2d(ctx, H.cinv.get(), inv.data(), inv.size() * sizeof(double));
    }

    // FP16 smoothing operators on the levels between the finest and the
    // coarsest (made as the levels were built): their scales
    if (o.coarse_fp16 && H.L.size() > 2) {
        std::vector<double> hs(H.L.size());
        d2h(ctx, hs.data(), scales.get(), sizeof(double) * hs.size());
        for (size_t l = 1; l + 1 < H.L.size(); ++l) {
            if (!(hs[l] > 0.0)) fail(SFEM_ERR_SINGULAR, "multigrid: zero coarse operator");
            H.L[l].scale = hs[l];
        }
    }

    // FP16 finest-level smoothing operator
    if (o.fine_fp16 && !H.mf) {
        const int64_t nv = S->sell_blocks * B2;
        ensure_dinv(ctx, S);  // its largest entry sets the scale
        DevBuf<unsigned long long> mx(ctx, 1);
        memset_d(ctx, mx.get(), 0, sizeof(unsigned long long));
        hipLaunchKernelGGL(maxdiag_kernel, dim3(256), dim3(kThreads), 0, s, S->dinv64.get(), S->nf, mx.get());
        SFEM_LAUNCH_CHECK();
        unsigned long long hm = 0;
        d2h(ctx, &hm, mx.get(), sizeof(hm));
        double amax;
        std::memcpy(&amax, &hm, sizeof(amax));
        if (S->partial) amax = std::max(amax, 1.0);  // unit diagonal of fixed dofs in the block rows
        if (!(amax > 0.0) || !std::isfinite(amax)) fail(SFEM_ERR_SINGULAR, "multigrid: zero or non-finite operator");
        int e = 0;
        std::frexp(amax, &e);  // amax = f 2^e, f in [1/2, 1)
        H.scale0 = std::ldexp(1.0, -e);
        H.h0.reset(ctx, (size_t)nv);
        hipLaunchKernelGGL(to_half_kernel, dim3(2048), dim3(kThreads), 0, s, S->sval32.get(), nv, (float)H.scale0,
                           H.h0.get());
        SFEM_LAUNCH_CHECK();
        H.binv0.reset(ctx, (size_t)S->nfb * B2);
        hipLaunchKernelGGL(binv_half_kernel<BS>, dim3(rows_grid(S->nfb)), dim3(kThreads), 0, s, S->sl_ptr.get(),
                           S->sl_w.get(), S->scol.get(), (const _Float16*)H.h0.get(), S->nfb, H.binv0.get());
        SFEM_LAUNCH_CHECK();
        H.fp16 = true;
    }

    // work vectors (level 0: and a zero row, read for fixed nodes by the
    // matrix-free operator)
    for (size_t l = 0; l < H.L.size(); ++l) {
        MgLevel& v = H.L[l];
        const size_t n = (size_t)v.nrows * BS, na = l == 0 ? n + BS : n;
        v.r.reset(ctx, na);
        v.x.reset(ctx, na);
        if (!v.da.get()) {  // a level's power iteration may be using its own
            v.da.reset(ctx, na);
            v.db.reset(ctx, na);
        }
        if (l == 0)
            for (float* b : {v.r.get(), v.x.get(), v.da.get(), v.db.get()}) memset_d(ctx, b, 0, na * sizeof(float));
        if (l > 0) v.b.reset(ctx, n);
    }

    // lambda_max(B^-1 A) of the finest level when it is not matrix-free (on
    // the context's stream); then every level's result
    const int nl = (int)H.L.size();
    if (!H.mf) {
        MgLevel& v = H.L[0];
        const int64_t n = v.nrows * BS;
        const Launch ls = plan_sell(S->nslices);
        parts[0].reset(ctx, std::max({ls.grid, init_grid(n), 1024u}));
        double* partials = parts[0].get();
        hipLaunchKernelGGL(pow_init_kernel, dim3(init_grid(n)), dim3(kThreads), 0, s, n, v.fixed.get(), v.da.get(),
                           ps.get(), partials);
        SFEM_LAUNCH_CHECK();
        float* vin = v.da.get();
        float* vout = v.db.get();
        for (int k = 0; k < o.power_iters; ++k) {
            if (H.fp16) {
                const SellOp<_Float16> op{S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), H.h0.get(), S->nslices,
                                          S->nfb, ls.per_block};
                hipLaunchKernelGGL((pow_sell_kernel<BS, _Float16>), dim3(ls.grid), dim3(kThreads), 0, s, op, vin,
                                   (const float*)H.binv0.get(), vout, ps.get(), partials);
            } else {
                const SellOp<float> op{S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval32.get(), S->nslices,
                                       S->nfb, ls.per_block};
                hipLaunchKernelGGL((pow_sell_kernel<BS, float>), dim3(ls.grid), dim3(kThreads), 0, s, op, vin,
                                   S->binv32.get(), vout, ps.get(), partials);
            }
            SFEM_LAUNCH_CHECK();
            std::swap(vin, vout);
        }
    }
    for (hipEvent_t d : join.done) SFEM_HIP(hipStreamWaitEvent(s, d, 0));
    std::vector<PowState> hps((size_t)nl);
    d2h(ctx, hps.data(), ps.get(), sizeof(PowState) * nl);  // the side streams joined: every level done
    for (int l = 0; l + 1 < nl; ++l) {
        H.L[l].lmax = hps[(size_t)l].lam;
        if (!(H.L[l].lmax > 0.0) || !std::isfinite(H.L[l].lmax))
            fail(SFEM_ERR_SINGULAR, "multigrid: no spectrum estimate on level " + std::to_string(l));
        H.L[l].degree = l == 0 || o.coarse_degree == 0 ? o.degree : o.coarse_degree;
        cheb_coefficients(H.L[l], H.L[l].degree, o.lo, o.hi);
    }
}

// One preconditioned CG solve; TV: outer operator values.
template <int ND, typename TV>
struct MgRun {
    static constexpr int BS = ND;
    sfem_ctx* ctx;
    sfem_system* S;
    MgHierarchy& H;
    hipStream_t s;
    Launch ls;
    const int* done;
    bool nt;
    std::vector<hipEvent_t>* ev = nullptr;  // non-null: time the finest-level smoothing steps
    size_t ev_next = 0;
    bool time_vc = false;  // time this V-cycle's steps
    void step(int l, int mode, const float* v, const float* d_in, float* d_out, float c1, float c2) {
        MgLevel& L = H.L[(size_t)l];
        // timed: the finest-level smoothing steps
        const bool timed = ev && time_vc && l == 0 && mode == 0 && 2 * ev_next + 1 < ev->size();
        if (timed) SFEM_HIP(hipEventRecord((*ev)[2 * ev_next], s));
        if (l == 0 && H.mf && H.mfma) {  // the matrix cores (3D)
            const MfTiles T = mf_tiles(H);
            const dim3 grid((unsigned)((T.tx * T.ty * T.tz + 7) / 8 * 8));
            const half4_t* kf = (const half4_t*)H.kfrag.get();
            const float* bi = (const float*)S->binv32.get();
            if (mode == 0)
                hipLaunchKernelGGL((cheb_mf_mfma_kernel<0>), grid, dim3(kThreads), 0, s, mf_op32(S, H), T, kf,
                                   H.kscale, v, L.r.get(), L.x.get(), d_in, d_out, c1, c2, bi, done);
            else
                hipLaunchKernelGGL((cheb_mf_mfma_kernel<1>), grid, dim3(kThreads), 0, s, mf_op32(S, H), T, kf,
                                   H.kscale, v, L.r.get(), L.x.get(), d_in, d_out, c1, c2, bi, done);
        } else if (l == 0 && H.mf) {
            auto go = [&](auto k) {
                hipLaunchKernelGGL(k, dim3(xcd_rows_grid(L.nrows)), dim3(kThreads), 0, s, mf_op32(S, H), v, L.r.get(),
                                   L.x.get(), d_in, d_out, c1, c2, (const float*)S->binv32.get(), done);
            };
            if (mode == 0) go(cheb_mf_kernel<ND, 0>);
            else go(cheb_mf_kernel<ND, 1>);
        } else if (l == 0 && H.fp16) {
            const SellOp<_Float16> op{S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), H.h0.get(), S->nslices, S->nfb,
                                      ls.per_block};
            auto go = [&](auto k) {
                hipLaunchKernelGGL(k, dim3(ls.grid), dim3(kThreads), 0, s, op, v, L.r.get(), L.x.get(), d_in, d_out,
                                   c1, c2, (const float*)H.binv0.get(), done);
            };
            if (mode == 0) nt ? go(cheb_sell_kernel<BS, 0, true, _Float16>) : go(cheb_sell_kernel<BS, 0, false, _Float16>);
            else nt ? go(cheb_sell_kernel<BS, 1, true, _Float16>) : go(cheb_sell_kernel<BS, 1, false, _Float16>);
        } else if (l == 0) {
            const SellOp<float> op{S->sl_ptr.get(), S->sl_w.get(), S->scol.get(), S->sval32.get(), S->nslices, S->nfb,
                                   ls.per_block};
            auto go = [&](auto k) {
                hipLaunchKernelGGL(k, dim3(ls.grid), dim3(kThreads), 0, s, op, v, L.r.get(), L.x.get(), d_in, d_out,
                                   c1, c2, (const float*)S->binv32.get(), done);
            };
            if (mode == 0) nt ? go(cheb_sell_kernel<BS, 0, true, float>) : go(cheb_sell_kernel<BS, 0, false, float>);
            else nt ? go(cheb_sell_kernel<BS, 1, true, float>) : go(cheb_sell_kernel<BS, 1, false, float>);
        } else {
            const bool grp = L.node_major;
            auto go = [&](auto k, auto op) {
                hipLaunchKernelGGL(k, dim3(rows_grid(L.nn * (grp ? Lanes<ND>::G : 1))), dim3(kThreads), 0, s, op, v,
                                   L.r.get(), L.x.get(), d_in, d_out, c1, c2, (const float*)L.binv.get(), done);
            };
            auto pick = [&](auto op, auto tv) {
                using TS = decltype(tv);
                if (grp) mode == 0 ? go(cheb_stencil_kernel<ND, 0, true, TS>, op) : go(cheb_stencil_kernel<ND, 1, true, TS>, op);
                else mode == 0 ? go(cheb_stencil_kernel<ND, 0, false, TS>, op) : go(cheb_stencil_kernel<ND, 1, false, TS>, op);
            };
            if (L.st16.size()) pick(StOp<_Float16>{L.st16.get(), grid_of(L), L.nn}, _Float16(0));
            else pick(StOp<float>{L.st32.get(), grid_of(L), L.nn}, 0.0f);
        }
        SFEM_LAUNCH_CHECK();
        if (timed) SFEM_HIP(hipEventRecord((*ev)[2 * ev_next++ + 1], s));
    }

    // x_l = V-cycle(b_l); level 0's right-hand side is the outer residual.
    // Levels below 0 start after the restriction that formed b_l, which also
    // made their first smoothing step.
    void vcycle(int l, const double* r_outer) {
        MgLevel& L = H.L[(size_t)l];
        const int nl = (int)H.L.size();
        if (l == nl - 1) {
            hipLaunchKernelGGL(dense_apply_kernel, dim3(1), dim3(kThreads), 0, s, H.cinv.get(), H.nc, L.b.get(),
                               L.x.get(), done);
            SFEM_LAUNCH_CHECK();
            return;
        }
        if (l == 0) {
            const float* binv = H.fp16 ? H.binv0.get() : S->binv32.get();
            hipLaunchKernelGGL((cheb_first_kernel<BS, double>), dim3(rows_grid(L.nrows)), dim3(kThreads), 0, s, r_outer,
                               H.scale0, L.nrows, binv, L.c2[0], L.r.get(), L.x.get(), L.da.get(), done);
            SFEM_LAUNCH_CHECK();
        }
        float* dp = L.da.get();
        float* dn = L.db.get();
        for (int k = 1; k < L.degree; ++k) {
            step(l, 0, dp, dp, dn, L.c1[k], L.c2[k]);
            std::swap(dp, dn);
        }
        step(l, 1, dp, nullptr, nullptr, 0.0f, 0.0f);
        MgLevel& C = H.L[(size_t)l + 1];
        const Grid gf = grid_of(L), gc = grid_of(C);
        auto rk = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(rows_grid(C.nn)), dim3(kThreads), 0, s, (const float*)L.r.get(),
                               (float)(1.0 / (l == 0 ? H.scale0 : L.scale)), gf,
                               l == 0 ? (const int32_t*)H.row_of_node.get() : nullptr, gc, C.f[0], C.f[1], C.f[2],
                               (const uint8_t*)C.fixed.get(), C.b.get(), (const float*)C.binv.get(), C.c2[0],
                               (float)C.scale, C.r.get(), C.x.get(), C.da.get(), done);
        };
        if (l + 1 < nl - 1) rk(restrict_kernel<ND, true>);
        else rk(restrict_kernel<ND, false>);
        SFEM_LAUNCH_CHECK();
        vcycle(l + 1, nullptr);
        hipLaunchKernelGGL(prolong_kernel<ND>, dim3(rows_grid(L.nrows)), dim3(kThreads), 0, s,
                           (const float*)C.x.get(), gc, gf, C.f[0], C.f[1], C.f[2],
                           l == 0 ? (const int32_t*)H.node_of_row.get() : nullptr, L.nrows,
                           (const uint8_t*)L.fixed.get(), L.x.get(), L.da.get(), done);
        SFEM_LAUNCH_CHECK();
        step(l, 0, L.da.get(), nullptr, L.db.get(), 0.0f, L.c2[0]);
        dp = L.db.get();
        dn = L.da.get();
        for (int k = 1; k < L.degree; ++k) {
            step(l, 0, dp, dp, dn, L.c1[k], L.c2[k]);
            std::swap(dp, dn);
        }
    }
};

template <typename TV>
Vals<TV> outer_vals(const sfem_system* S) {
    if constexpr (std::is_same<TV, SplitVal>::value) return {S->sval32.get(), S->sval16.get()};
    else if constexpr (std::is_same<TV, float>::value) return {S->sval32.get()};
    else return {S->sval64.get()};
}

template <int ND, typename TV>
void mg_pcg(sfem_ctx* ctx, sfem_system* S, MgHierarchy& H, const sfem_solve_opts* o, sfem_solve_stats* stats) {
    constexpr int BS = ND;
    hipStream_t s = ctx->stream;
    const bool padded = S->partial;
    const int64_t n = S->nfb * BS;
    double* x = padded ? S->px.get() : S->x.get();
    const double* rhs = padded ? S->prhs.get() : S->rhs.get();
    const Launch ls = plan_sell(S->nslices);
    const Launch lv = plan_vec(n, (unsigned)std::max(8, ctx->n_cu * 4));
    // p: and a zero row, read for fixed nodes by the matrix-free operator
    DevBuf<double> r(ctx, (size_t)n), p(ctx, (size_t)n + BS), q(ctx, (size_t)n);
    // FP64 outer operator matrix-free when the hierarchy allows it
    const bool mf_outer = H.mf == 2 && std::is_same<TV, double>::value;
    const unsigned mf_grid = std::min(xcd_rows_grid(S->nfb), 8192u);
    MfOp<double> mf64{};
    if (mf_outer) mf64 = mf_op<double>(S, H, H.kc64.get(), H.es64.get());
    else ensure_sell(ctx, S);  // the outer operator reads the sliced values
    unsigned tiles = 0;
    if (H.mfma) {
        const MfTiles T = mf_tiles(H);
        tiles = (unsigned)((T.tx * T.ty * T.tz + 7) / 8 * 8);
    }
    DevBuf<double> partials(ctx, 2 * (size_t)std::max({ls.grid, lv.grid, mf_grid, 2048u, xcd_rows_grid(S->nfb),
                                                        tiles}));
    DevBuf<MgState> dst(ctx, 1);
    if (o->x0_zero) memset_d(ctx, S->x.get(), 0, sizeof(double) * S->nf);
    if (padded) pad_gather(ctx, S, S->x.get(), x);
    memset_d(ctx, p.get(), 0, sizeof(double) * (n + BS));
    MgState init{};
    init.best_true = HUGE_VAL;
    init.need_true = 1;
    init.restart = 1;
    if (o->x0_zero) {  // x = 0: the residual is rhs, no operator pass
        SFEM_HIP(hipMemcpyAsync(r.get(), rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        init.need_true = 0;
    }
    h2d(ctx, dst.get(), &init, sizeof(init));
    MgState* st = dst.get();
    MgRun<ND, TV> run{ctx, S, H, s, ls, &st->done, ctx->tune_nt};
    const float* z = H.L[0].x.get();
    Vals<TV> vals = outer_vals<TV>(S);

    hipEvent_t t0, t1;
    SFEM_HIP(hipEventCreate(&t0));
    SFEM_HIP(hipEventCreate(&t1));
    SFEM_HIP(hipEventRecord(t0, s));
    hipLaunchKernelGGL(mg_norm_kernel, dim3(lv.grid), dim3(kThreads), 0, s, rhs, n, lv.per_block, st, partials.get(),
                       o->rtol);
    SFEM_LAUNCH_CHECK();
    // FP64 residual of the assembled K: from K's block columns when the
    // finest level is matrix-free (K assembled on the grid, exactly
    // symmetric; the sliced values need not exist), else the sliced values
    const unsigned tk_grid = std::min(rows_grid(S->nfb * 32), 2048u);  // 32 lanes per row
    auto true_resid = [&] {
        if (H.mf)
            hipLaunchKernelGGL(mg_true_k_kernel<BS>, dim3(tk_grid), dim3(kThreads), 0, s, S->K->bcolptr.get(),
                               S->K->browidx.get(), S->K->bval.get(), H.node_of_row.get(), H.row_of_node.get(),
                               S->free_of.get(), S->nfb, (const double*)x, rhs, r.get(), st, partials.get());
        else
            hipLaunchKernelGGL(mg_true_kernel<BS>, dim3(ls.grid), dim3(kThreads), 0, s, S->sl_ptr.get(),
                               S->sl_w.get(), S->scol.get(), S->sval64.get(), S->nslices, S->nfb, ls.per_block,
                               (const double*)x, rhs, r.get(), st, partials.get());
        SFEM_LAUNCH_CHECK();
    };
    MgState host{};
    int64_t enq = 0, counted = 0, timed = 0;
    double smooth_ms = 0.0;
    const int chunk_max = std::max(1, std::min(o->check_every, 8));
    // finest-level smoothing steps per V-cycle; with time_kernels those of
    // every third V-cycle are timed (each timed launch adds two timestamps)
    const int per_iter = 2 * H.L[0].degree - 1;
    constexpr int kTimeEvery = 3;
    std::vector<hipEvent_t> ev;
    if (o->time_kernels) {
        ev.resize(2 * (size_t)per_iter * ((chunk_max + kTimeEvery - 1) / kTimeEvery));
        for (auto& e : ev) SFEM_HIP(hipEventCreate(&e));
        run.ev = &ev;
    }
    d2h(ctx, &host, st, sizeof(host));
    // iterations enqueued per poll: chunk_max, then the count the observed
    // contraction of r.r predicts to the stop (V-cycles enqueued past the
    // stop exit at once but still cost their launches)
    int64_t next = chunk_max, seen_iters = 0;
    double seen_rr = host.bb;
    while (!host.done && enq < o->max_iter) {
        const int64_t chunk = std::min<int64_t>(next, o->max_iter - enq);
        run.ev_next = 0;
        for (int64_t k = 0; k < chunk; ++k) {
            true_resid();
            run.time_vc = k % kTimeEvery == 0;
            run.vcycle(0, r.get());
            hipLaunchKernelGGL(mg_rz_kernel, dim3(lv.grid), dim3(kThreads), 0, s, (const double*)r.get(), z, n,
                               lv.per_block, st, partials.get());
            SFEM_LAUNCH_CHECK();
            hipLaunchKernelGGL(mg_p_kernel, dim3(lv.grid), dim3(kThreads), 0, s, p.get(), z, n, lv.per_block,
                               (const MgState*)st);
            SFEM_LAUNCH_CHECK();
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(ls.grid), dim3(kThreads), 0, s, S->sl_ptr.get(), S->sl_w.get(),
                                   S->scol.get(), vals, S->nslices, S->nfb, ls.per_block, (const double*)p.get(),
                                   q.get(), st, partials.get());
            };
            if (mf_outer) hipLaunchKernelGGL(mg_spmv_mf_kernel<ND>, dim3(mf_grid), dim3(kThreads), 0, s, mf64,
                                             (const double*)p.get(), q.get(), st, partials.get());
            else if (ctx->tune_nt) go(mg_spmv_kernel<BS, TV, true>);
            else go(mg_spmv_kernel<BS, TV, false>);
            SFEM_LAUNCH_CHECK();
            hipLaunchKernelGGL(mg_update_kernel, dim3(lv.grid), dim3(kThreads), 0, s, x, r.get(),
                               (const double*)p.get(), (const double*)q.get(), n, lv.per_block, st, partials.get());
            SFEM_LAUNCH_CHECK();
        }
        true_resid();  // the check a converged recurrence asked for
        enq += chunk;
        d2h(ctx, &host, st, sizeof(host));
        if (o->time_kernels) {
            // V-cycles after the stop exit at once; time those of real iterations
            const int64_t real = std::min<int64_t>(chunk, std::max<int64_t>(0, host.iters - counted));
            const int64_t real_timed = (real + kTimeEvery - 1) / kTimeEvery;
            for (int64_t k = 0; k < real_timed * per_iter && k < (int64_t)run.ev_next; ++k) {
                float ms = 0;
                SFEM_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                smooth_ms += ms;
                ++timed;
            }
        }
        counted = host.iters;
        next = chunk_max;
        if (host.iters > seen_iters && host.rr > 0.0 && host.rr < seen_rr && host.tol2bb > 0.0 &&
            std::isfinite(host.rr)) {
            const double per_iter_log = std::log(host.rr / seen_rr) / (double)(host.iters - seen_iters);
            const double left = std::log(host.tol2bb / host.rr) / per_iter_log;
            if (std::isfinite(left)) next = std::max<int64_t>(1, std::min<int64_t>(chunk_max, (int64_t)std::ceil(left)));
        }
        seen_iters = host.iters;
        seen_rr = host.rr;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (!host.done) {  // iteration cap: the true residual of the final iterate
        MgState h2 = host;
        h2.need_true = 1;
        h2d(ctx, st, &h2, sizeof(h2));
        true_resid();
        d2h(ctx, &host, st, sizeof(host));
        if (host.done != 1) host.done = 0;  // reported as the iteration cap
    }
    if (padded) pad_scatter(ctx, S, x, S->x.get());
    SFEM_HIP(hipEventRecord(t1, s));
    sync(ctx);
    float total = 0;
    SFEM_HIP(hipEventElapsedTime(&total, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);

    stats->iterations = host.iters;
    stats->fp64_updates = host.trues;
    stats->spmv_launches = host.iters;
    stats->spmv64_launches = host.trues;
    stats->rel_residual = host.bb > 0 ? std::sqrt(host.rr_true / host.bb) : 0.0;
    stats->solve_ms = total;
    stats->spmv_ms = 0.0;
    stats->spmv_timed = 0;
    const double vb = std::is_same<TV, SplitVal>::value ? 6.0 : (double)sizeof(TV);
    stats->spmv_bytes = (double)S->nnzb * (BS * BS * vb + 4.0) + (double)(S->nslices + 1) * 8.0 +
                        (double)S->nslices * 4.0 + 2.0 * (double)n * 8.0;
    stats->converged = host.done == 1 ? 1 : 0;
    stats->format = SFEM_FMT_BSR;
    stats->stop_reason = host.done == 0 ? 4 : host.done;
    stats->floor_residual = host.bb > 0 ? std::sqrt(host.floor_rr / 16.0 / host.bb) : 0.0;
    // a smoothing step: FP16 or FP32 operator once (matrix-free: the node map,
    // ghost-layer row and one cell scale per row, the fixed flags), v
    // gathered, r and x read and written, d read and written, the 3x3 (2x2)
    // inverse blocks read
    stats->smooth_ms = smooth_ms;
    stats->smooth_timed = timed;
    if (H.mf)
        stats->smooth_bytes = (double)S->nfb * (3 * 4.0 + BS) + (double)n * 4.0 * (6.0 + BS);
    else
        stats->smooth_bytes = (double)S->nnzb * (BS * BS * (H.fp16 ? 2.0 : 4.0) + 4.0) +
                              (double)(S->nslices + 1) * 8.0 + (double)S->nslices * 4.0 + (double)n * 4.0 * (7.0 + BS);
}

template <int ND>
void mg_solve_nd(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st) {
    MgHierarchy& H = *S->mg;
    switch (o->precision) {
        case SFEM_PREC_FP64: mg_pcg<ND, double>(ctx, S, H, o, st); break;
        case SFEM_PREC_SPLIT: mg_pcg<ND, SplitVal>(ctx, S, H, o, st); break;
        case SFEM_PREC_FP32_MATRIX: mg_pcg<ND, float>(ctx, S, H, o, st); break;
        case SFEM_PREC_MIXED:
            fail(SFEM_ERR_UNSUPPORTED, "multigrid: precision fp64, split or fp32_matrix (the V-cycle is FP32)");
        default: fail(SFEM_ERR_ARGUMENT, "solve: unknown precision");
    }
}

}  // namespace

void mg_opts_default(sfem_mg_opts* o) {
    o->degree = 2;
    o->coarse_degree = 0;
    o->lo = 0.1;
    o->hi = 1.1;
    o->power_iters = 8;
    o->coarse_nodes = 64;
    o->fine_fp16 = 1;
    o->coarse_fp16 = 1;
    o->matrix_free = 2;
}

void system_set_grid(sfem_ctx* ctx, sfem_system* S, int ndim, const int64_t* dims, const sfem_mg_opts* opts) {
    if (ndim == 0 && !dims) {  // the grid K was assembled on
        if (!S->K || !S->K->grid_nd)
            fail(SFEM_ERR_ARGUMENT, "set_grid: no element counts given and K was not assembled on a grid");
        ndim = S->K->grid_nd;
        dims = S->K->grid_ne;
    }
    if (ndim != 2 && ndim != 3) fail(SFEM_ERR_DIMENSION, "set_grid: ndim must be 2 or 3");
    if (!dims) fail(SFEM_ERR_ARGUMENT, "set_grid: dims is NULL");
    sfem_mg_opts o;
    mg_opts_default(&o);
    if (opts) o = *opts;
    if (o.degree < 1 || o.degree > kMgMaxDegree || o.coarse_degree < 0 || o.coarse_degree > kMgMaxDegree ||
        !(o.lo > 0.0) || !(o.hi > o.lo) || o.power_iters < 1 ||
        o.coarse_nodes < 1)
        fail(SFEM_ERR_ARGUMENT, "set_grid: bad multigrid options");
    if (!S->has_block || S->bs != ndim || !S->K || S->K->bs != ndim)
        fail(SFEM_ERR_UNSUPPORTED, "multigrid: needs a node-block system with n_dof = ndim (vector elasticity)");
    int64_t nn = 1;
    for (int d = 0; d < ndim; ++d) {
        if (dims[d] < 1) fail(SFEM_ERR_ARGUMENT, "set_grid: element counts must be >= 1");
        nn *= dims[d] + 1;
    }
    if (nn != S->K->nunits)
        fail(SFEM_ERR_ARGUMENT, "set_grid: the grid has " + std::to_string(nn) + " nodes, K has " +
                                    std::to_string(S->K->nunits));
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<MgHierarchy, MgDelete> H(new MgHierarchy());
    H->nd = ndim;
    H->bs = ndim;
    H->degree = o.degree;
    H->L.reserve(SFEM_MG_MAX_LEVELS);  // level references stay valid while levels are added
    MgLevel L0;
    for (int d = 0; d < 3; ++d) L0.n[d] = d < ndim ? dims[d] + 1 : 1;
    L0.nn = nn;
    H->L.push_back(std::move(L0));
    S->mg.reset();
    if (ndim == 3) build_levels<3>(ctx, S, *H, o);
    else build_levels<2>(ctx, S, *H, o);
    sync(ctx);
    H->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    S->grid_nd = ndim;
    for (int d = 0; d < 3; ++d) S->grid_dims[d] = d < ndim ? dims[d] : 0;
    S->mg_opts = o;
    S->mg = std::move(H);
}

void mg_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* o, sfem_solve_stats* st) {
    if (!S->mg) fail(SFEM_ERR_STATE, "multigrid: call sfem_system_set_grid first");
    if (S->mg->nd == 3) mg_solve_nd<3>(ctx, S, o, st);
    else mg_solve_nd<2>(ctx, S, o, st);
}

void mg_info(const sfem_system* S, sfem_mg_info* info) {
    *info = sfem_mg_info{};
    if (!S->mg) fail(SFEM_ERR_STATE, "multigrid: no hierarchy (sfem_system_set_grid)");
    const MgHierarchy& H = *S->mg;
    info->levels = (int)H.L.size();
    for (size_t l = 0; l < H.L.size(); ++l) {
        info->nodes[l] = H.L[l].nn;
        for (int d = 0; d < 3; ++d) info->dims[l][d] = H.L[l].n[d] - 1;
        info->lambda_max[l] = H.L[l].lmax;
    }
    info->coarse_dofs = H.nc;
    info->setup_ms = H.setup_ms;
    info->degree = H.degree;
    info->fine_fp16 = H.fp16 ? 1 : 0;
    info->matrix_free = H.mf;
    info->ke_deviation = H.ke_dev;
    info->coarse_fp16 = 0;
    for (size_t l = 1; l + 1 < H.L.size(); ++l)
        if (H.L[l].st16.size()) info->coarse_fp16 = 1;
}

void mg_stencil(sfem_ctx* ctx, const sfem_system* S, int level, double* stencil, uint8_t* fixed) {
    if (!S->mg) fail(SFEM_ERR_STATE, "multigrid: no hierarchy (sfem_system_set_grid)");
    const MgHierarchy& H = *S->mg;
    if (level < 1 || level >= (int)H.L.size()) fail(SFEM_ERR_ARGUMENT, "mg_stencil: level out of range");
    const MgLevel& L = H.L[(size_t)level];
    if (stencil) d2h(ctx, stencil, L.st64.get(), L.st64.bytes());
    if (fixed) d2h(ctx, fixed, L.fixed.get(), L.fixed.bytes());
}

}  // namespace sfem
