// element.hip -- per-element work of assemble_system on gfx950:
//   * validation of the connectivity / dof numbering (the errors the reference
//     raises while it walks the elements),
//   * element stiffness by 2x2(x2) Gauss quadrature, written in COO order,
//   * (key, payload) emission for the deterministic sort-by-key assembly.
//
// Reference: smearFEM src/fem.jl:135-252.  Element loop :179, Gauss loop :183,
// Jacobian / det / inverse / dNdX :192-196, Laplace branch :198-208, plane
// stress and 3D isotropic B'DB :209-233, COO index and (row, col) :236-249.
#include "basis.hpp"
#include "sfem_objects.hpp"

// Bitwise symmetry of Ke relies on every product being rounded on its own, so
// no multiply-add contraction in this file.
#pragma clang fp contract(off)

namespace sfem {

namespace {

// ------------------------------------------------------------ validation --
// One thread per element.  Error keys are element*8 + step, where step orders
// the checks the way the reference meets them inside one element:
//   0 NodeList[:,IEN[e,:]] out of range (:180)      -> BoundsError
//   2 singular Jacobian at the first Gauss point    -> SingularException
//   4 ID[IEN[e,a], d] out of range (:243)           -> BoundsError
//   5 singular Jacobian at a later Gauss point      -> SingularException
// (steps 1 and 3 are the structural errors found on the host for element 1).
__global__ __launch_bounds__(256) void validate_kernel(const int32_t* __restrict__ ien, int nen,
                                                       int64_t n_elem, int ndof,
                                                       const int32_t* __restrict__ dof,
                                                       int64_t id_rows, int64_t period, ValidateOut* out) {
    __shared__ unsigned long long s_key[4];
    __shared__ int s_max[4], s_min[4], s_blk[4];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = ~0ull;
    int mx = -1, mn = INT32_MAX, blk = 1;
    if (e < n_elem) {
        bool bad = false;
        for (int a = 0; a < nen; ++a)
            if (ien[e * nen + a] < 0) bad = true;
        const unsigned long long ek = (unsigned long long)(period > 0 ? e % period : e);
        if (bad) {
            key = ek * 8 + 0;
        } else if (ndof == 1) {
            for (int a = 0; a < nen; ++a) mx = max(mx, ien[e * nen + a]);
        } else {
            for (int a = 0; a < nen; ++a) {
                const int node = ien[e * nen + a];
                if (node >= id_rows) { key = ek * 8 + 4; break; }
                const int d0 = dof[(int64_t)node * ndof];
                if (d0 < 0 || d0 % ndof != 0) blk = 0;
                for (int d = 0; d < ndof; ++d) {
                    const int v = dof[(int64_t)node * ndof + d];
                    mn = min(mn, v);
                    mx = max(mx, v);
                    if (v != d0 + d) blk = 0;
                }
            }
        }
    }
    // wave then block reduction, one atomic per block into its group's copy
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long k2 = __shfl_xor(key, off, 64);
        key = k2 < key ? k2 : key;
        mx = max(mx, __shfl_xor(mx, off, 64));
        mn = min(mn, __shfl_xor(mn, off, 64));
        blk = min(blk, __shfl_xor(blk, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_key[w] = key; s_max[w] = mx; s_min[w] = mn; s_blk[w] = blk; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            key = s_key[i] < key ? s_key[i] : key;
            mx = max(mx, s_max[i]);
            mn = min(mn, s_min[i]);
            blk = min(blk, s_blk[i]);
        }
        ValidateOut* o = out + blockIdx.x % kValidateCopies;
        if (key != ~0ull) atomicMin(&o->err_key, key);
        if (mx >= 0) atomicMax(&o->max_unit, mx);
        if (mn != INT32_MAX) atomicMin(&o->min_dof, mn);
        if (!blk) atomicAnd(&o->blocked, 0);
    }
}

// ------------------------------------------------ element arithmetic --
// The steps of one element's quadrature.

// Jacobian at Gauss point g from the element's node coordinates (Jac =
// coords * dN, :192), its inverse (zero when det = 0, the caller reports the
// SingularException, :195) and the weight |det| (weights are 1.0, :194).
template <int NDIM, int NEN>
__device__ __forceinline__ double elem_jacobian(const double (*x)[NDIM], int g, double* inv, double& det) {
    double xi, eta, zeta;
    gauss_point(g, xi, eta, zeta);
    double J[NDIM][NDIM];
#pragma unroll
    for (int i = 0; i < NDIM; ++i)
#pragma unroll
        for (int j = 0; j < NDIM; ++j) J[i][j] = 0.0;
    for (int a = 0; a < NEN; ++a) {
        double dn[NDIM];
        dshape<NDIM, NEN>(a, xi, eta, zeta, dn);
#pragma unroll
        for (int i = 0; i < NDIM; ++i)
#pragma unroll
            for (int j = 0; j < NDIM; ++j) J[i][j] = J[i][j] + x[a][i] * dn[j];
    }
    if constexpr (NDIM == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double r = 1.0 / det;
        inv[0] = J[1][1] * r;  inv[1] = -J[0][1] * r;
        inv[2] = -J[1][0] * r; inv[3] = J[0][0] * r;
    } else {
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double r = 1.0 / det;
        inv[0] = c00 * r;
        inv[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
        inv[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
        inv[3] = c01 * r;
        inv[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
        inv[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
        inv[6] = c02 * r;
        inv[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
        inv[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    }
    if (det == 0.0)
#pragma unroll
        for (int i = 0; i < NDIM * NDIM; ++i) inv[i] = 0.0;
    return 1.0 * fabs(det);
}

// dN_a/dX at Gauss point g (dNdX = dN * invJ, :196)
template <int NDIM, int NEN>
__device__ __forceinline__ void elem_dndx(int a, int g, const double* inv, double* out) {
    double xi, eta, zeta;
    gauss_point(g, xi, eta, zeta);
    double dn[NDIM];
    dshape<NDIM, NEN>(a, xi, eta, zeta, dn);
#pragma unroll
    for (int k = 0; k < NDIM; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NDIM; ++j) s = s + dn[j] * inv[j * NDIM + k];
        out[k] = s;
    }
}

// the isotropic elasticity constants: plane stress (:217), 3D (:230)
template <int NDIM>
__device__ __forceinline__ void elem_d(double E, double nu, double& D11, double& D12, double& D44) {
    if constexpr (NDIM == 2) {
        D11 = E / (1 - nu * nu);
        D12 = nu * E / (1 - nu * nu);
        D44 = E / (2 * (1 + nu));
    } else {
        const double c = E / ((1 + nu) * (1 - 2 * nu));
        D11 = (1 - nu) * c;
        D12 = nu * c;
        D44 = (1 - 2 * nu) / 2 * c;
    }
}

// one Gauss point's w B_a' D B_b added to the NDIM x NDIM block (a, b), A =
// dN_a/dX, B = dN_b/dX, written out per entry (block (a,b) entry (i,j) and
// block (b,a) entry (j,i) are the same rounded expression)
template <int NDIM>
__device__ __forceinline__ void elem_block_gp(const double* A, const double* B, double w, double D11, double D12,
                                              double D44, double (&acc)[NDIM][NDIM]) {
#pragma unroll
    for (int i = 0; i < NDIM; ++i)
#pragma unroll
        for (int j = 0; j < NDIM; ++j) {
            double ent;
            if (i == j) {
                double o = 0.0;
                bool first = true;
#pragma unroll
                for (int k = 0; k < NDIM; ++k)
                    if (k != i) { o = first ? A[k] * B[k] : o + A[k] * B[k]; first = false; }
                ent = D11 * (A[i] * B[i]) + D44 * o;
            } else {
                ent = D12 * (A[i] * B[j]) + D44 * (A[j] * B[i]);
            }
            acc[i][j] = acc[i][j] + ent * w;
        }
}

// ------------------------------------------------------- element kernel --
// Ke is bitwise symmetric: the block formula below is B'DB written out per
// node pair, and with the reference's D matrices (:217, :230) and contraction
// off, block (a,b) entry (i,j) and block (b,a) entry (j,i) are the same
// rounded expression.  So only the node pairs a <= b are computed.
// A block holds EPB elements and one thread per (element, a <= b).
//   phase 0: gather the element coordinates into LDS
//   phase 1: per (element, gp) Jacobian, |det|, inverse; per (element, gp,
//            node) dN/dX
//   phase 2: per (element, a <= b) the ndof x ndof block of Ke summed over
//            the Gauss points in order (the reference's V[inz] += Ke[i,j],
//            :245) into LDS; with chk.ke0 set (structured grids) each block
//            is also compared with s_e Ke(0) (ElemCheck)
//   phase 3: coalesced store of the EPB elements' Ke in COO order, block
//            (b, a) as the transpose of (a, b) for a < b (V NULL: no store,
//            the check alone).
template <int NEN>
constexpr int elem_pairs() { return NEN * (NEN + 1) / 2; }
template <int NEN>
constexpr int elem_per_block() { return 256 / elem_pairs<NEN>(); }

template <int NDIM, int NEN, int NDOF>
__global__ __launch_bounds__(256) void element_kernel(const double* __restrict__ coords,
                                                      const int32_t* __restrict__ ien,
                                                      const double* __restrict__ young,
                                                      int64_t n_young, double nu_scalar,
                                                      const double* __restrict__ nu_e, int64_t n_elem,
                                                      double* __restrict__ V, ValidateOut* out, ElemCheck chk) {
    constexpr int NGP = 1 << NDIM;
    constexpr int NPS = elem_pairs<NEN>();
    constexpr int EPB = elem_per_block<NEN>();
    constexpr int B2 = NDOF * NDOF;
    constexpr int KE = NEN * NEN * B2;
    // the Gauss-point data (inverse Jacobians, weights, dN/dX) and, once the
    // blocks are formed, the blocks themselves share one LDS region
    constexpr int NINV = EPB * NGP * NDIM * NDIM, NW = EPB * NGP, ND_ = EPB * NGP * NEN * NDIM;
    constexpr int NKE = EPB * NPS * B2;
    constexpr int NPOOL = NINV + NW + ND_ > NKE ? NINV + NW + ND_ : NKE;
    __shared__ double sx[EPB][NEN][NDIM];
    __shared__ int svalid[EPB];
    __shared__ double pool[NPOOL];
    __shared__ double swmax[4];
    __shared__ int16_t stab[KE];  // COO entry of an element -> its value in ske
    auto sinv = reinterpret_cast<double (*)[NGP][NDIM * NDIM]>(pool);
    auto sw = reinterpret_cast<double (*)[NGP]>(pool + NINV);
    auto sd = reinterpret_cast<double (*)[NGP][NEN][NDIM]>(pool + NINV + NW);
    double* const ske = pool;  // the computed blocks, (element, a <= b), after the Gauss-point data

    const int tid = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * EPB;
    if (tid < EPB) svalid[tid] = (e0 + tid) < n_elem;
    for (int q = tid; q < KE; q += blockDim.x) {
        const int p = q / B2, m = q % B2, a = p / NEN, b = p % NEN, i = m / NDOF, j = m % NDOF;
        const int lo = min(a, b), hi = max(a, b);
        const int pair = lo * NEN - lo * (lo - 1) / 2 + (hi - lo);
        stab[q] = (int16_t)(pair * B2 + (a <= b ? i * NDOF + j : j * NDOF + i));
    }
    __syncthreads();
    for (int t = tid; t < EPB * NEN; t += blockDim.x) {
        const int el = t / NEN, a = t % NEN;
        const int64_t e = e0 + el;
        const int node = e < n_elem ? ien[e * NEN + a] : -1;
        if (node < 0) svalid[el] = 0;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) sx[el][a][d] = node >= 0 ? coords[(int64_t)node * NDIM + d] : 0.0;
    }
    __syncthreads();

    for (int t = tid; t < EPB * NGP; t += blockDim.x) {
        const int el = t / NGP, g = t % NGP;
        double det, inv[NDIM * NDIM];
        const double w = elem_jacobian<NDIM, NEN>(sx[el], g, inv, det);
        if (det == 0.0 && svalid[el])  // inv(Jac) throws SingularException, :195
            atomicMin(&out->err_key, (unsigned long long)(e0 + el) * 8 + (g == 0 ? 2 : 5));
        sw[el][g] = w;
#pragma unroll
        for (int i = 0; i < NDIM * NDIM; ++i) sinv[el][g][i] = inv[i];
    }
    __syncthreads();

    for (int t = tid; t < EPB * NGP * NEN; t += blockDim.x) {  // dNdX = ΔN * invJ, :196
        const int el = t / (NGP * NEN), g = (t / NEN) % NGP, a = t % NEN;
        elem_dndx<NDIM, NEN>(a, g, sinv[el][g], sd[el][g][a]);
    }
    __syncthreads();

    const bool busy = tid < EPB * NPS;
    const int el = busy ? tid / NPS : 0;
    int a = 0, b = busy ? tid % NPS : 0;  // pair index -> (a, b), a <= b
    while (b >= NEN - a) {
        b -= NEN - a;
        ++a;
    }
    b += a;
    const int64_t e = e0 + el;
    double acc[NDOF][NDOF];
#pragma unroll
    for (int i = 0; i < NDOF; ++i)
#pragma unroll
        for (int j = 0; j < NDOF; ++j) acc[i][j] = 0.0;
    if (busy) {
        double D11 = 0, D12 = 0, D44 = 0;
        if constexpr (NDOF > 1) {
            const double E = (e < n_elem) ? young[n_young == 1 ? 0 : e] : 0.0;
            const double nu = (nu_e && e < n_elem) ? nu_e[e] : nu_scalar;
            elem_d<NDIM>(E, nu, D11, D12, D44);
        }
        for (int g = 0; g < NGP; ++g) {
            const double w = sw[el][g];
            double A[NDIM], B[NDIM];
#pragma unroll
            for (int k = 0; k < NDIM; ++k) { A[k] = sd[el][g][a][k]; B[k] = sd[el][g][b][k]; }
            if constexpr (NDOF == 1) {  // w * dot(∇Na, ∇Nb), :206
                double s = A[0] * B[0];
#pragma unroll
                for (int k = 1; k < NDIM; ++k) s = s + A[k] * B[k];
                acc[0][0] = acc[0][0] + s * w;
            } else {
                elem_block_gp<NDIM>(A, B, w, D11, D12, D44, acc);
            }
        }
    }
    __syncthreads();  // every thread is past its Gauss-point reads
    if (busy)
#pragma unroll
        for (int i = 0; i < NDOF; ++i)
#pragma unroll
            for (int j = 0; j < NDOF; ++j) ske[tid * B2 + i * NDOF + j] = acc[i][j];
    __syncthreads();
    if (chk.ke0) {
        // s_e = Ke(e)[0][0] / Ke(0)[0][0]; |Ke(e) (1 / s_e) - Ke(0)| rounded
        // once (a fused multiply-add); block (b, a) holds the same values
        double dmax = 0.0;
        if (busy && e < n_elem) {
            const double s = ske[el * NPS * B2] / chk.ke0[0];
            const double ise = 1.0 / s;
            if (a == 0 && b == 0) chk.escale[e] = s;
#pragma unroll
            for (int i = 0; i < NDOF; ++i)
#pragma unroll
                for (int j = 0; j < NDOF; ++j)
                    dmax = max_nan(dmax, fabs(fma(acc[i][j], ise, -chk.ke0[(a * NEN + b) * B2 + i * NDOF + j])));
        }
        for (int off = 32; off > 0; off >>= 1) dmax = max_nan(dmax, __shfl_xor(dmax, off, 64));
        if ((tid & 63) == 0) swmax[tid >> 6] = dmax;
        __syncthreads();
        if (tid == 0) {
            double bm = swmax[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) bm = max_nan(bm, swmax[w]);
            chk.bmax[blockIdx.x] = bm;
        }
    }
    if (!V) return;  // the check alone: K is summed from the scales
    const int64_t nel = min((int64_t)EPB, n_elem - e0);
    const int count = (int)(nel * KE);
    double* dst = V + e0 * KE;
    for (int t = tid; t < count; t += blockDim.x) {
        const int te = t / KE;
        dst[t] = ske[te * (NPS * B2) + stab[t - te * KE]];
    }
}

// ------------------------------------------------------------- key emission --
// Node-block mode (bs == ndof, or ndof == 1): one key per (element, a, b);
// payload = COO index of the block's first value, e*KE + (a*nen+b)*ndof^2.
// Scalar mode (bs == 1 < ndof): one key per COO entry, payload = COO index.
// key = col_unit << row_bits | row_unit, so ascending keys are CSC order.
__global__ __launch_bounds__(256) void keys_block_kernel(const int32_t* __restrict__ ien, int nen,
                                                         int ndof, const int32_t* __restrict__ dof,
                                                         int64_t n_elem, int row_bits,
                                                         uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ payload) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int P = nen * nen;
    if (t >= n_elem * P) return;
    const int64_t e = t / P;
    const int p = (int)(t - e * P), a = p / nen, b = p % nen;
    const int na = ien[e * nen + a], nb = ien[e * nen + b];
    uint32_t ua, ub;
    if (ndof == 1) {
        ua = (uint32_t)na;
        ub = (uint32_t)nb;
    } else {
        ua = (uint32_t)(dof[(int64_t)na * ndof] / ndof);
        ub = (uint32_t)(dof[(int64_t)nb * ndof] / ndof);
    }
    keys[t] = ((uint64_t)ub << row_bits) | ua;
    payload[t] = (uint32_t)(e * (int64_t)P * ndof * ndof + (int64_t)p * ndof * ndof);
}

__global__ __launch_bounds__(256) void keys_scalar_kernel(const int32_t* __restrict__ ien, int nen,
                                                          int ndof, const int32_t* __restrict__ dof,
                                                          int64_t n_elem, int row_bits,
                                                          uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ payload) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nd2 = ndof * ndof, KE = nen * nen * nd2;
    if (t >= n_elem * KE) return;
    const int64_t e = t / KE;
    const int q = (int)(t - e * KE);
    const int p = q / nd2, i = (q % nd2) / ndof, j = q % ndof;
    const int na = ien[e * nen + p / nen], nb = ien[e * nen + p % nen];
    const uint32_t row = (uint32_t)dof[(int64_t)na * ndof + i];
    const uint32_t col = (uint32_t)dof[(int64_t)nb * ndof + j];
    keys[t] = ((uint64_t)col << row_bits) | row;
    payload[t] = (uint32_t)t;
}

template <int NDIM, int NEN, int NDOF>
void launch_elem(sfem_ctx* ctx, const sfem_mesh* m, double* V, ValidateOut* out, int64_t n_elem = -1,
                 ElemCheck chk = {}) {
    constexpr int EPB = elem_per_block<NEN>();
    if (n_elem < 0) n_elem = m->n_elem;
    const unsigned grid = grid_for(n_elem, EPB);
    hipLaunchKernelGGL((element_kernel<NDIM, NEN, NDOF>), dim3(grid), dim3(256), 0, ctx->stream,
                       m->coords.get(), m->ien.get(), m->young.get(), m->n_young, m->nu, m->nu_e.get(), n_elem,
                       V, out, chk);
    SFEM_LAUNCH_CHECK();
}

int grid_kind(const sfem_mesh* m) {
    const int kind = m->ndim * 100 + m->nen * 10 + m->ndof;
    if (m->face || (kind != 383 && kind != 242)) fail(SFEM_ERR_UNSUPPORTED, "element check: Hex8 / Q4 vector meshes");
    return kind;
}

}  // namespace

void launch_validate(sfem_ctx* ctx, const sfem_mesh* m, ValidateOut* d_out) {
    if (m->n_elem == 0) return;
    hipLaunchKernelGGL(validate_kernel, dim3(grid_for(m->n_elem, 256)), dim3(256), 0, ctx->stream,
                       m->ien.get(), m->nen, m->n_elem, m->ndof, m->dof.get(), m->id_rows, m->key_period, d_out);
    SFEM_LAUNCH_CHECK();
}

void launch_element_stiffness(sfem_ctx* ctx, const sfem_mesh* m, double* V, ValidateOut* d_out) {
    if (m->n_elem == 0) return;
    if (m->face) {
        launch_face_mass(ctx, m, V);
        return;
    }
    const int kind = m->ndim * 100 + m->nen * 10 + m->ndof;
    switch (kind) {
        case 242: launch_elem<2, 4, 2>(ctx, m, V, d_out); break;  // Q4 plane stress
        case 383: launch_elem<3, 8, 3>(ctx, m, V, d_out); break;  // Hex8 elasticity
        case 241: launch_elem<2, 4, 1>(ctx, m, V, d_out); break;  // Q4 Laplace
        case 291: launch_elem<2, 9, 1>(ctx, m, V, d_out); break;  // Q2 Laplace
        // Q2 with nDof=2 always overruns the COO (a BoundsError); the Laplace
        // kernel runs only to find singular elements that come first.
        case 292: launch_elem<2, 9, 1>(ctx, m, V, d_out); break;
        case 381: launch_elem<3, 8, 1>(ctx, m, V, d_out); break;  // Hex8 Laplace
        default: fail(SFEM_ERR_DIMENSION, "no element kernel for this (ndim, nen, nDof)");
    }
}

unsigned element_launch_blocks(const sfem_mesh* m) {
    return grid_for(m->n_elem, grid_kind(m) == 383 ? elem_per_block<8>() : elem_per_block<4>());
}

void launch_element_zero(sfem_ctx* ctx, const sfem_mesh* m, double* ke0, ValidateOut* d_out) {
    if (grid_kind(m) == 383)
        launch_elem<3, 8, 3>(ctx, m, ke0, d_out, 1);
    else
        launch_elem<2, 4, 2>(ctx, m, ke0, d_out, 1);
}

void launch_element_checked(sfem_ctx* ctx, const sfem_mesh* m, double* V, ValidateOut* d_out,
                            const ElemCheck& chk) {
    if (grid_kind(m) == 383)
        launch_elem<3, 8, 3>(ctx, m, V, d_out, -1, chk);
    else
        launch_elem<2, 4, 2>(ctx, m, V, d_out, -1, chk);
}

int64_t launch_keys(sfem_ctx* ctx, const sfem_mesh* m, int bs, int row_bits, uint64_t* keys,
                    uint32_t* payload) {
    const int nen = m->nen, ndof = m->ndof;
    if (bs == ndof || ndof == 1) {
        const int64_t n = m->n_elem * nen * nen;
        if (n)
            hipLaunchKernelGGL(keys_block_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                               m->ien.get(), nen, ndof, m->dof.get(), m->n_elem, row_bits, keys,
                               payload);
        SFEM_LAUNCH_CHECK();
        return n;
    }
    const int64_t n = m->n_elem * nen * nen * ndof * ndof;
    if (n)
        hipLaunchKernelGGL(keys_scalar_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                           m->ien.get(), nen, ndof, m->dof.get(), m->n_elem, row_bits, keys, payload);
    SFEM_LAUNCH_CHECK();
    return n;
}

}  // namespace sfem
